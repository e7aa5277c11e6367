// Host-side plan logic of the library, checked on the CPU under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_host_sanitize.py builds and runs this file):
//   * the quarter transform's cyclic-window schedule (qt_window, qt_row, qt_owns): every
//     unordered orbital pair {p, q}, diagonal included, is owned by exactly one batch;
//   * the FP64 quarter transform's window copy of C: every column a tile reads is inside
//     qt_window_cols(n) and holds the orbital the window row needs;
//   * tri_pair, the lower-triangle tile-pair decode of the SYRK problems, and syrk_tile,
//     the same with diagonal pair tiles: every lower-triangle tile covered exactly once;
//   * GroupBatch::xcd_bounds, the per-XCD cut of a grouped launch's tile table;
//   * GroupCollector::add, the valid-tile enumeration of a grouped launch;
//   * ws_carve, the workspace layout of a build (also the no-allocation build's scratch).
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "jk_common.h"
#include "jk_kernels.h"

using namespace pnjk;

static int g_fail = 0;
#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      if (g_fail < 20) {                                  \
        std::fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
        std::fprintf(stderr, __VA_ARGS__);                \
        std::fprintf(stderr, "\n");                       \
      }                                                   \
      ++g_fail;                                           \
    }                                                     \
  } while (0)

static void check_window() {
  for (int n = 1; n <= 700; ++n) {
    const int h = qt_window(n);
    std::vector<int> owned((size_t)n * n, 0);
    for (int p = 0; p < n; ++p) {
      CHECK(qt_row(p, 0, n) == p && qt_owns(p, 0, n), "n=%d p=%d: window row 0 is not q = p", n, p);
      for (int r = 0; r < h; ++r) {
        const int q = qt_row(p, r, n);
        CHECK(q >= 0 && q < n, "n=%d p=%d r=%d: q=%d out of range", n, p, r, q);
        if (!qt_owns(p, r, n)) continue;
        const int hi = p > q ? p : q, lo = p > q ? q : p;
        ++owned[(size_t)hi * n + lo];
      }
    }
    for (int a = 0; a < n; ++a)
      for (int b = 0; b <= a; ++b)
        CHECK(owned[(size_t)a * n + b] == 1, "n=%d pair (%d,%d) owned %d times", n, a, b, owned[(size_t)a * n + b]);
  }
}

static void check_window_copy() {
  for (int n = 1; n <= 700; ++n) {
    const int h = qt_window(n), ldw = qt_window_cols(n);
    CHECK(ldw % 2 == 0, "n=%d: odd window stride", n);
    for (int BM : {64, 128, 160}) {
      const int mtiles = (h + BM - 1) / BM;
      CHECK((long)mtiles * BM + n - 1 <= ldw, "n=%d BM=%d: launcher bound fails", n, BM);
      for (int p = 0; p < n; ++p) {
        const int base = (p & 1) ? p - 1 : p, shift = p & 1;  // Cw1[j] = C[(j + 1) mod n]
        CHECK(base % 2 == 0, "n=%d p=%d: odd base", n, p);
        const int last = base + mtiles * BM - 1;
        CHECK(last < ldw, "n=%d BM=%d p=%d: column %d past %d", n, BM, p, last, ldw);
        for (int r = 0; r < h; ++r)
          CHECK((base + r + shift) % n == qt_row(p, r, n), "n=%d p=%d r=%d: wrong orbital", n, p, r);
      }
    }
  }
}

static void check_tri_pair() {
  for (int rem = 0; rem < (1 << 22); ++rem) {
    int mt = -1, nt = -1;
    tri_pair(rem, mt, nt);
    CHECK(nt >= 0 && nt <= mt && mt * (mt + 1) / 2 + nt == rem, "rem=%d -> (%d,%d)", rem, mt, nt);
  }
}

static void check_syrk_tile() {
  for (int T = 1; T <= 64; ++T)
    for (int npair = 0; npair <= T / 2; ++npair) {
      const int count = T * (T + 1) / 2 - 2 * npair;
      std::vector<int> cover((size_t)T * T, 0);
      int last = -1;  // row-major position of the previous regular tile
      for (int rem = 0; rem < count; ++rem) {
        int mt = -9, nt = -9;
        syrk_tile(rem, npair, mt, nt);
        if (rem < npair) {
          CHECK(nt == -1 && mt == 2 * rem && mt + 1 < T, "T=%d npair=%d rem=%d -> (%d,%d)", T, npair, rem, mt, nt);
          if (nt != -1 || mt < 0 || mt + 1 >= T) continue;
          ++cover[(size_t)mt * T + mt];
          ++cover[(size_t)(mt + 1) * T + mt];
          ++cover[(size_t)(mt + 1) * T + mt + 1];
        } else {
          CHECK(nt >= 0 && nt <= mt && mt < T, "T=%d npair=%d rem=%d -> (%d,%d)", T, npair, rem, mt, nt);
          if (nt < 0 || nt > mt || mt >= T) continue;
          CHECK(mt * T + nt > last, "T=%d npair=%d rem=%d: not in row-major order", T, npair, rem);
          last = mt * T + nt;
          ++cover[(size_t)mt * T + nt];
        }
      }
      for (int a = 0; a < T; ++a)
        for (int b = 0; b <= a; ++b)
          CHECK(cover[(size_t)a * T + b] == 1, "T=%d npair=%d: tile (%d,%d) covered %d times", T, npair, a, b,
                cover[(size_t)a * T + b]);
    }
}

static void check_xcd_bounds(std::mt19937& rng) {
  for (int trial = 0; trial < 400; ++trial) {
    GroupBatch b;
    const int n = (int)(rng() % 5000);
    int maxc = 0;
    long total = 0;
    for (int i = 0; i < n; ++i) {
      const int c = 1 + (int)(rng() % 200);
      b.tiles.push_back(make_int2(0, i));
      b.ksteps.push_back(c);
      maxc = c > maxc ? c : maxc;
      total += c;
    }
    const std::vector<int> xb = b.xcd_bounds();
    CHECK((int)xb.size() == kXcds + 1 && xb[0] == 0 && xb[kXcds] == n, "trial %d: bad ends", trial);
    for (int x = 0; x < kXcds; ++x) {
      CHECK(xb[x] <= xb[x + 1], "trial %d: cut points not monotone", trial);
      long run = 0;
      for (int i = xb[x]; i < xb[x + 1]; ++i) run += b.ksteps[i];
      CHECK(run <= total / kXcds + maxc + 1, "trial %d xcd %d: run %ld of total %ld", trial, x, run, total);
    }
  }
}

struct FakeProb {  // every third tile of a batch is not needed
  int per;
  bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = id / per;
    t.mt = id % per;
    t.nt = 0;
    t.m0 = t.mt * BM;
    t.n0 = 0;
    return t.mt % 3 != 2;
  }
};

static hipError_t fake_launch(const void*, const int2*, const int*, int, hipStream_t) { return hipSuccess; }

static void check_collector() {
  GroupCollector col;
  FakeProb a{7}, b{5};
  col.add(a, TileF64{2, 2}, 70, 64, 64, 11, &fake_launch, "fake");
  col.add(b, TileF64{2, 2}, 50, 64, 64, 13, &fake_launch, "fake");
  col.add(a, TileF64{5, 4}, 14, 160, 128, 3, &fake_launch, "fake");
  CHECK(col.batches.size() == 2, "two tile shapes give two batches (%zu)", col.batches.size());
  const GroupBatch& g = col.batches[0];
  CHECK(g.probs.size() == 2 * sizeof(FakeProb), "two problem slots");
  CHECK(g.tiles.size() == g.ksteps.size(), "one cost per tile");
  size_t want = 0;
  for (int id = 0; id < 70; ++id) want += (id % 7) % 3 != 2;
  for (int id = 0; id < 50; ++id) want += (id % 5) % 3 != 2;
  CHECK(g.tiles.size() == want, "valid tiles %zu, expected %zu", g.tiles.size(), want);
  for (size_t i = 0; i < g.tiles.size(); ++i) {
    const int slot = g.tiles[i].x, id = g.tiles[i].y;
    const int per = slot == 0 ? 7 : 5;
    CHECK((id % per) % 3 != 2 && g.ksteps[i] == (slot == 0 ? 11 : 13), "tile %zu (%d,%d)", i, slot, id);
  }
}

static void check_carve(std::mt19937& rng) {
  std::vector<char> arena(1 << 24);
  for (int trial = 0; trial < 2000; ++trial) {
    size_t bytes[14];
    void* ptr[14];
    for (int i = 0; i < 14; ++i) bytes[i] = (rng() % 4 == 0) ? 0 : (size_t)(rng() % 100000);
    const size_t need = ws_carve(bytes, 14, nullptr, ptr);
    CHECK(need <= arena.size(), "trial %d: need %zu", trial, need);
    // the arena's start is 256-aligned in the same way the scratch array is
    char* base = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(arena.data()) + 255) / 256 * 256);
    if (need + 256 > arena.size()) continue;
    const size_t used = ws_carve(bytes, 14, base, ptr);
    CHECK(used == need, "trial %d: sizes differ", trial);
    char* prev_end = base;
    for (int i = 0; i < 14; ++i) {
      if (!bytes[i]) {
        CHECK(ptr[i] == nullptr, "trial %d piece %d: zero-size piece has an address", trial, i);
        continue;
      }
      char* p = static_cast<char*>(ptr[i]);
      CHECK((reinterpret_cast<uintptr_t>(p) & 255) == 0, "trial %d piece %d: not 256-aligned", trial, i);
      CHECK(p >= prev_end && p + bytes[i] <= base + used, "trial %d piece %d: overlap or past the end", trial, i);
      std::fill(p, p + bytes[i], (char)i);  // ASan catches a piece outside the arena
      prev_end = p + bytes[i];
    }
  }
}

int main() {
  std::mt19937 rng(20261204);
  check_window();
  check_window_copy();
  check_tri_pair();
  check_syrk_tile();
  check_xcd_bounds(rng);
  check_collector();
  check_carve(rng);
  if (g_fail) {
    std::fprintf(stderr, "%d host plan checks failed\n", g_fail);
    return 1;
  }
  std::printf("all host plan checks passed\n");
  return 0;
}
