// Split-precision (BF16x3) kernels of the RI JK build for gfx950.
//
// Each FP64 operand x is stored as two bf16 planes, x_hi = bf16(x) and
// x_lo = bf16(x - x_hi); a product a*b is formed as a_hi*b_hi + a_hi*b_lo +
// a_lo*b_hi on v_mfma_f32_16x16x32_bf16 with FP32 accumulation. The GEMM core
// stages both planes of both operands through double-buffered LDS in a
// row-major, k-contiguous layout (one 16-byte fragment per lane per MFMA), so
// every operand needs its contraction index contiguous. Layouts chosen for that:
//   Ct16[q][mu]        C^T in 16-row x 32-mu blocks, per build (A of HT, QT)
//   bht16[nu][P][mu]   b with mu contiguous, in 16-row x 32-mu blocks (blk_index),
//                      at upload (B of HT)
//   btri16[t][P]       packed lower triangle in 16-row x 32-P blocks, at upload (B of JN)
//   X16[q][nu][P]      half-transformed b, per build  (A/B of K^n; B of QT with
//                      P contiguous: copied to LDS as is and read with the
//                      LDS transpose read)
// Contractions (reference lines as in the FP64 kernels): HT integrals.py:211,
// QT :354 (-> d :212, K_MO :358), JN :214, K^n :216.
#include <type_traits>
#include <utility>

#include "jk_common.h"
#include "jk_kernels.h"

namespace pnjk {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

constexpr int kThreads = 256;
constexpr int kBK = 32;
constexpr int kLDK = kBK;  // LDS row: 32 bf16 = 64 bytes = four 16-byte chunks, unpadded

// LDS placement of 16-byte chunk `c` (0..3) of tile row `r`: XOR-swizzled so
// that 16 lanes reading the same logical chunk of 16 consecutive rows (one
// MFMA fragment read) and 4 lanes storing the 4 chunks of each of 4 rows (the
// loaders) touch distinct banks, for 32 and for 64 4-byte banks alike.
__device__ __forceinline__ int swz(int r, int c) { return r * kLDK + 8 * (c ^ (((r >> 1) ^ (r >> 2)) & 3)); }

// Output row of accumulator register r of `lane` inside a 16x16 f32 MFMA block.
__device__ __forceinline__ int out_row(int lane, int r) { return 4 * (lane >> 4) + r; }

// A bf16 operand: two planes with identical layout.
struct Planes {
  const uint16_t* hi;
  const uint16_t* lo;
};

// Operand layouts: RK = rows with the contraction index k contiguous (copied to
// LDS row by row); KN = k-major with the tile's columns contiguous (copied to LDS
// in that orientation and transposed by the fragment read).
enum LdKind { RK = 0, KN = 2 };

// Direct HBM -> LDS copy of both planes of a ROWS x 32 tile whose contraction
// index k is contiguous in memory (the 32 elements from k0 of row r at at(r, k0)), with
// global_load_lds_dwordx4: lane L of one wave instruction writes 16 bytes at the
// LDS base + 16 L, so one instruction fills 16 tile rows (1 KB). Lane L fetches
// the global chunk that belongs in swizzled slot L & 3 of row L >> 2. The
// instructions are dealt round-robin over the NW waves of the workgroup. No registers are
// staged and no LDS store is issued; the loads complete under vmcnt.
template <int ROWS, int NW>
struct Dma {
  static_assert(ROWS % 16 == 0, "a tile is whole 16-row blocks");
  static constexpr int kInst = ROWS / 16;
  template <class S>
  static __device__ __forceinline__ void issue(const S& s, int r0, int k0, uint16_t* sh, uint16_t* sl, int wave,
                                               int lane) {
    const int rr = lane >> 2, slot = lane & 3;
#pragma unroll
    for (int j = 0; j < (kInst + NW - 1) / NW; ++j) {
      const int i = NW * j + wave;  // wave-uniform
      if (kInst % NW == 0 || i < kInst) {
        const int r = 16 * i + rr;
        const int c = slot ^ (((r >> 1) ^ (r >> 2)) & 3);
        const int64_t o = s.at(r0 + r, k0) + 8 * c;
        __builtin_amdgcn_global_load_lds((const PNJK_GLOBAL void*)(s.p.hi + o),
                                         (__attribute__((address_space(3))) void*)(sh + 16 * i * kLDK), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const PNJK_GLOBAL void*)(s.p.lo + o),
                                         (__attribute__((address_space(3))) void*)(sl + 16 * i * kLDK), 16, 0, 0);
      }
    }
  }
};

// 16-column group ("quad", 32 bytes) swizzle of row k of a KN tile of COLS
// columns: quad Q of row k sits at Q ^ qswz(k). One transpose read (below) of a
// half wave touches rows {r, r+1, r+2, r+3, r+8, .., r+11} (r a multiple of 4) at one
// logical quad; the swizzle puts those eight 32-byte pieces on distinct banks.
template <int COLS>
__device__ __forceinline__ int qswz(int k) {
  static_assert(COLS == 64 || COLS == 128, "KN tiles are 64 or 128 columns wide");
  if constexpr (COLS == 128) return (k & 3) | ((k >> 1) & 4);
  else return ((k >> 1) & 1) | ((k >> 2) & 2);  // rows of 128 bytes: odd rows sit 32 banks on
}

// Direct HBM -> LDS copy of both planes of a 32 (k) x COLS tile; the source maps (k, n) to its
// element offset, eight consecutive n from a multiple of 8 being contiguous. LDS keeps the
// orientation: row k holds the COLS elements, quads swizzled by qswz. Lane L of
// instruction i writes the 16 bytes at LDS chunk 64 i + L, i.e. row k =
// (64 i + L) / (COLS / 8); it fetches the global chunk that belongs there.
template <int COLS, int NW>
struct DmaKN {
  static constexpr int kInst = COLS / 16;  // 32 rows x 2 COLS bytes per plane, 1 KB per instruction
  static constexpr int kRowChunks = COLS / 8;
  template <class S>
  static __device__ __forceinline__ void issue(const S& s, int n0, int k0, uint16_t* sh, uint16_t* sl, int wave,
                                               int lane) {
#pragma unroll
    for (int j = 0; j < (kInst + NW - 1) / NW; ++j) {
      const int i = NW * j + wave;  // wave-uniform
      if (kInst % NW == 0 || i < kInst) {
        const int o = 64 * i + lane;
        const int k = o / kRowChunks, pc = o % kRowChunks;
        const int64_t g = s.at(k0 + k, n0 + 16 * ((pc >> 1) ^ qswz<COLS>(k)) + 8 * (pc & 1));
        __builtin_amdgcn_global_load_lds((const PNJK_GLOBAL void*)(s.p.hi + g),
                                         (__attribute__((address_space(3))) void*)(sh + 512 * i), 16, 0, 0);
        __builtin_amdgcn_global_load_lds((const PNJK_GLOBAL void*)(s.p.lo + g),
                                         (__attribute__((address_space(3))) void*)(sl + 512 * i), 16, 0, 0);
      }
    }
  }
};

// The MFMA B fragment of columns [n, n + 16) from a KN tile in LDS: lane l needs
// elements k = 8 (l / 16) + t, t < 8, of column n + l % 16. ds_read_b64_tr_b16 over
// each group of 16 lanes hands lane i element e of the 8 bytes that lane
// 4 e + i / 4 of the group addressed; so lane 4 j + c of group g addresses row
// r = 8 g + j (then r + 4 for t >= 4), columns n + 4 c .. n + 4 c + 3. Rows r and
// r + 4 share their swizzle, so the second read is the first one's address plus a
// constant.
template <int COLS>
struct TrRead {
  int base, sw;  // this lane's element offset (row r, column 4 c) and its quad swizzle
  __device__ __forceinline__ explicit TrRead(int lane) {
    const int r = 8 * (lane >> 4) + ((lane >> 2) & 3);
    base = r * COLS + 4 * (lane & 3);
    sw = qswz<COLS>(r);
  }
  __device__ __forceinline__ bf16x8 read(const uint16_t* tile, int n) const {
    typedef __attribute__((address_space(3))) bf16x4 lds4;
    const uint16_t* p = tile + base + 16 * ((n >> 4) ^ sw);
    const bf16x4 x = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds4*)p);
    const bf16x4 y = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds4*)(p + 4 * COLS));
    return __builtin_shufflevector(x, y, 0, 1, 2, 3, 4, 5, 6, 7);
  }
};

// One output tile of (16 FM WGM) x (16 FN WGN): WGM x WGN waves, each with an
// FM x FN grid of 16x16 MFMA blocks. LDS: two buffers per operand (separate
// arrays, so the compiler sees that fragment reads of one never alias copies
// into the other), each holding the hi plane then the lo plane. k-step kt+1 is
// fetched into the other buffer while k-step kt computes; one barrier per k-step.
// (Three stages in eight-wave workgroups, and in four-wave ones at one workgroup
// per CU, measured no faster at nbf = 1000.)
// Problems whose tiles have k-step counts of their own (Prob::ksteps(tile); ksteps() is then
// the largest, for the grouped launches' cost).
template <class P, class = void>
struct HasTileKsteps : std::false_type {};
template <class P>
struct HasTileKsteps<P, std::void_t<decltype(std::declval<const P&>().ksteps(std::declval<const Tile&>()))>>
    : std::true_type {};

template <int FM, int FN, int WGM, int WGN, class Prob>
__device__ __forceinline__ void gemm_tile_x3(const Prob& prob, const Tile& t, uint16_t* sA0, uint16_t* sA1,
                                             uint16_t* sB0, uint16_t* sB1) {
  constexpr int BM = 16 * FM * WGM, BN = 16 * FN * WGN, NW = WGM * WGN;
  static_assert(Prob::kA == RK, "the A operand is k-contiguous");
  constexpr bool kBkn = Prob::kB == KN;
  const auto as = prob.a_src(t);
  const auto bs = prob.b_src(t);
  int nk;
  if constexpr (HasTileKsteps<Prob>::value)
    nk = prob.ksteps(t);
  else
    nk = prob.ksteps();

  // The wave index is uniform over the wave; readfirstlane makes that visible to
  // the compiler, so per-wave offsets stay scalar.
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int wm = (wave / WGN) * (16 * FM), wn = (wave % WGN) * (16 * FN);

  auto fetch = [&](int k0, uint16_t* a, uint16_t* b) {
    Dma<BM, NW>::issue(as, t.m0, k0, a, a + BM * kLDK, wave, lane);
    if constexpr (kBkn)
      DmaKN<BN, NW>::issue(bs, t.n0, k0, b, b + BN * kLDK, wave, lane);
    else
      Dma<BN, NW>::issue(bs, t.n0, k0, b, b + BN * kLDK, wave, lane);
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  fetch(0, sA0, sB0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  v4f acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

  const int frow = lane & 15, fk = 8 * (lane >> 4);
  const TrRead<kBkn ? BN : 64> tr(lane);
  // MFMA row blocks [i0, nact) of this wave hold a needed output row: with Prob::kRowSkip,
  // rows from the problem's row limit on (past the diagonal q <= p of the quarter
  // transform, padding past nbf5 or M) get no fragment reads and no MFMAs; with
  // Prob::kLower, neither do row blocks wholly above the diagonal of the output (their
  // last row before this wave's first column). Wave-uniform.
  const int nact = Prob::kRowSkip ? min(FM, max(0, (prob.row_limit(t) - (t.m0 + wm) + 15) >> 4)) : FM;
  const int i0 = Prob::kLower ? min(FM, max(0, (t.n0 + wn - (t.m0 + wm)) >> 4)) : 0;
  auto step = [&](const int kt, auto stage) {
    constexpr int cur = decltype(stage)::value;
    uint16_t* const a_c = cur ? sA1 : sA0;
    uint16_t* const b_c = cur ? sB1 : sB0;
    if (kt + 1 < nk) fetch((kt + 1) * kBK, cur ? sA0 : sA1, cur ? sB0 : sB1);  // wave-uniform
    bf16x8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      if (i >= i0 && i < nact) {
        const int o = swz(wm + 16 * i + frow, fk >> 3);
        ah[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(a_c + o));
        al[i] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(a_c + BM * kLDK + o));
      }
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      if constexpr (kBkn) {
        bh[j] = tr.read(b_c, wn + 16 * j);
        bl[j] = tr.read(b_c + BN * kLDK, wn + 16 * j);
      } else {
        const int o = swz(wn + 16 * j + frow, fk >> 3);
        bh[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(b_c + o));
        bl[j] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(b_c + BN * kLDK + o));
      }
    }
    // While this wave issues its MFMA block it runs at raised priority, so the SIMD's other
    // wave (usually in its copy/wait phase) does not take the issue slots: 3-5% on the HT
    // and K^n cores.
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int i = 0; i < FM; ++i)
      if (i >= i0 && i < nact) {
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
      }
    __builtin_amdgcn_s_setprio(0);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies into the next buffers have landed
    __syncthreads();
  };
  for (int kt = 0; kt < nk; kt += 2) {
    step(kt, S0{});
    if (kt + 1 < nk) step(kt + 1, S1{});
  }
  prob.template epilogue<FM, FN>(t, acc, wm, wn, lane, reinterpret_cast<double*>(sA0));
}

// Problems with diagonal pair tiles (Prob::kDiagPairs): a symmetric product A A^T whose
// lower triangle of BM x BM tiles would have half-empty diagonal tiles.
template <class P, class = void>
struct HasDiagPairs : std::false_type {};
template <class P>
struct HasDiagPairs<P, std::void_t<decltype(P::kDiagPairs)>> : std::bool_constant<P::kDiagPairs> {};
template <int FM, int FN, int WGM, int WGN>
constexpr bool kDiagPairShape = FM == FN && WGM == 2 && WGN == 2 && FM % 2 == 0;

// Diagonal pair tile of a symmetric product O = A A^T (t.nt < 0): the two diagonal tiles
// of panels P0 = rows [m0, m0 + BM) and P1 = [m0 + BM, m0 + 2 BM) and the tile between
// them, i.e. the lower triangle of one 2BM x 2BM block, from one copy of its 2BM rows per
// k-step (P0 into the A buffers, P1 into the B buffers, both as the regular tile copies
// them). In 16-row strips g = 0..S-1 (S = 2BM / 16) the triangle holds blocks (g, c),
// c <= g. Folded row f < S/2 pairs strip S-1-f (columns 0..S-1-f) with strip f (columns
// 0..f): S + 1 blocks for every f, accumulator j = c for the long strip and j = S - c for
// the short one, so a block's accumulator index is fixed by its column and each B
// fragment, read once per k-step, serves every strip of the wave that needs it. Wave w
// takes folded rows w, w + 4, ...: every wave issues the same MFMA count. Each output
// (m, n), n <= m, comes from the same fragments in the same order as in the regular
// tiles (A = strip of m, B = strip of n), so the results are those of the regular tiles
// bit for bit; the copy round trip that sets a k-step's time is paid once for three
// tiles' work (K^n at nbf = 1000 on 128 x 128 tiles: 3.50 -> 3.18 ms, 28 instead of 36
// workgroups per orbital).
template <int FM, int WGM, int WGN, class Prob>
__device__ __forceinline__ void diag_pair_x3(const Prob& prob, const Tile& t, uint16_t* sA0, uint16_t* sA1,
                                             uint16_t* sB0, uint16_t* sB1) {
  constexpr int BM = 16 * FM * WGM, NW = WGM * WGN;
  constexpr int S = 2 * BM / 16, H = S / 2, R = H / NW;  // strips, strips per panel, folded rows per wave
  static_assert(H % NW == 0, "folded rows are dealt evenly over the waves");
  const auto as = prob.a_src(t);
  const int nk = prob.ksteps();
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  auto fetch = [&](int k0, uint16_t* a, uint16_t* b) {
    Dma<BM, NW>::issue(as, t.m0, k0, a, a + BM * kLDK, wave, lane);
    Dma<BM, NW>::issue(as, t.m0 + BM, k0, b, b + BM * kLDK, wave, lane);
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  fetch(0, sA0, sB0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  v4f acc[R][S + 1];
#pragma unroll
  for (int k = 0; k < R; ++k)
#pragma unroll
    for (int j = 0; j <= S; ++j) acc[k][j] = v4f{0.f, 0.f, 0.f, 0.f};
  // Folded row k of this wave is f = wave + NW k: short strip f, long strip S - 1 - f. The
  // block tests compare the wave index with constants, so most fold at compile time.
  // Strips past the last row (in the last pair tile) compute on clamped rows like the others
  // and are not stored.
  __builtin_assume(wave >= 0 && wave < NW);
  const int frow = lane & 15, fk = 8 * (lane >> 4);
  auto step = [&](const int kt, auto stage) {
    constexpr int cur = decltype(stage)::value;
    uint16_t* const p0 = cur ? sA1 : sA0;
    uint16_t* const p1 = cur ? sB1 : sB0;
    if (kt + 1 < nk) fetch((kt + 1) * kBK, cur ? sA0 : sA1, cur ? sB0 : sB1);  // wave-uniform
    auto strip = [&](int g, bf16x8& h, bf16x8& l) {
      const uint16_t* b = g < H ? p0 : p1;
      const int o = swz(16 * (g < H ? g : g - H) + frow, fk >> 3);
      h = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(b + o));
      l = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(b + BM * kLDK + o));
    };
    bf16x8 lh[R], ll[R], sh[R], sl[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      strip(S - 1 - (wave + NW * k), lh[k], ll[k]);
      strip(wave + NW * k, sh[k], sl[k]);
    }
    // B fragments one column ahead, read whether or not the wave needs the column.
    bf16x8 bh, bl;
    strip(0, bh, bl);
    __builtin_amdgcn_s_setprio(1);
#pragma unroll
    for (int c = 0; c < S; ++c) {
      bf16x8 nh, nl;
      if (c + 1 < S) strip(c + 1, nh, nl);
      if (c + wave <= S - 1) {  // column c is in the wave's longest strip
#pragma unroll
        for (int k = 0; k < R; ++k) {
          if (c + wave + NW * k <= S - 1) {  // block (long strip, c)
            acc[k][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ll[k], bh, acc[k][c], 0, 0, 0);
            acc[k][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lh[k], bl, acc[k][c], 0, 0, 0);
            acc[k][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(lh[k], bh, acc[k][c], 0, 0, 0);
          }
          if (c <= wave + NW * k) {  // block (short strip, c)
            acc[k][S - c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sl[k], bh, acc[k][S - c], 0, 0, 0);
            acc[k][S - c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sh[k], bl, acc[k][S - c], 0, 0, 0);
            acc[k][S - c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sh[k], bh, acc[k][S - c], 0, 0, 0);
          }
        }
      }
      if (c + 1 < S) {
        bh = nh;
        bl = nl;
      }
    }
    __builtin_amdgcn_s_setprio(0);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies into the next buffers have landed
    __syncthreads();
  };
  for (int kt = 0; kt < nk; kt += 2) {
    step(kt, S0{});
    if (kt + 1 < nk) step(kt + 1, S1{});
  }
#pragma unroll
  for (int k = 0; k < R; ++k)
#pragma unroll
    for (int j = 0; j <= S; ++j) {
      const int f = wave + NW * k;
      const bool lng = j + f <= S - 1;
      prob.store_block(t, acc[k][j], lng ? S - 1 - f : f, lng ? j : S - j, lane);
    }
}

// A tile of the problem: a diagonal pair tile where the problem has them (t.nt < 0, only
// for the shapes that take them), else the regular tile.
template <int FM, int FN, int WGM, int WGN, class Prob>
__device__ __forceinline__ void run_tile_x3(const Prob& prob, const Tile& t, uint16_t* sA0, uint16_t* sA1,
                                            uint16_t* sB0, uint16_t* sB1) {
  if constexpr (HasDiagPairs<Prob>::value && kDiagPairShape<FM, FN, WGM, WGN>) {
    if (t.nt < 0) {
      diag_pair_x3<FM, WGM, WGN>(prob, t, sA0, sA1, sB0, sB1);
      return;
    }
  }
  gemm_tile_x3<FM, FN, WGM, WGN>(prob, t, sA0, sA1, sB0, sB1);
}

// One output tile per workgroup. Two waves per SIMD: four-wave workgroups run
// two per CU, eight-wave workgroups one.
template <int FM, int FN, int WGM, int WGN, class Prob>
__global__ void __launch_bounds__(64 * WGM * WGN, 2) k_gemm_x3(const Prob prob) {
  constexpr int BM = 16 * FM * WGM, BN = 16 * FN * WGN;
  __shared__ __align__(16) uint16_t sA0[2 * BM * kLDK], sA1[2 * BM * kLDK];
  __shared__ __align__(16) uint16_t sB0[2 * BN * kLDK], sB1[2 * BN * kLDK];
  Tile t;
  if (!prob.tile(xcd_remap(blockIdx.x, gridDim.x), BM, BN, t)) return;
  run_tile_x3<FM, FN, WGM, WGN>(prob, t, sA0, sA1, sB0, sB1);
}

// Grouped form over a (slot, local tile) table (see jk_kernels.h): one workgroup
// per tile. Workgroup b runs on XCD x = b % 8 and takes entry b / 8 of that XCD's
// run [xcd_bounds[x], xcd_bounds[x + 1]) of the table; the grid covers the longest
// run and the surplus workgroups of shorter runs exit at once.
template <int FM, int FN, int WGM, int WGN, class Prob>
__global__ void __launch_bounds__(64 * WGM * WGN, 2)
    k_gemm_x3_grouped(const Prob* __restrict__ probs, const int2* __restrict__ tiles,
                      const int* __restrict__ xcd_bounds) {
  constexpr int BM = 16 * FM * WGM, BN = 16 * FN * WGN;
  __shared__ __align__(16) uint16_t sA0[2 * BM * kLDK], sA1[2 * BM * kLDK];
  __shared__ __align__(16) uint16_t sB0[2 * BN * kLDK], sB1[2 * BN * kLDK];
  const int x = blockIdx.x % kXcds;
  const int id = xcd_bounds[x] + (int)(blockIdx.x / kXcds);
  if (id >= xcd_bounds[x + 1]) return;
  // The entry is uniform over the workgroup: readfirstlane lets the problem
  // struct come in through scalar loads.
  const int2 e = tiles[id];
  const int slot = __builtin_amdgcn_readfirstlane(e.x), local = __builtin_amdgcn_readfirstlane(e.y);
  const Prob prob = probs[slot];
  Tile t;
  if (prob.tile(local, BM, BN, t)) run_tile_x3<FM, FN, WGM, WGN>(prob, t, sA0, sA1, sB0, sB1);
}

// ---------------------------------------------------------------- sources
struct SrcRows {  // element (r, k) at r_clamped*ld + k
  Planes p;
  int64_t base, ld;
  int rmax;
  __device__ __forceinline__ int64_t off(int r) const { return base + (int64_t)(r < rmax ? r : rmax - 1) * ld; }
  __device__ __forceinline__ int64_t at(int r, int k0) const { return off(r) + k0; }
};

// Blocked rows: 16-row x 32-k blocks of 512 elements (1 KB per plane), block (r / 16, k / 32)
// at ((r / 16) * kblocks + k / 32) * 512, row r % 16 of it at (r % 16) * 32. One DMA instruction
// (16 rows x 32 k) then reads one contiguous 1 KB, whole cache lines, where row-major rows give
// it 16 half lines: the split half-transform measured 5-8% faster with blocked operands.
__host__ __device__ inline int64_t blk_index(int64_t r, int64_t k, int64_t kblocks) {
  return ((r >> 4) * kblocks + (k >> 5)) * 512 + (r & 15) * 32 + (k & 31);
}
struct SrcBlocked {
  Planes p;
  int64_t kblocks;
  int64_t rmax;
  __device__ __forceinline__ int64_t at(int64_t r, int k0) const { return blk_index(r < rmax ? r : rmax - 1, k0, kblocks); }
};

template <int FM, int FN>
__device__ __forceinline__ void coords(int lane, int wm, int wn, int i, int j, int r, int& dm, int& dn) {
  dm = wm + 16 * i + out_row(lane, r);
  dn = wn + 16 * j + (lane & 15);
}

// X16[q][n] = split(sum_mu Ct[q][mu] bht[n][mu]), n = blk_index(nu, P, naux_p / 32): row q of
// X16 is the (nu, P) matrix of orbital q in 16 x 32 blocks, written contiguously.
struct ProbHT {
  static constexpr int kA = RK, kB = RK;
  static constexpr bool kRowSkip = false;
  static constexpr bool kLower = false;
  Planes ct, bht;
  uint16_t* xhi;
  uint16_t* xlo;
  int nbf5, ldk, mtiles, m_base;  // rows [m_base, m_base + mtiles*BM) of this launch
  int64_t N;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = m_base + t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __host__ __device__ int ksteps() const { return ldk / kBK; }
  __device__ int row_limit(const Tile&) const { return nbf5; }
  __device__ SrcBlocked a_src(const Tile&) const { return {ct, ldk / kBK, nbf5}; }
  __device__ SrcBlocked b_src(const Tile&) const { return {bht, ldk / kBK, N}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4f (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int q = t.m0 + dm;
          const int64_t n = (int64_t)t.n0 + dn;
          if (q < nbf5 && n < N) {
            const float v = acc[i][j][r];
            const uint16_t h = bf16_rne(v);
            gmem(xhi)[(int64_t)q * N + n] = h;
            gmem(xlo)[(int64_t)q * N + n] = bf16_rne(v - bf16_to_f(h));
          }
        }
  }
};

// The split triangular MO route (builds without J and K at nbf5 < nbf), as the FP64 one:
// b = L + L^T with L[mu][nu] = b[mu][nu] for nu < mu and b[mu][mu] / 2 on the diagonal, so
// Y_pq = W_pq + W_qp with W = C^T T and T = L C.
//
// T16[mu][q][P] = split(sum_{nu <= mu} Ct[q][nu] L[mu][nu][P]): batch mu is a GEMM with
// K = mu + 1. Its B rows (k = nu, KN: P contiguous) are the packed rows tri(mu, nu), nu < mu,
// of bt16, then row mu of the halved diagonal rows stored after the packed rows (row
// hd + mu), then their zero row (hd + nbf) for nu > mu in the last k-step. Columns past
// naux_p (a partial last tile) read the row's last 8 and are not stored.
struct SrcTriL {
  Planes p;
  int64_t kb, base, hd;  // base = tri(mu, 0)
  int mu, nbf, naux_p;
  __device__ __forceinline__ int64_t at(int k, int n) const {
    const int64_t row = k < mu ? base + k : (k == mu ? hd + mu : hd + nbf);
    return blk_index(row, n < naux_p ? n : naux_p - 8, kb);
  }
};

// Batches are dealt so that every XCD's contiguous range of workgroups holds short and long
// ones alike: batch b is mu = b / 2 for even b and nbf - 1 - b / 2 for odd b.
struct ProbTL {
  static constexpr int kA = RK, kB = KN;
  static constexpr bool kRowSkip = true;  // rows q >= nbf5
  static constexpr bool kLower = false;
  Planes ct, bt;
  uint16_t* thi;
  uint16_t* tlo;
  int nbf, nbf5, ldk, naux_p, mtiles, ntiles;
  int64_t kb, hd;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    const int per = mtiles * ntiles, b = id / per, r = id - b * per;
    t.batch = (b & 1) ? nbf - 1 - (b >> 1) : (b >> 1);
    t.mt = r % mtiles;
    t.nt = r / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __host__ __device__ int ksteps() const { return ldk / kBK; }
  __device__ int ksteps(const Tile& t) const { return t.batch / kBK + 1; }
  __device__ int row_limit(const Tile&) const { return nbf5; }
  __device__ SrcBlocked a_src(const Tile&) const { return {ct, ldk / kBK, nbf5}; }
  __device__ SrcTriL b_src(const Tile& t) const {
    return {bt, kb, (int64_t)t.batch * (t.batch + 1) / 2, hd, t.batch, nbf, naux_p};
  }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4f (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
    const int64_t row0 = (int64_t)t.batch * nbf5;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int q = t.m0 + dm, P = t.n0 + dn;
          if (q < nbf5 && P < naux_p) {
            const float v = acc[i][j][r];
            const uint16_t h = bf16_rne(v);
            const int64_t o = (row0 + q) * naux_p + P;
            gmem(thi)[o] = h;
            gmem(tlo)[o] = bf16_rne(v - bf16_to_f(h));
          }
        }
  }
};

// W[p][n] = sum_mu Ct[p][mu] T16[mu][n] over the N = nbf5 naux_p columns n = (q, P) (FP32
// output). T16 rows mu >= nbf (the zero tail of Ct's k range) read row nbf - 1; columns past N
// (a partial last tile) read the row's last 8 and are not stored.
struct SrcT16 {
  Planes p;
  int64_t N;
  int nbf;
  __device__ __forceinline__ int64_t at(int k, int64_t n) const {
    return (int64_t)(k < nbf ? k : nbf - 1) * N + (n < N ? n : N - 8);
  }
};

struct ProbWS {
  static constexpr int kA = RK, kB = KN;
  static constexpr bool kRowSkip = true;  // rows p >= nbf5
  static constexpr bool kLower = false;
  Planes ct, tt;
  float* W;
  int nbf, nbf5, ldk, mtiles;
  int64_t N;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __host__ __device__ int ksteps() const { return ldk / kBK; }
  __device__ int row_limit(const Tile&) const { return nbf5; }
  __device__ SrcBlocked a_src(const Tile&) const { return {ct, ldk / kBK, nbf5}; }
  __device__ SrcT16 b_src(const Tile&) const { return {tt, N, nbf}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4f (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int p = t.m0 + dm;
          const int64_t n = (int64_t)t.n0 + dn;
          if (p < nbf5 && n < N) gmem(W)[(int64_t)p * N + n] = acc[i][j][r];
        }
  }
};

// Batch p: Y_p[q][P] = sum_nu Ct[q][nu] X16[p][nu][P] for the cyclic window of rows
// q = (p + r) mod nbf5, r < qt_window(nbf5) (jk_common.h): every batch carries nbf5 / 2 + 1
// live rows, so a 160-row tile holds a whole window for nbf5 <= 318 and copies only the C^T
// rows it uses. Writes d[p][P] = Y_p[p][P] (window row 0; FP64 and split planes) and the
// per-64-column sums of Y_p[q]^2 (FP64) of the pairs the batch owns into kpart at
// (max(p, q), min(p, q)). Tiles are 64 or 128 aux columns wide (FN = 2 or 4); both give the
// same kpart bits: a 64-column partial is always S(first 32 columns) + S(last 32 columns),
// S being one wave's sum over two 16-column blocks reduced across 16 lanes.
// Element (k = nu, n = P) of X16[p]: base + blk_index(min(nu, nbf - 1), P, naux_p / 32).
struct SrcXT {
  Planes p;
  int64_t base, kblocks;
  int kmax;
  __device__ __forceinline__ int64_t at(int k, int n) const { return base + blk_index(k < kmax ? k : kmax - 1, n, kblocks); }
};

struct SrcWindow {  // window row r of batch b: C^T row qt_row(b, min(r, h - 1), n) of the blocked C^T planes
  Planes p;
  int64_t kblocks;
  int b, h, n;
  __device__ __forceinline__ int64_t at(int r, int k0) const {
    return blk_index(qt_row(b, r < h ? r : h - 1, n), k0, kblocks);
  }
};

struct ProbQT {
  static constexpr int kA = RK, kB = KN;
  static constexpr bool kRowSkip = true;
  static constexpr bool kLower = false;
  Planes ct, x;
  double* kpart;
  double* d;
  uint16_t* dhi;
  uint16_t* dlo;
  int nbf, naux_p, nbf5, ldk, mtiles, ntiles, nt64, h;  // nt64 = naux_p / 64 (kpart stride); h window rows
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.mt = id % mtiles;
    const int rest = id / mtiles;
    t.nt = rest % ntiles;
    t.batch = rest / ntiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __host__ __device__ int ksteps() const { return ldk / kBK; }
  __device__ int row_limit(const Tile&) const { return h; }
  __device__ SrcWindow a_src(const Tile& t) const { return {ct, ldk / kBK, t.batch, h, nbf5}; }
  __device__ SrcXT b_src(const Tile& t) const { return {x, t.batch * x16_stride(nbf, naux_p), naux_p / 32, nbf}; }
  // kpart slot of window row w of batch p, or -1 when the row is past the window or its pair
  // belongs to the other batch.
  __device__ __forceinline__ int64_t kslot(int p, int w) const {
    if (w >= h || !qt_owns(p, w, nbf5)) return -1;
    const int q = qt_row(p, w, nbf5);
    const int hi = p > q ? p : q, lo = p > q ? q : p;
    return ((int64_t)hi * nbf5 + lo) * nt64;
  }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4f (&acc)[FM][FN], int wm, int wn, int lane, double* lds) const {
    static_assert(FN == 2 || FN == 4, "quarter-transform tiles are 64 or 128 aux columns wide");
    const int p = t.batch;
    if (t.m0 == 0 && wm == 0) {  // window row 0 is q = p: d[p][P] = Y_p[p][P]
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          coords<FM, FN>(lane, wm, wn, 0, j, r, dm, dn);
          if (dm == 0 && t.n0 + dn < naux_p) {
            const int64_t o = (int64_t)p * naux_p + t.n0 + dn, ob = blk_index(p, t.n0 + dn, naux_p / 32);
            const float v = acc[0][j][r];
            gmem(d)[o] = v;
            const uint16_t hb = bf16_rne(v);
            gmem(dhi)[ob] = hb;  // the split planes of d are blocked (J^n's A operand)
            gmem(dlo)[ob] = bf16_rne(v - bf16_to_f(hb));
          }
        }
    }
    double s[FM][4];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        double tot = 0.0;
#pragma unroll
        for (int hh = 0; hh < FN / 2; ++hh) {
          double v = 0.0;
#pragma unroll
          for (int j = 2 * hh; j < 2 * hh + 2; ++j) v += (double)acc[i][j][r] * (double)acc[i][j][r];
#pragma unroll
          for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
          tot = hh == 0 ? v : tot + v;
        }
        s[i][r] = tot;
      }
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if constexpr (FN == 4) {  // each wave covers one 64-column block
      const int c0 = t.n0 + wn;
      if ((lane & 15) == 0 && c0 < naux_p) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t k = kslot(p, t.m0 + wm + 16 * i + out_row(lane, r));
            if (k >= 0) gmem(kpart)[k + c0 / 64] = s[i][r];
          }
      }
      (void)wave;
      (void)lds;
    } else {  // the two waves of a row pair cover one 64-column block: add through LDS
      if ((wave & 1) && (lane & 15) == 0) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) lds[wm + 16 * i + out_row(lane, r)] = s[i][r];
      }
      __syncthreads();
      if (!(wave & 1) && (lane & 15) == 0) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int dm = wm + 16 * i + out_row(lane, r);
            const int64_t k = kslot(p, t.m0 + dm);
            if (k >= 0) gmem(kpart)[k + t.n0 / 64] = s[i][r] + lds[dm];
          }
      }
    }
  }
};

// Orbital-gradient MO integrals in BF16x3 (integrals.py:703-716): batch q,
//   b_MO[p][q][P] = sum_nu Ctf16[p][nu] X16[q][nu][P]   for p < nbf, q < nbf5,
// with Ctf16 the split planes of C^T over all nbf columns; FP64 output.
// kPlanes: the epilogue also writes the bf16 split planes of b_MO at the same index (the
// Lagrangian's exchange operand), bit for bit what a conversion pass over b_MO would give.
template <bool kPlanes>
struct ProbMOT {
  static constexpr int kA = RK, kB = KN;
  static constexpr bool kRowSkip = true;  // rows p >= nbf of the last row tile
  static constexpr bool kLower = false;
  Planes ct, x;
  double* out;
  uint16_t* oh;
  uint16_t* ol;
  int64_t ldrow;  // output row stride (nbf5 * naux for the reference layout)
  int nbf, naux, naux_p, nbf5, ldk, mtiles, ntiles;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.mt = id % mtiles;
    const int rest = id / mtiles;
    t.nt = rest % ntiles;
    t.batch = rest / ntiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __host__ __device__ int ksteps() const { return ldk / kBK; }
  __device__ int row_limit(const Tile&) const { return nbf; }
  __device__ SrcBlocked a_src(const Tile&) const { return {ct, ldk / kBK, nbf}; }
  __device__ SrcXT b_src(const Tile& t) const { return {x, t.batch * x16_stride(nbf, naux_p), naux_p / 32, nbf}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4f (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < nbf && n < naux) {
            const int64_t idx = (int64_t)m * ldrow + (int64_t)t.batch * naux + n;
            const double v = acc[i][j][r];
            gmem(out)[idx] = v;
            if (kPlanes) {
              uint16_t h, l;
              split2(v, h, l);
              gmem(oh)[idx] = h;
              gmem(ol)[idx] = l;
            }
          }
        }
  }
};

// Exchange term of the orbital-gradient Lagrangian in BF16x3 (utils.py:353-355), split
// over k' = (q, P) in chunks of kc: part[c][a][b] = sum_{k' in chunk c} B16[a][k'] W16[b][k'],
// W = ck[b][q] B[b][k'], both operands rows of stride ldrow (zero past nbf5 * naux).
struct ProbLagX {
  static constexpr int kA = RK, kB = RK;
  static constexpr bool kRowSkip = false;
  static constexpr bool kLower = false;
  Planes bm, w;
  double* part;
  int64_t ldrow, kc;
  int nbf, nbf5, mtiles, ntiles;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.mt = id % mtiles;
    const int rest = id / mtiles;
    t.nt = rest % ntiles;
    t.batch = rest / ntiles;  // k chunk
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __host__ __device__ int ksteps() const { return (int)(kc / kBK); }
  __device__ int row_limit(const Tile&) const { return nbf; }
  __device__ SrcRows a_src(const Tile& t) const { return {bm, (int64_t)t.batch * kc, ldrow, nbf}; }
  __device__ SrcRows b_src(const Tile& t) const { return {w, (int64_t)t.batch * kc, ldrow, nbf5}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4f (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < nbf && n < nbf5) gmem(part)[((int64_t)t.batch * nbf + m) * nbf5 + n] = acc[i][j][r];
        }
  }
};

// J[q][mu][nu] = sum_P d16[q][P] btri16[t][P] for mu >= nu (FP64 output; the
// upper triangle is filled afterwards by launch_symmetrize)
struct ProbJN {
  static constexpr int kA = RK, kB = RK;
  static constexpr bool kRowSkip = true;
  static constexpr bool kLower = false;
  Planes dd, bt;
  double* J;
  int nbf, naux_p, nbf5, mtiles, m_base;  // rows [m_base, m_base + mtiles*BM) of this launch
  int64_t ntri;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = m_base + t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __host__ __device__ int ksteps() const { return naux_p / kBK; }
  __device__ int row_limit(const Tile&) const { return nbf5; }
  __device__ SrcBlocked a_src(const Tile&) const { return {dd, naux_p / kBK, nbf5}; }  // d16, bt16 are blocked
  __device__ SrcBlocked b_src(const Tile&) const { return {bt, naux_p / kBK, ntri}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4f (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
    const int64_t nn = (int64_t)nbf * nbf;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      int dm, dn;
      coords<FM, FN>(lane, wm, wn, 0, j, 0, dm, dn);
      const int64_t tt = (int64_t)t.n0 + dn;
      if (tt >= ntri) continue;
      const int mu = tri_row(tt);
      const int nu = (int)(tt - (int64_t)mu * (mu + 1) / 2);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int q = t.m0 + dm;
          if (q < nbf5) {
            gmem(J)[(int64_t)q * nn + (int64_t)mu * nbf + nu] = acc[i][j][r];
          }
        }
    }
  }
};

// K[b][i][j] = sum_P X16[b][i][P] X16[b][j][P] on the lower triangle of tiles (FP64
// output). Each value with i >= j goes to (i, j) and, for i > j, also to (j, i): the
// output is exactly symmetric without a separate pass (the upper half of a diagonal
// tile, whose own products pair the planes the other way round, is not stored).
// With npair > 0 (the tile shapes of kDiagPairShape) the first npair tiles of a batch are
// diagonal pair tiles (diag_pair_x3: panels 2s and 2s + 1 in one tile, nt = -1), and the
// others are the lower-triangle tiles they do not cover (syrk_tile).
struct ProbSYRK {
  static constexpr int kA = RK, kB = RK;
  static constexpr bool kRowSkip = true;
  static constexpr bool kLower = true;  // rows above the diagonal are never stored
  static constexpr bool kDiagPairs = true;
  Planes x;
  double* O;
  int64_t sa, ldo, so;  // sa: X16 elements per batch (blocked rows of K columns)
  int M, K, tiles_per_batch, npair;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = id / tiles_per_batch;
    syrk_tile(id - t.batch * tiles_per_batch, npair, t.mt, t.nt);
    t.m0 = t.mt * BM;
    t.n0 = t.nt < 0 ? t.m0 + BM : t.nt * BN;
    return true;
  }
  __host__ __device__ int ksteps() const { return K / kBK; }
  __device__ int row_limit(const Tile&) const { return M; }
  __device__ SrcBlocked a_src(const Tile& t) const { return {{x.hi + t.batch * sa, x.lo + t.batch * sa}, K / kBK, M}; }
  __device__ SrcBlocked b_src(const Tile& t) const { return a_src(t); }
  // The 16 x 16 block of rows m0 + 16 g and columns m0 + 16 c of a diagonal pair tile.
  __device__ void store_block(const Tile& t, const v4f& v, int g, int c, int lane) const {
    double* o = O + (int64_t)t.batch * so;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = t.m0 + 16 * g + out_row(lane, r), n = t.m0 + 16 * c + (lane & 15);
      if (m < M && n <= m) {
        gmem(o)[(int64_t)m * ldo + n] = v[r];
        if (n < m) gmem(o)[(int64_t)n * ldo + m] = v[r];
      }
    }
  }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4f (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
    double* o = O + (int64_t)t.batch * so;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < M && n <= m) {
            const double v = acc[i][j][r];
            gmem(o)[(int64_t)m * ldo + n] = v;
            if (n < m) gmem(o)[(int64_t)n * ldo + m] = v;
          }
        }
  }
};

// ---------------------------------------------------------------- conversions
__global__ void k_split_packed(const double* __restrict__ src, int64_t n, uint16_t* __restrict__ hi,
                               uint16_t* __restrict__ lo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint16_t h, l;
  split2(src[i], h, l);
  hi[i] = h;
  lo[i] = l;
}

// Split planes of a row-major rows x cols FP64 matrix in 16-row x 32-column blocks
// (blk_index, cols a multiple of 32), rows padded with zeros to rows_p (a multiple of 16).
// nb such matrices one after the other: source stride rows x cols, destination rows_p x cols.
// The destination is written in its own order: thread g fills elements [8 g, 8 g + 8) of
// both planes (one 16-byte store each) from 8 consecutive source values of one row.
__global__ void k_split_blocked(const double* __restrict__ src, int nb, int64_t rows, int64_t cols, int64_t rows_p,
                                uint16_t* __restrict__ hi, uint16_t* __restrict__ lo) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int kb = (int)(cols / 32), rb = (int)(rows_p / 16);
  const int64_t blk = g >> 6;  // 16 x 32 block: 64 chunks of 8
  if (blk >= (int64_t)nb * rb * kb) return;
  const int in = (int)(g & 63);
  const int64_t b = blk / ((int64_t)rb * kb);
  const int64_t rest = blk - b * rb * kb;
  const int64_t r = (rest / kb) * 16 + (in >> 2), c = (rest % kb) * 32 + 8 * (in & 3);
  u16x8 h, l;
  if (r < rows) {
    const double* p = src + (b * rows + r) * cols + c;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      uint16_t x, y;
      split2(p[e], x, y);
      h[e] = x;
      l[e] = y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 8; ++e) h[e] = l[e] = 0;
  }
  *reinterpret_cast<u16x8*>(hi + 8 * g) = h;
  *reinterpret_cast<u16x8*>(lo + 8 * g) = l;
}

// bht row n = blk_index(nu, P, naux_p / 32) holds b[mu][nu][P] over mu < ldk (zero for mu >= nbf
// and for nu >= nbf), rows in 16 x 32 blocks over (n, mu); from btri[tri(mu,nu)][P]:
// a 32 (mu) x 64 (P) tile through LDS per workgroup; grid (P tiles, mu tiles, nu < nbf
// rounded up to 16).
__global__ void k_build_bht(const double* __restrict__ btri, int nbf, int naux_p, int ldk, uint16_t* __restrict__ hi,
                            uint16_t* __restrict__ lo) {
  __shared__ double tileb[32][65];
  const int nu = blockIdx.z, mu0 = blockIdx.y * 32, p0 = blockIdx.x * 64;
  for (int e = threadIdx.x; e < 32 * 64; e += blockDim.x) {
    const int m = e / 64, pp = e % 64;
    const int mu = mu0 + m, p = p0 + pp;
    double v = 0.0;
    if (mu < nbf && nu < nbf && p < naux_p) v = btri[tri_index(mu, nu) * naux_p + p];
    tileb[m][pp] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 64; e += blockDim.x) {
    const int pp = e / 32, m = e % 32;
    const int mu = mu0 + m, p = p0 + pp;
    if (p < naux_p && mu < ldk) {
      uint16_t h, l;
      split2(tileb[m][pp], h, l);
      const int64_t o = blk_index(blk_index(nu, p, naux_p / 32), mu, ldk / 32);
      hi[o] = h;
      lo[o] = l;
    }
  }
}

// Ct[q][mu] (q < rows, mu < ldk) from Cp[mu][q] (zero outside nbf x nbf5).
__global__ void k_split_ct(const double* __restrict__ Cp, int ldcp, int nbf, int nbf5, int rows_p, int ldk,
                           uint16_t* __restrict__ hi, uint16_t* __restrict__ lo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows_p * ldk) return;
  const int q = (int)(i / ldk), mu = (int)(i % ldk);
  const double v = (q < nbf5 && mu < nbf) ? Cp[(int64_t)mu * ldcp + q] : 0.0;
  uint16_t h, l;
  split2(v, h, l);
  const int64_t o = blk_index(q, mu, ldk / 32);
  hi[o] = h;
  lo[o] = l;
}

template <int FM, int FN, int WGM, int WGN, class Prob>
hipError_t grouped_launch(const void* probs, const int2* tiles, const int* xcd_bounds, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_gemm_x3_grouped<FM, FN, WGM, WGN, Prob>), dim3(grid), dim3(64 * WGM * WGN), 0, s,
                     static_cast<const Prob*>(probs), tiles, xcd_bounds);
  return hipGetLastError();
}

// Tile shapes are named (fm, fn) for a BM x BN = (32 fm) x (32 fn) output tile.
// All but (10, 4) use 2 x 2 waves of (fm, fn) MFMA blocks; (10, 4) is 4 x 2 waves
// of (5, 4) blocks (8 waves), offered to the half-transform and J^n, whose M
// extent (nbf5) it covers in one tile for nbf5 <= 320.
template <class Prob>
hipError_t launch_x3(const Prob& prob, TileF64 tc, int nblocks, hipStream_t s, GroupCollector* col = nullptr) {
  if (nblocks <= 0) return hipSuccess;
  const dim3 grid(nblocks);
#define PNJK_CASE(FMV, FNV, WFM, WFN, WGMV, WGNV)                                                          \
  if (tc.fm == FMV && tc.fn == FNV) {                                                                      \
    if (col) {                                                                                             \
      col->add(prob, tc, nblocks, 32 * FMV, 32 * FNV, prob.ksteps(), &grouped_launch<WFM, WFN, WGMV, WGNV, Prob>, \
               __PRETTY_FUNCTION__);                                                                       \
      return hipSuccess;                                                                                   \
    }                                                                                                      \
    hipLaunchKernelGGL((k_gemm_x3<WFM, WFN, WGMV, WGNV, Prob>), grid, dim3(64 * WGMV * WGNV), 0, s, prob); \
    return hipGetLastError();                                                                              \
  }
  PNJK_CASE(2, 2, 2, 2, 2, 2)
  PNJK_CASE(4, 4, 4, 4, 2, 2)
  PNJK_CASE(5, 4, 5, 4, 2, 2)
  PNJK_CASE(4, 2, 4, 2, 2, 2)
  PNJK_CASE(5, 2, 5, 2, 2, 2)
  PNJK_CASE(2, 4, 2, 4, 2, 2)
  if constexpr (std::is_same<Prob, ProbHT>::value || std::is_same<Prob, ProbJN>::value) {
    PNJK_CASE(10, 4, 5, 4, 4, 2)
  }
  if constexpr (std::is_same<Prob, ProbMOT<true>>::value || std::is_same<Prob, ProbMOT<false>>::value) {
    PNJK_CASE(6, 4, 6, 4, 2, 2)
  }
  if constexpr (std::is_same<Prob, ProbHT>::value || std::is_same<Prob, ProbJN>::value ||
                std::is_same<Prob, ProbSYRK>::value || std::is_same<Prob, ProbLagX>::value) {
    PNJK_CASE(5, 5, 5, 5, 2, 2)
  }
#undef PNJK_CASE
  return hipErrorInvalidValue;
}

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace

hipError_t launch_split_blocked(const double* src, int64_t rows, int64_t cols, uint16_t* hi, uint16_t* lo,
                                hipStream_t s) {
  if (cols % 32 != 0) return hipErrorInvalidValue;
  const int64_t rows_p = (rows + 15) / 16 * 16, n = rows_p * cols / 8;
  hipLaunchKernelGGL(k_split_blocked, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, src, 1, rows, cols, rows_p, hi,
                     lo);
  return hipGetLastError();
}

hipError_t launch_split_x16(const double* X, const SplitGeom& g, uint16_t* hi, uint16_t* lo, hipStream_t s) {
  if (g.naux_p % 32 != 0) return hipErrorInvalidValue;
  const int64_t rows_p = (g.nbf + 15) / 16 * 16, n = (int64_t)g.nbf5 * rows_p * g.naux_p / 8;
  hipLaunchKernelGGL(k_split_blocked, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, X, g.nbf5, (int64_t)g.nbf,
                     (int64_t)g.naux_p, rows_p, hi, lo);
  return hipGetLastError();
}

// W16[b][k'] = split(ck[b][q] B[b][k']) for k' = q naux + P < nbf5 naux, zero up to ldrow.
__global__ void k_lag_w16(const double* __restrict__ B, int64_t ldrow, int naux, int nbf5,
                          const double* __restrict__ ck, uint16_t* __restrict__ hi, uint16_t* __restrict__ lo) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nbf5 * ldrow) return;
  const int b = (int)(i / ldrow);
  const int q = (int)((i - (int64_t)b * ldrow) / naux);
  uint16_t h, l;
  split2(q < nbf5 ? ck[(int64_t)b * nbf5 + q] * B[i] : 0.0, h, l);
  hi[i] = h;
  lo[i] = l;
}

hipError_t launch_lag_x3(int nbf, int nbf5, int naux, const double* B, int64_t ldrow, int64_t kc, int nchunks,
                         const double* ck, uint16_t* bh, uint16_t* bl, uint16_t* wh, uint16_t* wl, double* part,
                         bool planes_ready, hipStream_t s) {
  if (kc % kBK != 0 || (int64_t)nchunks * kc != ldrow) return hipErrorInvalidValue;
  const int64_t nb = (int64_t)nbf * ldrow, nw = (int64_t)nbf5 * ldrow;
  if (!planes_ready) hipLaunchKernelGGL(k_split_packed, dim3((unsigned)cdiv(nb, 256)), dim3(256), 0, s, B, nb, bh, bl);
  hipLaunchKernelGGL(k_lag_w16, dim3((unsigned)cdiv(nw, 256)), dim3(256), 0, s, B, ldrow, naux, nbf5, ck, wh, wl);
  // 160 x 160 tiles: each b_MO row block is read by 2 column tiles over nbf5 = 300, not by 5
  // as with 160 x 64 (the whole split Lagrangian 24.9 -> 22.9 ms at nbf = 1000).
  const TileF64 tc{5, 5};
  ProbLagX pr;
  pr.bm = {bh, bl};
  pr.w = {wh, wl};
  pr.part = part;
  pr.ldrow = ldrow;
  pr.kc = kc;
  pr.nbf = nbf;
  pr.nbf5 = nbf5;
  pr.mtiles = cdiv(nbf, 32 * tc.fm);
  pr.ntiles = cdiv(nbf5, 32 * tc.fn);
  return launch_x3(pr, tc, pr.mtiles * pr.ntiles * nchunks, s);
}

hipError_t launch_build_bht(const double* btri, int nbf, int naux_p, int ldk, uint16_t* hi, uint16_t* lo,
                            hipStream_t s) {
  const dim3 grid(cdiv(naux_p, 64), cdiv(ldk, 32), (nbf + 15) / 16 * 16);
  hipLaunchKernelGGL(k_build_bht, grid, dim3(256), 0, s, btri, nbf, naux_p, ldk, hi, lo);
  return hipGetLastError();
}

hipError_t launch_split_ct(const double* Cp, int ldcp, int nbf, int nbf5, int rows, int ldk, uint16_t* hi,
                           uint16_t* lo, hipStream_t s) {
  const int rows_p = (rows + 15) / 16 * 16;  // whole 16-row blocks
  const int64_t n = (int64_t)rows_p * ldk;
  hipLaunchKernelGGL(k_split_ct, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, Cp, ldcp, nbf, nbf5, rows_p, ldk, hi,
                     lo);
  return hipGetLastError();
}

hipError_t launch_ht_x3(const SplitGeom& g, TileF64 tc, const uint16_t* ct_hi, const uint16_t* ct_lo,
                        const uint16_t* bht_hi, const uint16_t* bht_lo, uint16_t* x_hi, uint16_t* x_lo,
                        hipStream_t s, GroupCollector* col) {
  ProbHT pr;
  pr.ct = {ct_hi, ct_lo};
  pr.bht = {bht_hi, bht_lo};
  pr.xhi = x_hi;
  pr.xlo = x_lo;
  pr.nbf5 = g.nbf5;
  pr.ldk = g.ldk;
  pr.N = x16_stride(g.nbf, g.naux_p);
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  pr.m_base = 0;
  pr.mtiles = cdiv(g.nbf5, BM);
  return launch_x3(pr, tc, pr.mtiles * cdiv(pr.N, BN), s, col);
}

hipError_t launch_tl_x3(const SplitGeom& g, const uint16_t* ct_hi, const uint16_t* ct_lo, const uint16_t* bt_hi,
                        const uint16_t* bt_lo, int64_t hd, uint16_t* t_hi, uint16_t* t_lo, hipStream_t s) {
  static_assert(HasTileKsteps<ProbTL>::value, "batch mu runs mu / 32 + 1 k-steps");
  static_assert(!HasTileKsteps<ProbWS>::value && !HasTileKsteps<ProbHT>::value, "one k-step count per launch");
  if (g.naux_p % 32 != 0 || g.naux_p < 32) return hipErrorInvalidValue;
  const TileF64 tc{5, 4};  // 160 x 128 (KN tiles are 64 or 128 columns wide)
  ProbTL pr;
  pr.ct = {ct_hi, ct_lo};
  pr.bt = {bt_hi, bt_lo};
  pr.thi = t_hi;
  pr.tlo = t_lo;
  pr.nbf = g.nbf;
  pr.nbf5 = g.nbf5;
  pr.ldk = g.ldk;
  pr.naux_p = g.naux_p;
  pr.mtiles = cdiv(g.nbf5, 32 * tc.fm);
  pr.ntiles = cdiv(g.naux_p, 32 * tc.fn);
  pr.kb = g.naux_p / 32;
  pr.hd = hd;
  const int64_t blocks = (int64_t)g.nbf * pr.mtiles * pr.ntiles;
  if (blocks > INT32_MAX) return hipErrorInvalidValue;
  return launch_x3(pr, tc, (int)blocks, s);
}

hipError_t launch_ws_x3(const SplitGeom& g, const uint16_t* ct_hi, const uint16_t* ct_lo, const uint16_t* t_hi,
                        const uint16_t* t_lo, float* W, hipStream_t s) {
  const TileF64 tc{5, 4};
  ProbWS pr;
  pr.ct = {ct_hi, ct_lo};
  pr.tt = {t_hi, t_lo};
  pr.W = W;
  pr.nbf = g.nbf;
  pr.nbf5 = g.nbf5;
  pr.ldk = g.ldk;
  pr.N = (int64_t)g.nbf5 * g.naux_p;
  pr.mtiles = cdiv(g.nbf5, 32 * tc.fm);
  const int64_t blocks = (int64_t)pr.mtiles * cdiv(pr.N, 32 * tc.fn);
  if (pr.N + 32 * tc.fn > INT32_MAX || blocks > INT32_MAX) return hipErrorInvalidValue;
  return launch_x3(pr, tc, (int)blocks, s);
}

hipError_t launch_qt_x3(const SplitGeom& g, TileF64 tc, const uint16_t* ct_hi, const uint16_t* ct_lo,
                        const uint16_t* x_hi, const uint16_t* x_lo, double* kpart, double* d, uint16_t* d_hi,
                        uint16_t* d_lo, int* ntiles, hipStream_t s, GroupCollector* col, int window) {
  // naux_p is a multiple of 64; a 128-column tile may cover 64 columns past it,
  // which the B loader reads from the workspace slack and the epilogue drops.
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  if (tc.fn != 2 && tc.fn != 4) return hipErrorInvalidValue;
  ProbQT pr;
  pr.ct = {ct_hi, ct_lo};
  pr.x = {x_hi, x_lo};
  pr.kpart = kpart;
  pr.d = d;
  pr.dhi = d_hi;
  pr.dlo = d_lo;
  pr.nbf = g.nbf;
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ldk = g.ldk;
  if (window < 0 || window > qt_window(g.nbf5)) return hipErrorInvalidValue;
  pr.h = window ? window : qt_window(g.nbf5);
  pr.mtiles = cdiv(pr.h, BM);
  pr.ntiles = cdiv(g.naux_p, BN);
  pr.nt64 = g.naux_p / 64;
  *ntiles = pr.nt64;
  return launch_x3(pr, tc, pr.mtiles * pr.ntiles * g.nbf5, s, col);
}

template <bool kPlanes>
hipError_t launch_mo_x3_t(const SplitGeom& g, int naux, TileF64 tc, const uint16_t* ct_hi, const uint16_t* ct_lo,
                          const uint16_t* x_hi, const uint16_t* x_lo, double* b_mo, int64_t ldrow, hipStream_t s,
                          uint16_t* bmo_hi, uint16_t* bmo_lo) {
  // A 128-column tile may read up to 64 columns past naux_p (X16 slack) and drops them.
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  if ((tc.fn != 2 && tc.fn != 4) || ldrow < (int64_t)g.nbf5 * naux) return hipErrorInvalidValue;
  ProbMOT<kPlanes> pr;
  pr.ct = {ct_hi, ct_lo};
  pr.x = {x_hi, x_lo};
  pr.out = b_mo;
  pr.oh = bmo_hi;
  pr.ol = bmo_lo;
  pr.ldrow = ldrow;
  pr.nbf = g.nbf;
  pr.naux = naux;
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ldk = g.ldk;
  pr.mtiles = cdiv(g.nbf, BM);
  pr.ntiles = cdiv(g.naux_p, BN);
  return launch_x3(pr, tc, pr.mtiles * pr.ntiles * g.nbf5, s);
}

hipError_t launch_mo_x3(const SplitGeom& g, int naux, TileF64 tc, const uint16_t* ct_hi, const uint16_t* ct_lo,
                        const uint16_t* x_hi, const uint16_t* x_lo, double* b_mo, int64_t ldrow, hipStream_t s,
                        uint16_t* bmo_hi, uint16_t* bmo_lo) {
  if (bmo_hi && bmo_lo)
    return launch_mo_x3_t<true>(g, naux, tc, ct_hi, ct_lo, x_hi, x_lo, b_mo, ldrow, s, bmo_hi, bmo_lo);
  return launch_mo_x3_t<false>(g, naux, tc, ct_hi, ct_lo, x_hi, x_lo, b_mo, ldrow, s, nullptr, nullptr);
}

// Zeroes columns [c0, c1) of `rows` rows of stride ld (the pad of the b_MO planes).
__global__ void k_fill_cols16(uint16_t* __restrict__ h, uint16_t* __restrict__ l, int rows, int64_t ld, int64_t c0,
                              int64_t c1) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t w = c1 - c0;
  if (idx >= (int64_t)rows * w) return;
  const int64_t off = (idx / w) * ld + c0 + idx % w;
  h[off] = 0;
  l[off] = 0;
}

hipError_t launch_fill_cols16(uint16_t* h, uint16_t* l, int rows, int64_t ld, int64_t c0, int64_t c1, hipStream_t s) {
  if (c1 <= c0 || rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_cols16, dim3((unsigned)cdiv((int64_t)rows * (c1 - c0), 256)), dim3(256), 0, s, h, l, rows,
                     ld, c0, c1);
  return hipGetLastError();
}

hipError_t launch_jn_x3(const SplitGeom& g, TileF64 tc, const uint16_t* d_hi, const uint16_t* d_lo,
                        const uint16_t* bt_hi, const uint16_t* bt_lo, double* J, hipStream_t s,
                        GroupCollector* col) {
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  ProbJN pr;
  pr.dd = {d_hi, d_lo};
  pr.bt = {bt_hi, bt_lo};
  pr.J = J;
  pr.nbf = g.nbf;
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ntri = (int64_t)g.nbf * (g.nbf + 1) / 2;
  pr.m_base = 0;
  pr.mtiles = cdiv(g.nbf5, BM);
  return launch_x3(pr, tc, pr.mtiles * cdiv(pr.ntri, BN), s, col);
}

hipError_t launch_kn_x3(const SplitGeom& g, TileF64 tc, const uint16_t* x_hi, const uint16_t* x_lo, double* K,
                        hipStream_t s, GroupCollector* col) {
  if (tc.fm != tc.fn) return hipErrorInvalidValue;
  const int BM = 32 * tc.fm;
  ProbSYRK pr;
  pr.x = {x_hi, x_lo};
  pr.O = K;
  pr.sa = x16_stride(g.nbf, g.naux_p);
  pr.ldo = g.nbf;
  pr.so = (int64_t)g.nbf * g.nbf;
  pr.M = g.nbf;
  pr.K = g.naux_p;
  const int side = cdiv(g.nbf, BM);
  // Diagonal pair tiles for the shapes whose kernels take them: 128 x 128 (4, 4), the
  // built-in choice, and 64 x 64 (2, 2).
  pr.npair = (tc.fm == 2 || tc.fm == 4) ? side / 2 : 0;
  pr.tiles_per_batch = side * (side + 1) / 2 - 2 * pr.npair;
  return launch_x3(pr, tc, pr.tiles_per_batch * g.nbf5, s, col);
}

}  // namespace pnjk
