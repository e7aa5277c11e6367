// FP64 kernels of the RI JK build for gfx950 (MI355X).
//
// Every dense contraction of the build is one instantiation of a single
// LDS-staged GEMM core on v_mfma_f64_16x16x4f64. A workgroup of 4 waves owns a
// BM x BN output tile (2 x 2 waves, each BM/2 x BN/2 as 16x16 MFMA blocks);
// operand tiles are staged through double-buffered LDS in k-major layout with
// a register prefetch of the next k-slice. Each contraction is a "problem"
// struct that supplies the tile schedule, the two operand sources and the
// epilogue; the contractions and their reference lines are:
//
//   HT   X[q][nu,P]  = sum_mu C[mu,q] b[mu,nu,P]          integrals.py:211, :353
//   QT   Y_p[q][P]   = sum_nu C[nu,q] X[p][nu][P]         integrals.py:354
//        -> d[p][P] = Y_p[p][P] (integrals.py:212), K_MO row sums (integrals.py:358)
//   SYRK K[q] = X[q] X[q]^T                               integrals.py:216
//        J_MO = d d^T                                     integrals.py:356
//   JN   J[q][mu,nu] = sum_P d[q][P] b[mu,nu,P]           integrals.py:214
//   HC   H_core[i] = sum C[mu,i] H[mu,nu] C[nu,i]         integrals.py:360
//
// b is stored once, as the packed lower triangle btri[tri(mu,nu)][P] (b is
// symmetric in mu,nu, integrals.py:35-49), halving its HBM footprint; the
// symmetric outputs are computed on one triangle of tiles; a tiled transpose
// (k_symmetrize) then fills the other triangle with coalesced reads and writes.
#include <type_traits>

#include "jk_common.h"
#include "jk_kernels.h"

namespace pnjk {
namespace {

constexpr int kThreads = 256;
constexpr int kPad = 2;  // LDS row padding (doubles); keeps 16-byte alignment

enum Kind { KM = 0, MK = 1, MKLANE = 2 };

// Global -> register -> LDS loader for one operand tile of ROWS x BK.
// LDS layout is k-major: lds[k * (ROWS + kPad) + r].
template <int KIND, int ROWS, int BK>
struct Loader;

// Source element (k, r) at src.at(k, r); r is contiguous in memory (pairs).
template <int ROWS, int BK>
struct Loader<KM, ROWS, BK> {
  static constexpr int kChunks = (ROWS / 2) * BK;
  static constexpr int kPer = (kChunks + kThreads - 1) / kThreads;
  v2d v[kPer];
  template <class S>
  __device__ __forceinline__ void load(const S& s, int r0, int k0) {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int c = threadIdx.x + i * kThreads;
      if (kChunks % kThreads == 0 || c < kChunks) {
        const int k = c / (ROWS / 2), rp = (c % (ROWS / 2)) * 2;
        v[i] = *gmem(reinterpret_cast<const v2d*>(s.at(k0 + k, r0 + rp)));
      }
    }
  }
  __device__ __forceinline__ void store(double* lds) const {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int c = threadIdx.x + i * kThreads;
      if (kChunks % kThreads == 0 || c < kChunks) {
        const int k = c / (ROWS / 2), rp = (c % (ROWS / 2)) * 2;
        *reinterpret_cast<v2d*>(lds + k * (ROWS + kPad) + rp) = v[i];
      }
    }
  }
};

// Source element (r, k) at src.row(r)[k]; k is contiguous in memory. MK gives
// consecutive lanes consecutive rows; MKLANE gives them consecutive k pairs of one
// row (BK/2 lanes per row), which coalesces the global loads of long k-contiguous
// rows: measured 5% faster on the whole FP64 Lagrangian (its split-K exchange
// GEMM) and 4% slower on the Coulomb GEMM, so each problem picks its kind.
template <int KIND, int ROWS, int BK>
struct LoaderMK {
  static constexpr int kChunks = ROWS * (BK / 2);
  static constexpr int kPer = (kChunks + kThreads - 1) / kThreads;
  v2d v[kPer];
  template <class S>
  __device__ __forceinline__ void load(const S& s, int r0, int k0) {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int c = threadIdx.x + i * kThreads;
      if (kChunks % kThreads == 0 || c < kChunks) {
        const int r = KIND == MK ? c % ROWS : c / (BK / 2), kp = KIND == MK ? (c / ROWS) * 2 : (c % (BK / 2)) * 2;
        v[i] = *gmem(reinterpret_cast<const v2d*>(s.row(r0 + r) + k0 + kp));
      }
    }
  }
  __device__ __forceinline__ void store(double* lds) const {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int c = threadIdx.x + i * kThreads;
      if (kChunks % kThreads == 0 || c < kChunks) {
        const int r = KIND == MK ? c % ROWS : c / (BK / 2), kp = KIND == MK ? (c / ROWS) * 2 : (c % (BK / 2)) * 2;
        lds[kp * (ROWS + kPad) + r] = v[i].x;
        lds[(kp + 1) * (ROWS + kPad) + r] = v[i].y;
      }
    }
  }
};
template <int ROWS, int BK>
struct Loader<MK, ROWS, BK> : LoaderMK<MK, ROWS, BK> {};
template <int ROWS, int BK>
struct Loader<MKLANE, ROWS, BK> : LoaderMK<MKLANE, ROWS, BK> {};

// k-steps of one tile: a problem whose k extent depends on the tile (a triangular
// iteration space) defines ksteps_tile(t, BK); the others have one extent for all tiles.
template <class Prob>
__device__ __forceinline__ auto tile_ksteps(const Prob& p, const Tile& t, int BK, int)
    -> decltype(p.ksteps_tile(t, BK)) {
  return p.ksteps_tile(t, BK);
}
template <class Prob>
__device__ __forceinline__ int tile_ksteps(const Prob& p, const Tile&, int BK, long) {
  return p.ksteps(BK);
}

// Problems whose diagonal tiles (mt == nt) hold only their lower triangle (Prob::kFoldDiag:
// the symmetric products A A^T), for the 128 x 128 tile.
template <class P, class = void>
struct HasFoldDiag : std::false_type {};
template <class P>
struct HasFoldDiag<P, std::void_t<decltype(P::kFoldDiag)>> : std::bool_constant<P::kFoldDiag> {};

// A diagonal tile of a symmetric product: its S = BM / 16 row strips hold the blocks
// (g, c), c <= g. Wave w takes strips S - 1 - w (columns 0..S-1-w) and w (columns 0..w),
// S + 1 blocks for every wave (accumulator c for the long strip, S - c for the short one),
// where the regular 2 x 2 wave split gives one wave the full 64 x 64 block below the
// diagonal and another nothing, so the tile took as long as a full one. Every block reads
// the same LDS fragments (A = its strip, B = its column) in the same order as in the
// regular split, so the values are the same bit for bit. The tile's A and B rows are the
// same, so one operand copy per k-step serves both.
template <int FM, int BK, class Prob>
__device__ __forceinline__ void gemm_diag_f64(const Prob& prob, const Tile& t, double* sA) {
  constexpr int BM = 32 * FM, S = BM / 16, NW = kThreads / 64;
  static_assert(S == 2 * NW, "two strips per wave");
  constexpr int SA = BK * (BM + kPad);
  const auto as = prob.a_src(t);  // A and B are the same rows here: one copy serves both
  const int nk = tile_ksteps(prob, t, BK, 0);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int gl = S - 1 - wave, gs = wave;

  Loader<Prob::kA, BM, BK> la;
  la.load(as, t.m0, 0);
  la.store(sA);
  __syncthreads();

  v4d acc[S + 1];
#pragma unroll
  for (int j = 0; j <= S; ++j) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  int kt = 0;
  do {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) la.load(as, t.m0, (kt + 1) * BK);
    const double* a_l = sA + cur * SA + (lane & 15);
    const double* b_l = a_l;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      const double al = a_l[kr * (BM + kPad) + 16 * gl], as_ = a_l[kr * (BM + kPad) + 16 * gs];
      double b[S];
#pragma unroll
      for (int c = 0; c < S; ++c) b[c] = b_l[kr * (BM + kPad) + 16 * c];
#pragma unroll
      for (int c = 0; c < S; ++c) {
        if (c <= gl) acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(al, b[c], acc[c], 0, 0, 0);
        if (c <= gs) acc[S - c] = __builtin_amdgcn_mfma_f64_16x16x4f64(as_, b[c], acc[S - c], 0, 0, 0);
      }
    }
    if (more) la.store(sA + (cur ^ 1) * SA);
    __syncthreads();
  } while (++kt < nk);
#pragma unroll
  for (int j = 0; j <= S; ++j) {
    const bool lng = j <= gl;
    prob.store_block(t, acc[j], lng ? gl : gs, lng ? j : S - j, lane);
  }
}

// The GEMM core. Accumulator element (fm, fn, r) of a lane is output element
//   m = m0 + wm + 16*fm + lane/16 + 4*r,  n = n0 + wn + 16*fn + lane%16
// (the v_mfma_f64_16x16x4f64 result layout).
template <int FM, int FN, int BK, class Prob>
__device__ __forceinline__ void gemm_tile_f64(const Prob& prob, const Tile& t, double* sA, double* sB) {
  constexpr int BM = 32 * FM, BN = 32 * FN;
  constexpr int SA = BK * (BM + kPad), SB = BK * (BN + kPad);
  const auto as = prob.a_src(t);
  const auto bs = prob.b_src(t);
  if constexpr (HasFoldDiag<Prob>::value && FM == 4 && FN == 4) {
    if (t.mt == t.nt) {
      gemm_diag_f64<FM, BK>(prob, t, sA);
      return;
    }
  }
  const int nk = tile_ksteps(prob, t, BK, 0);

  // (The wave index is left as a vector value here: making it scalar with
  // readfirstlane turns the MFMA block predicates into scalar branches, which
  // measured 6% slower on the half-transform and 11% on the exchange SYRK.)
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = (wave >> 1) * (BM / 2), wn = (wave & 1) * (BN / 2);

  Loader<Prob::kA, BM, BK> la;
  Loader<Prob::kB, BN, BK> lb;
  la.load(as, t.m0, 0);
  lb.load(bs, t.n0, 0);
  la.store(sA);
  lb.store(sB);
  __syncthreads();

  v4d acc[FM][FN];
  uint32_t active = 0;  // 16x16 blocks whose outputs are needed (wave-uniform)
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};
      if (prob.block_active(t, t.m0 + wm + 16 * i, t.n0 + wn + 16 * j)) active |= 1u << (i * FN + j);
    }

  // Every problem has at least one k-step (k extents are padded to >= 32). A
  // do-while gives the accumulators no zero-trip path, which would otherwise make
  // the compiler copy them out of AGPRs at the loop exit.
  int kt = 0;
  do {
    const int cur = kt & 1;
    const bool more = kt + 1 < nk;
    if (more) {
      la.load(as, t.m0, (kt + 1) * BK);
      lb.load(bs, t.n0, (kt + 1) * BK);
    }
    const double* a_l = sA + cur * SA + wm + (lane & 15);
    const double* b_l = sB + cur * SB + wn + (lane & 15);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4);
      double a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) a[i] = a_l[kr * (BM + kPad) + 16 * i];
#pragma unroll
      for (int j = 0; j < FN; ++j) b[j] = b_l[kr * (BN + kPad) + 16 * j];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          if (active & (1u << (i * FN + j)))
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) {
      la.store(sA + (cur ^ 1) * SA);
      lb.store(sB + (cur ^ 1) * SB);
    }
    __syncthreads();
  } while (++kt < nk);
  prob.template epilogue<FM, FN>(t, acc, wm, wn, lane, sA);
}

// One output tile per workgroup.
template <int FM, int FN, int BK, class Prob>
__global__ void __launch_bounds__(kThreads) k_gemm_f64(const Prob prob) {
  constexpr int BM = 32 * FM, BN = 32 * FN;
  __shared__ __align__(16) double sA[2 * BK * (BM + kPad)];
  __shared__ __align__(16) double sB[2 * BK * (BN + kPad)];
  Tile t;
  if (!prob.tile(xcd_remap(blockIdx.x, gridDim.x), BM, BN, t)) return;
  gemm_tile_f64<FM, FN, BK>(prob, t, sA, sB);
}

// Grouped form: one launch over the tiles of many instances of the same problem
// type, one workgroup per tile; tiles[i] = (slot of the instance's problem in
// probs, local tile id). Workgroup b runs on XCD x = b % 8 and takes entry b / 8 of
// that XCD's run [xcd_bounds[x], xcd_bounds[x + 1]) of the table; the grid covers
// the longest run and the surplus workgroups of shorter runs exit at once.
template <int FM, int FN, int BK, class Prob>
__global__ void __launch_bounds__(kThreads)
    k_gemm_f64_grouped(const Prob* __restrict__ probs, const int2* __restrict__ tiles,
                       const int* __restrict__ xcd_bounds) {
  constexpr int BM = 32 * FM, BN = 32 * FN;
  __shared__ __align__(16) double sA[2 * BK * (BM + kPad)];
  __shared__ __align__(16) double sB[2 * BK * (BN + kPad)];
  const int x = blockIdx.x % kXcds;
  const int id = xcd_bounds[x] + (int)(blockIdx.x / kXcds);
  if (id >= xcd_bounds[x + 1]) return;
  // The entry is uniform over the workgroup: readfirstlane lets the problem
  // struct come in through scalar loads.
  const int2 e = tiles[id];
  const int slot = __builtin_amdgcn_readfirstlane(e.x), local = __builtin_amdgcn_readfirstlane(e.y);
  const Prob prob = probs[slot];
  Tile t;
  if (prob.tile(local, BM, BN, t)) gemm_tile_f64<FM, FN, BK>(prob, t, sA, sB);
}

// ---------------------------------------------------------------- sources
struct SrcKMDense {  // element (k, r) at base[k*ld + r]
  const double* base;
  int64_t ld;
  __device__ __forceinline__ const double* at(int k, int r) const { return base + (int64_t)k * ld + r; }
};

struct SrcMKDense {  // element (r, k) at base[min(r, rmax-1)*ld + k]
  const double* base;
  int64_t ld;
  int rmax;
  __device__ __forceinline__ const double* row(int r) const {
    return base + (int64_t)(r < rmax ? r : rmax - 1) * ld;
  }
};

struct SrcKMRows {  // element (k, r) at base[min(k, kmax-1)*ld + r]
  const double* base;
  int64_t ld;
  int kmax;
  __device__ __forceinline__ const double* at(int k, int r) const {
    return base + (int64_t)(k < kmax ? k : kmax - 1) * ld + r;
  }
};

// b as a K x N operand of the half-transform: element (mu, n), n = nu*naux_p + P,
// at btri[tri(mu, nu)][P]. Rows past nbf are clamped (their C weights are zero).
struct SrcKMBtri {
  const double* btri;
  int nbf, naux_p;
  __device__ __forceinline__ const double* at(int k, int n) const {
    int nu = n / naux_p;
    const int p = n - nu * naux_p;
    nu = nu < nbf ? nu : nbf - 1;
    const int mu = k < nbf ? k : nbf - 1;
    return btri + tri_index(mu, nu) * naux_p + p;
  }
};

// ---------------------------------------------------------------- problems
template <int FM, int FN>
__device__ __forceinline__ void elem_coords(int lane, int wm, int wn, int i, int j, int r, int& dm, int& dn) {
  dm = wm + 16 * i + (lane >> 4) + 4 * r;
  dn = wn + 16 * j + (lane & 15);
}

struct ProbHT {
  static constexpr int kA = KM, kB = KM;
  const double* Cp;
  const double* btri;
  double* X;
  int nbf, naux_p, nbf5, ldcp, nbf_p, mtiles;
  int64_t N;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < nbf5; }
  __host__ __device__ int ksteps(int BK) const { return nbf_p / BK; }
  __device__ SrcKMDense a_src(const Tile&) const { return {Cp, ldcp}; }
  __device__ SrcKMBtri b_src(const Tile&) const { return {btri, nbf, naux_p}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int q = t.m0 + dm;
          const int64_t n = (int64_t)t.n0 + dn;
          if (q < nbf5 && n < N) gmem(X)[(int64_t)q * N + n] = acc[i][j][r];
        }
  }
};

// Orbital mu of output row slot m of the triangular half-transform: even slots take the
// rows from the end of the triangle, odd ones from its start (nbf - 1, 0, nbf - 2, 1, ...), so
// that any contiguous run of slots holds long and short rows alike.
__host__ __device__ __forceinline__ int tri_row_of_slot(int m, int nbf) {
  return (m & 1) ? (m - 1) / 2 : nbf - 1 - m / 2;
}

// b as the K x N operand of the triangular half-transform: element (nu, c), c = m naux_p + P
// over row slots m (tri_row_of_slot), is L[mu][nu][P] with L = b below the diagonal, b / 2 on
// it (dhalf, the diagonal rows halved) and 0 above (zeros, one zero row), so b = L + L^T.
struct SrcKMBtriL {
  const double* btri;
  const double* dhalf;
  const double* zeros;
  int nbf, naux_p;
  __device__ __forceinline__ const double* at(int k, int c) const {
    int m = c / naux_p;
    const int p = c - m * naux_p;
    m = m < nbf ? m : nbf - 1;
    const int mu = tri_row_of_slot(m, nbf);
    if (k < mu) return btri + tri_index(mu, k) * naux_p + p;
    if (k == mu) return dhalf + (int64_t)mu * naux_p + p;
    return zeros + p;
  }
};

// T[q][mu][P] = sum_{nu <= mu} C[nu][q] L[mu][nu][P] (L above): half the half-transform's
// work, since b = L + L^T gives Y_p[q] = W[p][q] + W[q][p] with W[p][q][P] = sum_mu C[mu][q]
// T[p][mu][P]. Columns run over row slots (c = m naux_p + P) and each tile's k-loop stops at
// the largest mu among its columns, so tiles carry from 1 to nbf / BK k-steps.
struct ProbHTL {
  static constexpr int kA = KM, kB = KM;
  const double* Cp;
  const double* btri;
  const double* dhalf;
  const double* zeros;
  double* T;
  int nbf, naux_p, nbf5, ldcp, nbf_p, mtiles, bn;  // bn: tile columns (<= naux_p)
  int64_t N;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < nbf5; }
  __host__ __device__ int ksteps(int BK) const { return nbf_p / BK; }
  __device__ int ksteps_tile(const Tile& t, int BK) const {
    // the tile's columns span row slots m_lo .. m_hi (two at most: bn <= naux_p)
    const int64_t c_hi = (int64_t)t.n0 + bn - 1;
    int m_lo = (int)(t.n0 / naux_p), m_hi = (int)((c_hi < N ? c_hi : N - 1) / naux_p);
    m_lo = m_lo < nbf ? m_lo : nbf - 1;
    m_hi = m_hi < nbf ? m_hi : nbf - 1;
    int mu = tri_row_of_slot(m_lo, nbf);
    for (int m = m_lo + 1; m <= m_hi; ++m) mu = max(mu, tri_row_of_slot(m, nbf));
    return (mu + BK) / BK;  // k = nu in [0, mu]
  }
  __device__ SrcKMDense a_src(const Tile&) const { return {Cp, ldcp}; }
  __device__ SrcKMBtriL b_src(const Tile&) const { return {btri, dhalf, zeros, nbf, naux_p}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int q = t.m0 + dm;
          const int64_t c = (int64_t)t.n0 + dn;
          if (q < nbf5 && c < N) {
            const int m = (int)(c / naux_p);
            const int64_t p = c - (int64_t)m * naux_p;
            gmem(T)[(int64_t)q * N + (int64_t)tri_row_of_slot(m, nbf) * naux_p + p] = acc[i][j][r];
          }
        }
  }
};

// Batch p: Y_p[q][P] = sum_nu C[nu][q] X[p][nu][P] for the cyclic window of rows
// q = (p + r) mod nbf5, r < h = qt_window(nbf5) (jk_common.h), so every batch has the same
// nbf5 / 2 + 1 rows. The A operand is the window copy of C: Cw0[nu][j] = C[nu][j mod nbf5]
// and Cw1[nu][j] = C[nu][(j + 1) mod nbf5] (rows nu >= nbf zero), so window row r of batch p
// is column p + r of Cw0, or column p - 1 + r of Cw1 for odd p: a contiguous column range at
// an even offset, as the pair loads need. Writes d[p][P] = Y_p[p][P] (window row 0) and the
// K_MO row-sum partials of the pairs the batch owns at (max(p, q), min(p, q)).
struct ProbQT {
  static constexpr int kA = KM, kB = KM;
  const double* Cw;  // Cw0, then Cw1 at Cw + nbf_p * ldw
  const double* X;
  double* kpart;
  double* d;
  int nbf, naux_p, nbf5, ldw, nbf_p, mtiles, ntiles, nt64, h;  // nt64 = naux_p / 64 (kpart stride)
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.mt = id % mtiles;
    const int rest = id / mtiles;
    t.nt = rest % ntiles;
    t.batch = rest / ntiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < h; }
  __host__ __device__ int ksteps(int BK) const { return nbf_p / BK; }
  __device__ SrcKMDense a_src(const Tile& t) const {
    const int p = t.batch;
    return {(p & 1) ? Cw + (int64_t)nbf_p * ldw + (p - 1) : Cw + p, ldw};
  }
  __device__ SrcKMRows b_src(const Tile& t) const {
    return {X + (int64_t)t.batch * nbf * naux_p, naux_p, nbf};
  }
  // kpart slot of window row w of batch p, or -1 when the row is past the window or its pair
  // belongs to the other batch.
  __device__ __forceinline__ int64_t kslot(int p, int w) const {
    if (w >= h || !qt_owns(p, w, nbf5)) return -1;
    const int q = qt_row(p, w, nbf5);
    const int hi = p > q ? p : q, lo = p > q ? q : p;
    return ((int64_t)hi * nbf5 + lo) * nt64;
  }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double* lds) const {
    const int p = t.batch;
    if (t.m0 == 0 && wm == 0 && lane < 16) {  // window row 0 is q = p: d[p][P] = Y_p[p][P]
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int n = t.n0 + wn + 16 * j + lane;
        if (n < naux_p) gmem(d)[(int64_t)p * naux_p + n] = acc[0][j][0];
      }
    }
    // Row sums of squares per 64 aux columns: a 64-column partial is always
    // S(first 32 columns) + S(last 32 columns), S being one wave's sum over two
    // 16-column blocks reduced across the 16 column lanes, so 64- and 128-column
    // tiles (FN = 2, 4) give the same bits (fixed order: deterministic).
    static_assert(FN == 2 || FN == 4, "quarter-transform tiles are 64 or 128 aux columns wide");
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
          for (int j = 2 * hh; j < 2 * hh + 2; ++j) v += acc[i][j][r] * acc[i][j][r];
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
            const int64_t k = kslot(p, t.m0 + wm + 16 * i + (lane >> 4) + 4 * r);
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
          for (int r = 0; r < 4; ++r) lds[wm + 16 * i + (lane >> 4) + 4 * r] = s[i][r];
      }
      __syncthreads();
      if (!(wave & 1) && (lane & 15) == 0) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int dm = wm + 16 * i + (lane >> 4) + 4 * r;
            const int64_t k = kslot(p, t.m0 + dm);
            if (k >= 0) gmem(kpart)[k + t.n0 / 64] = s[i][r] + lds[dm];
          }
      }
    }
  }
};

struct ProbSYRK {
  static constexpr int kA = MK, kB = MK;
  static constexpr bool kFoldDiag = true;  // diagonal 128 x 128 tiles: gemm_diag_f64
  const double* A;
  double* O;
  int64_t lda, sa, ldo, so;
  int M, K, tiles_side, tiles_per_batch;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = id / tiles_per_batch;
    const int rem = id - t.batch * tiles_per_batch;
    tri_pair(rem, t.mt, t.nt);
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  // blocks strictly above the diagonal of a diagonal tile are filled by the mirror writes
  __device__ bool block_active(const Tile& t, int r0, int c0) const {
    return r0 < M && c0 < M && !(t.mt == t.nt && r0 + 16 <= c0);
  }
  __host__ __device__ int ksteps(int BK) const { return K / BK; }
  __device__ SrcMKDense a_src(const Tile& t) const { return {A + (int64_t)t.batch * sa, lda, M}; }
  __device__ SrcMKDense b_src(const Tile& t) const { return {A + (int64_t)t.batch * sa, lda, M}; }
  // The 16 x 16 block of rows m0 + 16 g and columns n0 + 16 c of a diagonal tile.
  __device__ void store_block(const Tile& t, const v4d& v, int g, int c, int lane) const {
    double* o = O + (int64_t)t.batch * so;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = t.m0 + 16 * g + (lane >> 4) + 4 * r, n = t.n0 + 16 * c + (lane & 15);
      if (m < M && n <= m) {
        gmem(o)[(int64_t)m * ldo + n] = v[r];
        if (n < m) gmem(o)[(int64_t)n * ldo + m] = v[r];
      }
    }
  }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
    double* o = O + (int64_t)t.batch * so;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < M && n <= m) {  // (i, j) and its mirror (j, i): O exactly symmetric
            const double v = acc[i][j][r];
            gmem(o)[(int64_t)m * ldo + n] = v;
            if (n < m) gmem(o)[(int64_t)n * ldo + m] = v;
          }
        }
  }
};

struct ProbJN {
  static constexpr int kA = MK, kB = MK;
  const double* d;
  const double* btri;
  double* J;
  int nbf, naux_p, nbf5, mtiles;
  int64_t ntri;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < nbf5; }
  __host__ __device__ int ksteps(int BK) const { return naux_p / BK; }
  __device__ SrcMKDense a_src(const Tile&) const { return {d, naux_p, nbf5}; }
  __device__ SrcMKDense b_src(const Tile&) const { return {btri, naux_p, (int)(ntri < (1 << 30) ? ntri : (1 << 30))}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
    const int64_t nn = (int64_t)nbf * nbf;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      int dm, dn;
      elem_coords<FM, FN>(lane, wm, wn, 0, j, 0, dm, dn);
      const int64_t tt = (int64_t)t.n0 + dn;
      if (tt >= ntri) continue;
      const int mu = tri_row(tt);
      const int nu = (int)(tt - (int64_t)mu * (mu + 1) / 2);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int q = t.m0 + dm;
          if (q < nbf5) gmem(J)[(int64_t)q * nn + (int64_t)mu * nbf + nu] = acc[i][j][r];
        }
    }
  }
};

struct ProbGemmT {  // T[mu][i] = sum_nu Hp[mu][nu] Cp[nu][i]
  static constexpr int kA = MK, kB = KM;
  const double* Hp;
  const double* Cp;
  double* T;
  int ldh, ldcp, nbf, nbf_p, mtiles;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < nbf; }
  __host__ __device__ int ksteps(int BK) const { return nbf_p / BK; }
  __device__ SrcMKDense a_src(const Tile&) const { return {Hp, ldh, nbf}; }
  __device__ SrcKMDense b_src(const Tile&) const { return {Cp, ldcp}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < nbf && n < ldcp) gmem(T)[(int64_t)m * ldcp + n] = acc[i][j][r];
        }
  }
};

// Orbital-gradient MO integrals (integrals.py:703-716, Integrals_MO_RI_*): batch q,
//   b_MO[p][q][P] = sum_nu Cf[nu][p] X[q][nu][P]   for p < nbf, q < nbf5, P < naux,
// i.e. the quarter transform over every row p instead of q <= p. Cf holds all nbf
// columns of C (k-major, zero rows past nbf, zero columns past nbf).
struct ProbMO {
  static constexpr int kA = KM, kB = KM;
  const double* Cf;
  const double* X;
  double* out;
  int64_t ldrow;  // output row stride (nbf5 * naux for the reference layout)
  int nbf, naux, naux_p, nbf5, ldcf, nbf_p, mtiles, ntiles;
  int rows;  // output rows (columns of Cf): nbf for b_MO, nbf5 for the refinement's W
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.mt = id % mtiles;
    const int rest = id / mtiles;
    t.nt = rest % ntiles;
    t.batch = rest / ntiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < rows; }
  __host__ __device__ int ksteps(int BK) const { return nbf_p / BK; }
  __device__ SrcKMDense a_src(const Tile&) const { return {Cf, ldcf}; }
  __device__ SrcKMRows b_src(const Tile& t) const { return {X + (int64_t)t.batch * nbf * naux_p, naux_p, nbf}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < rows && n < naux) gmem(out)[(int64_t)m * ldrow + (int64_t)t.batch * naux + n] = acc[i][j][r];
        }
  }
};

// H_mat[i][j] = sum_mu Cf[mu][i] T[mu][j] = (C^T H C)[i][j] for i < nbf, j < nbf5
// (integrals.py:705), with T = H C[:, :nbf5] from ProbGemmT.
struct ProbHM {
  static constexpr int kA = KM, kB = KM;
  const double* Cf;
  const double* T;
  double* out;
  int nbf, nbf5, ldcf, ldcp, nbf_p, mtiles;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < nbf; }
  __host__ __device__ int ksteps(int BK) const { return nbf_p / BK; }
  __device__ SrcKMDense a_src(const Tile&) const { return {Cf, ldcf}; }
  __device__ SrcKMRows b_src(const Tile&) const { return {T, ldcp, nbf}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < nbf && n < nbf5) gmem(out)[(int64_t)m * nbf5 + n] = acc[i][j][r];
        }
  }
};

// Exchange term of the orbital-gradient Lagrangian (utils.py:353-355), split over
// k' = (q, P) in chunks of kc:
//   part[c][a][b] = sum_{k' in chunk c} B[a][k'] W[b][k'],   W[b][q naux + P] = ck[b][q] B[b][q][P],
// a < nbf, b < nbf5, with B = b_MO rows of stride ldrow (zero past nbf5 * naux).
struct ProbLagX {
  static constexpr int kA = MKLANE, kB = MKLANE;
  const double* B;
  const double* W;
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
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < nbf; }
  __host__ __device__ int ksteps(int BK) const { return (int)(kc / BK); }
  __device__ SrcMKDense a_src(const Tile& t) const { return {B + (int64_t)t.batch * kc, ldrow, nbf}; }
  __device__ SrcMKDense b_src(const Tile& t) const { return {W + (int64_t)t.batch * kc, ldrow, nbf5}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < nbf && n < nbf5)
            gmem(part)[((int64_t)t.batch * nbf + m) * nbf5 + n] = acc[i][j][r];
        }
  }
};

// Generalised Fock matrices, closed shell (utils.py:25-54), as one GEMM over the stacked
// J^n and K^n: F[i][mn] = sum_k A[i][k] JK[k][mn] + n_i H[mn], with A = [C^J' + diag(w n) | -C^K']
// (nbf5 x kp, zero past 2 nbf5) and JK rows J[0..nbf5) then K[0..nbf5) of length nn = nbf^2
// and stride ld (even; pairs at the row end are clamped inside the row).
struct SrcKMJK {
  const double* J;
  const double* K;
  int64_t ld;
  int n5;
  __device__ __forceinline__ const double* at(int k, int r) const {
    const int kk = k < 2 * n5 ? k : 2 * n5 - 1;
    const double* row = kk < n5 ? J + (int64_t)kk * ld : K + (int64_t)(kk - n5) * ld;
    return row + (r < ld - 1 ? (int64_t)r : ld - 2);
  }
};

struct ProbFock {
  static constexpr int kA = MK, kB = KM;
  const double* A;
  const double* J;
  const double* K;
  const double* H;   // nbf x nbf, row stride nbf
  const double* hc;  // coefficient of H in row i
  double* F;
  int64_t ld, nn;
  int nbf, nbf5, kp, mtiles;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < nbf5; }
  __host__ __device__ int ksteps(int BK) const { return kp / BK; }
  __device__ SrcMKDense a_src(const Tile&) const { return {A, kp, nbf5}; }
  __device__ SrcKMJK b_src(const Tile&) const { return {J, K, ld, nbf5}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm;
          const int64_t n = (int64_t)t.n0 + dn;
          if (m < nbf5 && n < nn) gmem(F)[(int64_t)m * nn + n] = acc[i][j][r] + hc[m] * H[n];
        }
  }
};

// out[t][n] = sum_k A[t][k] W[k][n] for rows t < rows, n < K: the packed b from the packed
// 3-centre integrals (mu nu|P) and the metric's inverse square root (integrals.py:49). A has
// row stride K; W is K x K (row stride K) with at least 128 doubles of slack at its end.
struct ProbRowGemm {
  static constexpr int kA = MK, kB = KM;
  const double* A;
  const double* W;
  double* out;
  int rows, K, mtiles;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < rows; }
  __host__ __device__ int ksteps(int BK) const { return K / BK; }
  __device__ SrcMKDense a_src(const Tile&) const { return {A, K, rows}; }
  __device__ SrcKMDense b_src(const Tile&) const { return {W, K}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < rows && n < K) gmem(out)[(int64_t)m * K + n] = acc[i][j][r];
        }
  }
};

// out[m][n] = sum_k A[k][m] B[k][n], all K x K with row stride K (K a multiple of 64):
// A^T B, for V diag(f) V^T from the eigenvectors.
struct ProbAtB {
  static constexpr int kA = KM, kB = KM;
  const double* A;
  const double* B;
  double* out;
  int K, mtiles;
  __host__ __device__ bool tile(int id, int BM, int BN, Tile& t) const {
    t.batch = 0;
    t.mt = id % mtiles;
    t.nt = id / mtiles;
    t.m0 = t.mt * BM;
    t.n0 = t.nt * BN;
    return true;
  }
  __device__ bool block_active(const Tile&, int r0, int) const { return r0 < K; }
  __host__ __device__ int ksteps(int BK) const { return K / BK; }
  __device__ SrcKMDense a_src(const Tile&) const { return {A, K}; }
  __device__ SrcKMDense b_src(const Tile&) const { return {B, K}; }
  template <int FM, int FN>
  __device__ void epilogue(const Tile& t, v4d (&acc)[FM][FN], int wm, int wn, int lane, double*) const {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int dm, dn;
          elem_coords<FM, FN>(lane, wm, wn, i, j, r, dm, dn);
          const int m = t.m0 + dm, n = t.n0 + dn;
          if (m < K && n < K) gmem(out)[(int64_t)m * K + n] = acc[i][j][r];
        }
  }
};

// ---------------------------------------------------------------- small kernels
__global__ void k_kmo_reduce(int nbf5, int nt, const double* __restrict__ kpart, double* __restrict__ kmo) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nbf5 * nbf5) return;
  const int p = (int)(idx / nbf5), q = (int)(idx % nbf5);
  if (q > p) return;
  const double* src = kpart + ((int64_t)p * nbf5 + q) * nt;
  double s = 0.0;
  for (int t = 0; t < nt; ++t) s += src[t];
  kmo[(int64_t)p * nbf5 + q] = s;
  kmo[(int64_t)q * nbf5 + p] = s;
}

// H_core[i] = sum_mu Cp[mu][i] T[mu][i]: one wave per orbital i; lane l sums mu = l + 64 j
// in order, then an xor-shuffle reduction (fixed order).
__global__ void __launch_bounds__(256) k_hcore(int nbf, int nbf5, int ldcp, const double* __restrict__ Cp,
                                               const double* __restrict__ T, double* __restrict__ hcore) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= nbf5) return;
  double s = 0.0;
  for (int mu = lane; mu < nbf; mu += 64) s += Cp[(int64_t)mu * ldcp + i] * T[(int64_t)mu * ldcp + i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) hcore[i] = s;
}

__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
  atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}

// Packs rows mu in [mu0, mu1) of a full b (bc = b[mu0:mu1, :, :]) into btri.
// One workgroup per (mu, nu) pair, looping over P. Pairs nu > mu are only
// checked against their mirror: inside the chunk when nu < mu1, otherwise
// against the packed row tri(nu, mu), which a chunk of higher mu has already
// written (chunks are packed in descending mu order on one stream).
__global__ void k_pack_b(int nbf, int naux, int naux_p, int mu0, int mu1, const double* __restrict__ bc,
                         double* __restrict__ btri, double* asym_max, unsigned long long* nonfinite) {
  const int64_t row = blockIdx.x;  // (mu - mu0) * nbf + nu
  const int mu = mu0 + (int)(row / nbf), nu = (int)(row % nbf);
  const double* src = bc + row * naux;
  double amax = 0.0;
  unsigned long long bad = 0;
  if (nu > mu) {
    const double* mir = nu < mu1 ? bc + ((int64_t)(nu - mu0) * nbf + mu) * naux : btri + tri_index(nu, mu) * naux_p;
    for (int p = threadIdx.x; p < naux; p += blockDim.x) {
      const double v = src[p];
      if (!isfinite(v)) ++bad;
      const double a = fabs(v - mir[p]);
      amax = a > amax ? a : amax;
    }
  } else {
    double* dst = btri + tri_index(mu, nu) * naux_p;
    for (int p = threadIdx.x; p < naux_p; p += blockDim.x) {
      double v = 0.0;
      if (p < naux) {
        v = src[p];
        if (!isfinite(v)) ++bad;
      }
      dst[p] = v;
    }
  }
  if (bad) atomicAdd(nonfinite, bad);
  if (amax > 0.0) atomic_max_nonneg(asym_max, amax);
}

// Fills the strict upper triangle of each n x n matrix from the lower one:
// O[b][j][i] = O[b][i][j] for i > j. One 32 x 32 lower tile pair per workgroup,
// read and written row-contiguously through LDS.
__global__ void k_symmetrize(double* __restrict__ O, int n, int64_t ld, int64_t stride) {
  __shared__ double tile[32][33];
  const int64_t pair = blockIdx.x;
  const int ti = tri_row(pair);
  const int tj = (int)(pair - (int64_t)ti * (ti + 1) / 2);
  double* o = O + (int64_t)blockIdx.y * stride;
  for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
    const int r = e >> 5, c = e & 31;
    const int i = ti * 32 + r, j = tj * 32 + c;
    tile[r][c] = (i < n && j < n) ? o[(int64_t)i * ld + j] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
    const int r = e >> 5, c = e & 31;
    const int j = tj * 32 + r, i = ti * 32 + c;
    if (i < n && j < n && i > j) o[(int64_t)j * ld + i] = tile[c][r];
  }
}

// k_symmetrize over a table of jobs (one batch of n x n matrices each): workgroup b
// finds its job by binary search on the jobs' first workgroups.
__global__ void k_symmetrize_grouped(const SymJob* __restrict__ jobs, int njobs) {
  __shared__ double tile[32][33];
  const int64_t b = blockIdx.x;
  int lo = 0, hi = njobs - 1;  // last job with first <= b
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].first <= b) lo = mid;
    else hi = mid - 1;
  }
  const SymJob jb = jobs[lo];
  const int64_t pairs = sym_blocks(jb.n);
  const int64_t local = b - jb.first;
  const int64_t m = local / pairs, pair = local - m * pairs;
  if (m >= jb.count) return;
  const int ti = tri_row(pair);
  const int tj = (int)(pair - (int64_t)ti * (ti + 1) / 2);
  double* o = jb.O + m * jb.stride;
  const int n = jb.n;
  for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
    const int r = e >> 5, c = e & 31;
    const int i = ti * 32 + r, j = tj * 32 + c;
    tile[r][c] = (i < n && j < n) ? o[(int64_t)i * jb.ld + j] : 0.0;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
    const int r = e >> 5, c = e & 31;
    const int j = tj * 32 + r, i = ti * 32 + c;
    if (i < n && j < n && i > j) o[(int64_t)j * jb.ld + i] = tile[c][r];
  }
}

template <int FM, int FN, int BK, class Prob>
hipError_t grouped_launch(const void* probs, const int2* tiles, const int* xcd_bounds, int grid, hipStream_t s) {
  hipLaunchKernelGGL((k_gemm_f64_grouped<FM, FN, BK, Prob>), dim3(grid), dim3(kThreads), 0, s,
                     static_cast<const Prob*>(probs), tiles, xcd_bounds);
  return hipGetLastError();
}

// Either launches `prob` over nblocks tiles now, or (col != null) appends it
// and its valid tiles to the grouped launch for its problem type and tile.
template <class Prob>
hipError_t launch_prob(const Prob& prob, TileF64 tc, int nblocks, hipStream_t s, GroupCollector* col = nullptr) {
  if (nblocks <= 0) return hipSuccess;
  const dim3 grid(nblocks), block(kThreads);
#define PNJK_CASE(FMV, FNV, BKV)                                                                            \
  if (tc.fm == FMV && tc.fn == FNV) {                                                                       \
    if (col) {                                                                                              \
      col->add(prob, tc, nblocks, 32 * FMV, 32 * FNV, prob.ksteps(BKV), &grouped_launch<FMV, FNV, BKV, Prob>, \
               __PRETTY_FUNCTION__);                                                                        \
      return hipSuccess;                                                                                    \
    }                                                                                                       \
    hipLaunchKernelGGL((k_gemm_f64<FMV, FNV, BKV, Prob>), grid, block, 0, s, prob);                         \
    return hipGetLastError();                                                                               \
  }
  PNJK_CASE(2, 2, 16)
  PNJK_CASE(4, 4, 8)
  PNJK_CASE(5, 4, 8)
  PNJK_CASE(4, 2, 16)
  PNJK_CASE(2, 4, 16)
  PNJK_CASE(5, 2, 16)
#undef PNJK_CASE
  return hipErrorInvalidValue;
}

inline int cdiv(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

}  // namespace

hipError_t launch_ht_f64(const BuildGeom& g, TileF64 tc, const double* Cp, const double* btri, double* X,
                         hipStream_t s, GroupCollector* col) {
  ProbHT pr;
  pr.Cp = Cp;
  pr.btri = btri;
  pr.X = X;
  pr.nbf = g.nbf;
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ldcp = g.ldcp;
  pr.nbf_p = g.nbf_p;
  pr.N = (int64_t)g.nbf * g.naux_p;
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  pr.mtiles = cdiv(g.nbf5, BM);
  return launch_prob(pr, tc, pr.mtiles * cdiv(pr.N, BN), s, col);
}


hipError_t launch_qt_f64(const BuildGeom& g, TileF64 tc, const double* Cw, int ldw, const double* X, double* kpart,
                         double* d, int* ntiles, hipStream_t s, GroupCollector* col, int window) {
  // naux_p is a multiple of 64; a 128-column tile may cover 64 columns past it,
  // which the B loader reads from the workspace slack of X and the epilogue drops.
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  if (tc.fn != 2 && tc.fn != 4) return hipErrorInvalidValue;
  ProbQT pr;
  pr.Cw = Cw;
  pr.X = X;
  pr.kpart = kpart;
  pr.d = d;
  pr.nbf = g.nbf;
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ldw = ldw;
  pr.nbf_p = g.nbf_p;
  if (window < 0 || window > qt_window(g.nbf5)) return hipErrorInvalidValue;
  pr.h = window ? window : qt_window(g.nbf5);
  pr.mtiles = cdiv(pr.h, BM);
  // The window copy must cover every column a tile reads, (p or p - 1) + mtiles * BM - 1 <=
  // nbf5 - 2 + mtiles * BM, at an even row stride.
  if ((int64_t)pr.mtiles * BM + g.nbf5 - 1 > ldw || (ldw & 1)) return hipErrorInvalidValue;
  pr.ntiles = cdiv(g.naux_p, BN);
  pr.nt64 = g.naux_p / 64;
  *ntiles = pr.nt64;
  return launch_prob(pr, tc, pr.mtiles * pr.ntiles * g.nbf5, s, col);
}

// Cw0[nu][j] = Cp[nu][j mod n], Cw1[nu][j] = Cp[nu][(j + 1) mod n] for nu < rows (Cp rows
// past nbf are zero).
__global__ void k_window_c(const double* __restrict__ Cp, int ldcp, int rows, int n, int ldw, double* __restrict__ Cw) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)rows * ldw;
  if (idx >= 2 * per) return;
  const int which = (int)(idx / per);
  const int64_t e = idx - which * per;
  const int nu = (int)(e / ldw), j = (int)(e % ldw);
  gmem(Cw)[idx] = Cp[(int64_t)nu * ldcp + (j + which) % n];
}

hipError_t launch_window_c(const double* Cp, int ldcp, int rows, int n, int ldw, double* Cw, hipStream_t s) {
  const int64_t total = 2 * (int64_t)rows * ldw;
  hipLaunchKernelGGL(k_window_c, dim3((unsigned)cdiv(total, 256)), dim3(256), 0, s, Cp, ldcp, rows, n, ldw, Cw);
  return hipGetLastError();
}

hipError_t launch_symmetrize(int batch, int n, double* O, int64_t ld, int64_t stride, hipStream_t s) {
  if (batch <= 0 || n <= 1) return hipSuccess;
  const int64_t t = (n + 31) / 32;
  const int64_t pairs = t * (t + 1) / 2;
  if (batch > 65535 || pairs > 2147483647LL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_symmetrize, dim3((unsigned)pairs, (unsigned)batch), dim3(256), 0, s, O, n, ld, stride);
  return hipGetLastError();
}

hipError_t launch_symmetrize_grouped(const SymJob* jobs, int njobs, int64_t blocks, hipStream_t s) {
  if (njobs <= 0 || blocks <= 0) return hipSuccess;
  if (blocks > 2147483647LL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_symmetrize_grouped, dim3((unsigned)blocks), dim3(256), 0, s, jobs, njobs);
  return hipGetLastError();
}

hipError_t launch_kmo_reduce(int nbf5, int nt, const double* kpart, double* kmo, hipStream_t s) {
  const int64_t n = (int64_t)nbf5 * nbf5;
  hipLaunchKernelGGL(k_kmo_reduce, dim3(cdiv(n, 256)), dim3(256), 0, s, nbf5, nt, kpart, kmo);
  return hipGetLastError();
}

hipError_t launch_syrk_f64(int batch, int M, int K, const double* A, int64_t lda, int64_t sa, double* O,
                           int64_t ldo, int64_t so, TileF64 tc, hipStream_t s, GroupCollector* col) {
  if (tc.fm != tc.fn) return hipErrorInvalidValue;
  const int BM = 32 * tc.fm;
  const int BK = (tc.fm >= 4 && tc.fn >= 4) ? 8 : 16;
  if (K % BK != 0) return hipErrorInvalidValue;
  ProbSYRK pr;
  pr.A = A;
  pr.O = O;
  pr.lda = lda;
  pr.sa = sa;
  pr.ldo = ldo;
  pr.so = so;
  pr.M = M;
  pr.K = K;
  pr.tiles_side = cdiv(M, BM);
  pr.tiles_per_batch = pr.tiles_side * (pr.tiles_side + 1) / 2;
  return launch_prob(pr, tc, pr.tiles_per_batch * batch, s, col);
}

hipError_t launch_jn_f64(const BuildGeom& g, TileF64 tc, const double* d, const double* btri, double* J,
                         hipStream_t s, GroupCollector* col) {
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  ProbJN pr;
  pr.d = d;
  pr.btri = btri;
  pr.J = J;
  pr.nbf = g.nbf;
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ntri = g.ntri;
  pr.mtiles = cdiv(g.nbf5, BM);
  return launch_prob(pr, tc, pr.mtiles * cdiv(g.ntri, BN), s, col);
}

// Lagrangian helpers (utils.py:262-355).
// Copies an nbf5 x nbf5 coefficient matrix with its diagonal set to zero (:269-270).
__global__ void k_lag_coef(const double* __restrict__ src, double* __restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * n) return;
  dst[i] = (i / n == i % n) ? 0.0 : src[i];
}

// G[b][P] = w_b n_b d[b][P] + sum_q cj[b][q] d[q][P], d[q][P] = B[q][q][P], w_b = 1 for
// b < nbeta or b >= nalpha (:337-350).
__global__ void k_lag_g(const double* __restrict__ B, int64_t ldrow, int naux, int nbf5, int nbeta, int nalpha,
                        const double* __restrict__ nocc, const double* __restrict__ cj, double* __restrict__ G) {
  const int P = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (P >= naux) return;
  double v = 0.0;
  for (int q = 0; q < nbf5; ++q) v += cj[(int64_t)b * nbf5 + q] * B[(int64_t)q * ldrow + (int64_t)q * naux + P];
  if (b < nbeta || b >= nalpha) v += nocc[b] * B[(int64_t)b * ldrow + (int64_t)b * naux + P];
  G[(int64_t)b * naux + P] = v;
}

// W[b][k'] = ck[b][q] B[b][k'] for k' = q naux + P < nbf5 naux, zero up to ldrow (:353-355).
__global__ void k_lag_w(const double* __restrict__ B, int64_t ldrow, int naux, int nbf5,
                        const double* __restrict__ ck, double* __restrict__ W) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nbf5 * ldrow) return;
  const int b = (int)(i / ldrow);
  const int64_t kp = i - (int64_t)b * ldrow;
  const int q = (int)(kp / naux);
  W[i] = q < nbf5 ? ck[(int64_t)b * nbf5 + q] * B[i] : 0.0;
}

// Zeroes columns [c0, c1) of `rows` rows of stride ld.
__global__ void k_fill_cols(double* __restrict__ p, int rows, int64_t ld, int64_t c0, int64_t c1) {
  const int64_t w = c1 - c0;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * w) return;
  p[(i / w) * ld + c0 + i % w] = 0.0;
}

// elag[a][b] = n_b Hm[a][b] + sum_P B[a][b][P] G[b][P] - sum_c part[c][a][b] for b < nbf5,
// zero for b >= nbf5. One wave per (a, b); fixed summation order (deterministic).
__global__ void __launch_bounds__(256) k_lag_final(const double* __restrict__ B, int64_t ldrow, int naux, int nbf,
                                                   int nbf5, const double* __restrict__ G,
                                                   const double* __restrict__ Hm, const double* __restrict__ nocc,
                                                   const double* __restrict__ part, int nchunks,
                                                   double* __restrict__ elag) {
  const int lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= (int64_t)nbf * nbf) return;
  const int a = (int)(w / nbf), b = (int)(w % nbf);
  if (b >= nbf5) {
    if (lane == 0) elag[w] = 0.0;
    return;
  }
  const double* row = B + (int64_t)a * ldrow + (int64_t)b * naux;
  const double* g = G + (int64_t)b * naux;
  double v = 0.0, x = 0.0;
  for (int P = lane; P < naux; P += 64) v += row[P] * g[P];
  for (int c = lane; c < nchunks; c += 64) x += part[((int64_t)c * nbf + a) * nbf5 + b];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v += __shfl_xor(v, off, 64);
    x += __shfl_xor(x, off, 64);
  }
  if (lane == 0) elag[w] = nocc[b] * Hm[(int64_t)a * nbf5 + b] + v - x;
}

hipError_t launch_mo_f64(const BuildGeom& g, TileF64 tc, const double* Cf, int ldcf, const double* X, double* b_mo,
                         int64_t ldrow, hipStream_t s) {
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  if (g.naux_p % BN != 0 || ldcf < cdiv(g.nbf, BM) * BM || ldrow < (int64_t)g.nbf5 * g.naux)
    return hipErrorInvalidValue;
  ProbMO pr;
  pr.Cf = Cf;
  pr.X = X;
  pr.out = b_mo;
  pr.ldrow = ldrow;
  pr.nbf = g.nbf;
  pr.naux = g.naux;
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ldcf = ldcf;
  pr.nbf_p = g.nbf_p;
  pr.rows = g.nbf;
  pr.mtiles = cdiv(g.nbf, BM);
  pr.ntiles = g.naux_p / BN;
  return launch_prob(pr, tc, pr.mtiles * pr.ntiles * g.nbf5, s);
}

hipError_t launch_htl_f64(const BuildGeom& g, TileF64 tc, const double* Cp, const double* btri, const double* dhalf,
                          const double* zeros, double* T, hipStream_t s) {
  if (32 * tc.fn > g.naux_p) return hipErrorInvalidValue;  // a tile spans at most two row slots
  ProbHTL pr;
  pr.Cp = Cp;
  pr.btri = btri;
  pr.dhalf = dhalf;
  pr.zeros = zeros;
  pr.T = T;
  pr.nbf = g.nbf;
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ldcp = g.ldcp;
  pr.nbf_p = g.nbf_p;
  pr.N = (int64_t)g.nbf * g.naux_p;
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  pr.bn = BN;
  pr.mtiles = cdiv(g.nbf5, BM);
  return launch_prob(pr, tc, pr.mtiles * cdiv(pr.N, BN), s);
}

hipError_t launch_w_f64(const BuildGeom& g, TileF64 tc, const double* Cp, const double* T, double* W,
                        hipStream_t s) {
  const int BM = 32 * tc.fm, BN = 32 * tc.fn;
  // A tile reads columns up to mtiles * BM of each C row: past ldcp they run into the next
  // row (Cp carries slack rows), and only output rows below nbf5 are written.
  if (g.naux_p % BN != 0) return hipErrorInvalidValue;
  ProbMO pr;
  pr.Cf = Cp;
  pr.X = T;
  pr.out = W;
  pr.ldrow = (int64_t)g.nbf5 * g.naux_p;
  pr.nbf = g.nbf;
  pr.naux = g.naux_p;  // W[q][p][P] over the padded aux range (T is zero there)
  pr.naux_p = g.naux_p;
  pr.nbf5 = g.nbf5;
  pr.ldcf = g.ldcp;
  pr.nbf_p = g.nbf_p;
  pr.rows = g.nbf5;
  pr.mtiles = cdiv(g.nbf5, BM);
  pr.ntiles = g.naux_p / BN;
  return launch_prob(pr, tc, pr.mtiles * pr.ntiles * g.nbf5, s);
}

hipError_t launch_hmat_f64(const BuildGeom& g, const double* Cf, int ldcf, const double* T, double* hmat,
                           hipStream_t s) {
  const TileF64 tc{2, 2};  // 64 x 64 tiles: BN spans at most ldcp + 32 columns of T (slack)
  if (ldcf < cdiv(g.nbf, 64) * 64) return hipErrorInvalidValue;
  ProbHM pr;
  pr.Cf = Cf;
  pr.T = T;
  pr.out = hmat;
  pr.nbf = g.nbf;
  pr.nbf5 = g.nbf5;
  pr.ldcf = ldcf;
  pr.ldcp = g.ldcp;
  pr.nbf_p = g.nbf_p;
  pr.mtiles = cdiv(g.nbf, 64);
  return launch_prob(pr, tc, pr.mtiles * cdiv(g.nbf5, 64), s);
}

hipError_t launch_lag_f64(const BuildGeom& g, const double* B, int64_t ldrow, int64_t kc, int nchunks,
                          const double* nocc, const double* cj, const double* ck, int nbeta, int nalpha,
                          const double* Hm, double* G, double* W, double* part, double* elag, bool exchange,
                          hipStream_t s) {
  const int nbf = g.nbf, n5 = g.nbf5, naux = g.naux;
  if (kc % 16 != 0 || (int64_t)nchunks * kc != ldrow || ldrow < (int64_t)n5 * naux) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_lag_g, dim3(cdiv(naux, 256), n5), dim3(256), 0, s, B, ldrow, naux, n5, nbeta, nalpha, nocc, cj,
                     G);
  if (exchange) {  // FP64 exchange partials (otherwise the caller formed them)
    const int64_t nw = (int64_t)n5 * ldrow;
    hipLaunchKernelGGL(k_lag_w, dim3((unsigned)cdiv(nw, 256)), dim3(256), 0, s, B, ldrow, naux, n5, ck, W);
    const TileF64 tc{4, 2};
    ProbLagX pr;
    pr.B = B;
    pr.W = W;
    pr.part = part;
    pr.ldrow = ldrow;
    pr.kc = kc;
    pr.nbf = nbf;
    pr.nbf5 = n5;
    pr.mtiles = cdiv(nbf, 32 * tc.fm);
    pr.ntiles = cdiv(n5, 32 * tc.fn);
    hipError_t e = launch_prob(pr, tc, pr.mtiles * pr.ntiles * nchunks, s);
    if (e != hipSuccess) return e;
  }
  const int64_t waves = (int64_t)nbf * nbf;
  hipLaunchKernelGGL(k_lag_final, dim3((unsigned)cdiv(waves, 4)), dim3(256), 0, s, B, ldrow, naux, nbf, n5, G, Hm, nocc,
                     part, nchunks, elag);
  return hipGetLastError();
}

hipError_t launch_lag_coef(const double* src, double* dst, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_lag_coef, dim3(cdiv(n * n, 256)), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t launch_fill_cols(double* p, int rows, int64_t ld, int64_t c0, int64_t c1, hipStream_t s) {
  if (c1 <= c0 || rows <= 0) return hipSuccess;
  const int64_t n = (int64_t)rows * (c1 - c0);
  hipLaunchKernelGGL(k_fill_cols, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, p, rows, ld, c0, c1);
  return hipGetLastError();
}

// Refinement MO integrals from W (row (x, y) of naux_p values holds W_xy = (C^T L C)_xy):
// Y_pq = W_pq + W_qp. One workgroup per pair p >= q (pair = p (p + 1) / 2 + q):
// K_MO[p][q] = K_MO[q][p] = sum_{P < naux} Y_pq[P]^2, each thread summing P = tid + 256 j in
// increasing j, then a fixed tree over the workgroup; on the diagonal also
// d[p][P] = Y_pp[P] = 2 W_pp[P] for P < naux_p (zero past naux, as W is). With jmo, also
// J_MO[p][q] = J_MO[q][p] = sum_P d[p][P] d[q][P] in the same order (the rows W_pp, W_qq are
// re-read from cache), so that diag(J_MO) = diag(K_MO) bit for bit.
template <class TW>
__global__ void __launch_bounds__(256) k_ymo(const TW* __restrict__ W, int n5, int naux, int naux_p,
                                             double* __restrict__ kmo, double* __restrict__ d,
                                             double* __restrict__ jmo) {
  __shared__ double red[256], redj[256];
  const int pair = blockIdx.x;
  int p = (int)((sqrt(8.0 * pair + 1.0) - 1.0) * 0.5);
  while ((int64_t)(p + 1) * (p + 2) / 2 <= pair) ++p;
  while ((int64_t)p * (p + 1) / 2 > pair) --p;
  const int q = pair - p * (p + 1) / 2;
  const TW* a = W + ((int64_t)p * n5 + q) * naux_p;
  const TW* b = W + ((int64_t)q * n5 + p) * naux_p;
  double acc = 0.0;
  for (int P = threadIdx.x; P < naux; P += 256) {
    const double y = (double)a[P] + (double)b[P];
    acc = fma(y, y, acc);
  }
  red[threadIdx.x] = acc;
  if (jmo) {  // uniform over the grid
    const TW* wp = W + ((int64_t)p * n5 + p) * naux_p;
    const TW* wq = W + ((int64_t)q * n5 + q) * naux_p;
    double accj = 0.0;
    for (int P = threadIdx.x; P < naux; P += 256) {
      const double dp = (double)wp[P] + (double)wp[P], dq = (double)wq[P] + (double)wq[P];
      accj = fma(dp, dq, accj);
    }
    redj[threadIdx.x] = accj;
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
      if (jmo) redj[threadIdx.x] += redj[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    kmo[(int64_t)p * n5 + q] = red[0];
    kmo[(int64_t)q * n5 + p] = red[0];
    if (jmo) {
      jmo[(int64_t)p * n5 + q] = redj[0];
      jmo[(int64_t)q * n5 + p] = redj[0];
    }
  }
  if (p == q)
    for (int P = threadIdx.x; P < naux_p; P += 256) d[(int64_t)p * naux_p + P] = 2.0 * (double)a[P];
}

// dhalf[mu][P] = b[mu][mu][P] / 2 (the diagonal rows of the packed b, halved) and one zero
// row of naux_p values: the diagonal and upper parts of L in the triangular half-transform.
__global__ void k_half_diag(const double* __restrict__ btri, int nbf, int naux_p, double* __restrict__ dhalf,
                            double* __restrict__ zeros) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)nbf * naux_p;
  if (i < n) {
    const int64_t mu = i / naux_p, p = i - mu * naux_p;
    dhalf[i] = 0.5 * btri[tri_index(mu, mu) * naux_p + p];
  } else if (i < n + naux_p) {
    zeros[i - n] = 0.0;
  }
}

// Row i of the Fock GEMM: F[i] = hc[i] H + sum_j A[i][j] J[j] + sum_j A[i][nbf5 + j] K[j] (A zero
// from 2 nbf5 up to kp). D = [0, nbeta), S = [nbeta, nalpha), W = [nalpha, nbf5); C^J', C^K' are
// the coefficients with a zero diagonal for i >= ini (ini = no1 if no1 > 1 else 0).
// Closed shell (utils.py:25-54): hc = n_i; A[i][j] = C^J'[i][j] + n_i [j == i and i in [ini, nbeta)
//   or W]; A[i][nbf5 + j] = -C^K'[i][j].
// Open shell (utils.py:96-171): rows in D and W as above but with C^J', C^K' over columns in D and W
//   only (:123-152), plus n_i J[s] - n_i K[s] / 2 for every s in S (SUMij, :155-161); rows in S:
//   hc = 1/2 (:104), (J[s] - K[s]) / 2 for s in S other than i (SUMij less its diagonal,
//   :158-159) and n_j J[j] - n_j K[j] / 2 for j in D and W (PRODWROij, :164-169).
__global__ void k_fock_coef(const double* __restrict__ cj, const double* __restrict__ ck,
                            const double* __restrict__ nocc, int n5, int ini, int nbeta, int nalpha, int kp,
                            int open, double* __restrict__ A, double* __restrict__ hc) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n5 * kp) return;
  const int i = idx / kp, k = idx % kp;
  const bool iS = i >= nbeta && i < nalpha;
  if (k == 0) hc[i] = (open && iS) ? 0.5 : nocc[i];
  double v = 0.0;
  if (k < 2 * n5) {
    const bool kj = k < n5;  // J part, else K part
    const int j = kj ? k : k - n5;
    const bool jS = j >= nbeta && j < nalpha;
    if (!open || (!iS && !jS)) {
      if (!(j == i && i >= ini)) v = kj ? cj[(int64_t)i * n5 + j] : -ck[(int64_t)i * n5 + j];
    }
    if (kj && j == i && ((i >= ini && i < nbeta) || i >= nalpha)) v += nocc[i];
    if (open) {
      if (!iS && jS) v += kj ? nocc[i] : -0.5 * nocc[i];
      if (iS && jS && j != i) v += kj ? 0.5 : -0.5;
      if (iS && !jS) v += kj ? nocc[j] : -0.5 * nocc[j];
    }
  }
  A[idx] = v;
}

// G[m][i] = sum_n F[i][m][n] C[n][i] (utils.py:251), i < nbf5, row stride ldg. Block
// (x, y) takes orbital i = y and rows m of 32 x .. 32 x + 31, eight per wave; column i of C
// (Cf, row stride ldcf) is staged through LDS in chunks. Each lane sums n = lane + 64 j in
// increasing order, then the wave reduces by xor-shuffle: the order is fixed.
__global__ void __launch_bounds__(256) k_fock_g(const double* __restrict__ F, const double* __restrict__ Cf,
                                                int ldcf, int nbf, int ldg, double* __restrict__ G) {
  constexpr int kRows = 8, kChunk = 2048;
  __shared__ double c[kChunk];
  const int i = blockIdx.y, lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int mw = blockIdx.x * 32 + wave * kRows;
  double acc[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) acc[r] = 0.0;
  for (int n0 = 0; n0 < nbf; n0 += kChunk) {
    const int len = nbf - n0 < kChunk ? nbf - n0 : kChunk;
    __syncthreads();
    for (int t = threadIdx.x; t < len; t += 256) c[t] = Cf[(int64_t)(n0 + t) * ldcf + i];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int m = mw + r;
      if (m < nbf) {
        const double* row = F + ((int64_t)i * nbf + m) * nbf + n0;
        double v = acc[r];
        for (int n = lane; n < len; n += 64) v += row[n] * c[n];
        acc[r] = v;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    double v = acc[r];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int m = mw + r;
    if (lane == 0 && m < nbf) G[(int64_t)m * ldg + i] = v;
  }
}

// elag[i][j] = src[i][j] (nbf x nbf5) for i < noptorb, j < nbf5; zero elsewhere (utils.py:254-257).
__global__ void k_elag_scatter(const double* __restrict__ src, int nbf, int nbf5, int noptorb,
                               double* __restrict__ elag) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)nbf * nbf) return;
  const int i = (int)(idx / nbf), j = (int)(idx % nbf);
  elag[idx] = (i < noptorb && j < nbf5) ? src[(int64_t)i * nbf5 + j] : 0.0;
}

hipError_t launch_fock_f64(int nbf, int n5, const double* J, const double* K, int64_t ld, const double* Hd,
                           const double* nocc, const double* cj, const double* ck, int ini, int nbeta, int nalpha,
                           bool open_shell, double* A, double* hc, double* F, hipStream_t s) {
  const int64_t nn = (int64_t)nbf * nbf;
  if (ld % 2 != 0 || ld < nn || ld < 2) return hipErrorInvalidValue;
  const int kp = (2 * n5 + 15) / 16 * 16;
  hipLaunchKernelGGL(k_fock_coef, dim3(cdiv((int64_t)n5 * kp, 256)), dim3(256), 0, s, cj, ck, nocc, n5, ini, nbeta,
                     nalpha, kp, (int)open_shell, A, hc);
  const bool big = n5 > 64 && nn >= (1 << 16);
  // 64 x 128 tiles: at nbf = 1000, nbf5 = 300 the GEMM takes 6.8 ms against 8.5 (160 x 128),
  // 7.4 (160 x 64, 128 x 64) and 9.7 (128 x 128): the K extent (2 nbf5) is short.
  const TileF64 tc = big ? TileF64{2, 4} : TileF64{2, 2};
  ProbFock pr;
  pr.A = A;
  pr.J = J;
  pr.K = K;
  pr.H = Hd;
  pr.hc = hc;
  pr.F = F;
  pr.ld = ld;
  pr.nn = nn;
  pr.nbf = nbf;
  pr.nbf5 = n5;
  pr.kp = kp;
  pr.mtiles = cdiv(n5, 32 * tc.fm);
  const int64_t tiles = (int64_t)pr.mtiles * cdiv(nn, 32 * tc.fn);
  if (tiles > 2147483647LL) return hipErrorInvalidValue;
  return launch_prob(pr, tc, (int)tiles, s);
}

hipError_t launch_ymo_f64(const BuildGeom& g, const double* W, double* kmo, double* d, double* jmo, hipStream_t s) {
  const int64_t pairs = (int64_t)g.nbf5 * (g.nbf5 + 1) / 2;
  hipLaunchKernelGGL(k_ymo<double>, dim3((unsigned)pairs), dim3(256), 0, s, W, g.nbf5, g.naux, g.naux_p, kmo, d,
                     jmo);
  return hipGetLastError();
}

hipError_t launch_ymo_f32(const BuildGeom& g, const float* W, double* kmo, double* d, double* jmo, hipStream_t s) {
  const int64_t pairs = (int64_t)g.nbf5 * (g.nbf5 + 1) / 2;
  hipLaunchKernelGGL(k_ymo<float>, dim3((unsigned)pairs), dim3(256), 0, s, W, g.nbf5, g.naux, g.naux_p, kmo, d,
                     jmo);
  return hipGetLastError();
}

hipError_t launch_half_diag(const double* btri, int nbf, int naux_p, double* dhalf, double* zeros, hipStream_t s) {
  const int64_t n = (int64_t)nbf * naux_p + naux_p;
  hipLaunchKernelGGL(k_half_diag, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, btri, nbf, naux_p, dhalf, zeros);
  return hipGetLastError();
}

hipError_t launch_lagrange_fock_f64(const BuildGeom& g, const double* F, const double* Cf, int ldcf, double* G,
                                    double* tmp, int noptorb, double* elag, hipStream_t s) {
  hipLaunchKernelGGL(k_fock_g, dim3(cdiv(g.nbf, 32), g.nbf5), dim3(256), 0, s, F, Cf, ldcf, g.nbf, g.ldcp, G);
  hipError_t e = launch_hmat_f64(g, Cf, ldcf, G, tmp, s);
  if (e != hipSuccess) return e;
  const int64_t n = (int64_t)g.nbf * g.nbf;
  hipLaunchKernelGGL(k_elag_scatter, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, s, tmp, g.nbf, g.nbf5, noptorb,
                     elag);
  return hipGetLastError();
}

// Eigenvectors of the metric (column-major from the eigensolver: vector i at M[i*n ..]) and
// eigenvalues D (ascending) -> Vp[i][r] = V[r][i], Vs[i][r] = f(D_i) V[r][i], both ld x ld,
// zero-padded. f is the inverse square root with the cut of the psi4 Matrix::power used
// at integrals.py:42: |D_i| < cutoff * max|D| -> 0; a non-finite power -> 0.
__global__ void k_metric_scale(const double* __restrict__ D, const double* __restrict__ M, int n, int ld,
                               double cutoff, double* __restrict__ Vp, double* __restrict__ Vs) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)ld * ld) return;
  const int i = (int)(idx / ld), r = (int)(idx % ld);
  double v = 0.0, f = 0.0;
  if (i < n && r < n) {
    v = M[(int64_t)i * n + r];
    const double lo = fabs(D[0]), hi = fabs(D[n - 1]);
    const double amax = hi > lo ? hi : lo;
    const double a = D[i];
    if (!(fabs(a) < cutoff * amax)) {
      f = pow(a, -0.5);
      if (!isfinite(f)) f = 0.0;
    }
  }
  Vp[idx] = v;
  Vs[idx] = f * v;
}

hipError_t launch_metric_scale(const double* D, const double* M, int n, int ld, double cutoff, double* Vp, double* Vs,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_metric_scale, dim3(cdiv((int64_t)ld * ld, 256)), dim3(256), 0, s, D, M, n, ld, cutoff, Vp, Vs);
  return hipGetLastError();
}

hipError_t launch_atb_f64(const double* A, const double* B, int K, double* out, hipStream_t s) {
  if (K % 64 != 0) return hipErrorInvalidValue;
  ProbAtB pr;
  pr.A = A;
  pr.B = B;
  pr.out = out;
  pr.K = K;
  pr.mtiles = K / 64;
  return launch_prob(pr, TileF64{2, 2}, pr.mtiles * pr.mtiles, s);
}

hipError_t launch_rowgemm_f64(const double* A, int64_t rows, const double* W, int K, double* out, hipStream_t s) {
  if (K % 64 != 0 || rows <= 0 || rows > 2000000000LL) return hipErrorInvalidValue;
  const bool big = rows >= 4096 && K >= 256;
  const TileF64 tc = big ? TileF64{5, 4} : TileF64{2, 2};
  ProbRowGemm pr;
  pr.A = A;
  pr.W = W;
  pr.out = out;
  pr.rows = (int)rows;
  pr.K = K;
  pr.mtiles = cdiv(rows, 32 * tc.fm);
  const int64_t tiles = (int64_t)pr.mtiles * cdiv(K, 32 * tc.fn);
  if (tiles > 2147483647LL) return hipErrorInvalidValue;
  return launch_prob(pr, tc, (int)tiles, s);
}

// Convergence measures of the Lagrange multipliers (utils.py:405-411): out[0] = sum |e - e^T|,
// out[1] = max |e - e^T| over the nbf x nbf matrix. One workgroup: thread t visits i = t,
// t + 1024, ... in order, so the sum's order is fixed.
__global__ void __launch_bounds__(1024) k_lag_conv(const double* __restrict__ e, int n, double* __restrict__ out) {
  __shared__ double ssum[16], smax[16];
  double sum = 0.0, mx = 0.0;
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = threadIdx.x; idx < nn; idx += blockDim.x) {
    const int i = (int)(idx / n), j = (int)(idx % n);
    const double d = fabs(e[idx] - e[(int64_t)j * n + i]);
    sum += d;
    mx = d > mx ? d : mx;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_xor(sum, off, 64);
    const double o = __shfl_xor(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    ssum[wave] = sum;
    smax[wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0, m = 0.0;
    for (int w = 0; w < 16; ++w) {
      s += ssum[w];
      m = smax[w] > m ? smax[w] : m;
    }
    out[0] = s;
    out[1] = m;
  }
}

hipError_t launch_lag_conv(const double* elag, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_lag_conv, dim3(1), dim3(1024), 0, s, elag, n, out);
  return hipGetLastError();
}

hipError_t launch_hcore_f64(const BuildGeom& g, const double* Hp, int ldh, const double* Cp, double* T,
                            double* hcore, hipStream_t s) {
  const TileF64 tc{2, 2};
  ProbGemmT pr;
  pr.Hp = Hp;
  pr.Cp = Cp;
  pr.T = T;
  pr.ldh = ldh;
  pr.ldcp = g.ldcp;
  pr.nbf = g.nbf;
  pr.nbf_p = g.nbf_p;
  pr.mtiles = cdiv(g.nbf, 64);
  hipError_t e = launch_prob(pr, tc, pr.mtiles * cdiv(g.ldcp, 64), s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_hcore, dim3(cdiv(g.nbf5, 4)), dim3(256), 0, s, g.nbf, g.nbf5, g.ldcp, Cp, T, hcore);
  return hipGetLastError();
}

// Inverse of k_pack_b for rows mu0 <= mu < mu1: out[(mu - mu0) * nbf + nu][P] = btri[tri(max, min)][P].
__global__ void k_unpack_b(int nbf, int naux, int naux_p, int mu0, const double* __restrict__ btri,
                           double* __restrict__ out) {
  const int64_t row = blockIdx.x;
  const int mu = mu0 + (int)(row / nbf), nu = (int)(row % nbf);
  const double* src = btri + (mu >= nu ? tri_index(mu, nu) : tri_index(nu, mu)) * naux_p;
  double* dst = out + row * naux;
  for (int p = threadIdx.x; p < naux; p += blockDim.x) dst[p] = src[p];
}

hipError_t launch_unpack_b(int nbf, int naux, int naux_p, int mu0, int mu1, const double* btri, double* out,
                           hipStream_t s) {
  const int64_t rows = (int64_t)(mu1 - mu0) * nbf;
  if (rows <= 0) return hipSuccess;
  if (rows > 2147483647LL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_unpack_b, dim3((unsigned)rows), dim3(256), 0, s, nbf, naux, naux_p, mu0, btri, out);
  return hipGetLastError();
}

hipError_t launch_pack_b(int nbf, int naux, int naux_p, int mu0, int mu1, const double* b_chunk, double* btri,
                         double* asym_max, unsigned long long* nonfinite, hipStream_t s) {
  const int64_t rows = (int64_t)(mu1 - mu0) * nbf;
  if (rows <= 0) return hipSuccess;
  if (rows > 2147483647LL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_pack_b, dim3((unsigned)rows), dim3(256), 0, s, nbf, naux, naux_p, mu0, mu1, b_chunk, btri,
                     asym_max, nonfinite);
  return hipGetLastError();
}

}  // namespace pnjk
