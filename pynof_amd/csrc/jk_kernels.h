// Host-side launch interface of the JK build kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "jk_common.h"

namespace pnjk {

// ---------------------------------------------------------------- grouped launches
// A grouped launch runs the tiles of many instances of one problem type and
// tile shape in one kernel, one workgroup per tile. Launch functions given a
// collector do not launch: they append the instance's problem struct and its valid
// tile ids. Workgroup b of a grouped launch runs on XCD b % 8 and takes entry b / 8
// of that XCD's range [xcd_bounds[b % 8], xcd_bounds[b % 8 + 1]) of the table: one
// contiguous run per XCD (tiles that share an operand stay in one L2), the runs cut
// at equal k-step totals.
constexpr int kXcds = 8;
typedef hipError_t (*GroupedFn)(const void* probs, const int2* tiles, const int* xcd_bounds, int grid,
                                hipStream_t s);

struct GroupBatch {
  std::string key;        // problem type and tile shape
  size_t prob_size = 0;
  std::vector<char> probs;  // packed problem structs, one per instance slot
  std::vector<int2> tiles;  // (slot, local tile id)
  std::vector<int> ksteps;  // k-steps of each tile: its cost
  GroupedFn launch = nullptr;
  size_t probs_off = 0, tiles_off = 0, bounds_off = 0;  // offsets of the device copies in the plan buffer
  int grid = 0;                                         // workgroups of the launch

  // kXcds + 1 cut points of the table with equal k-step totals between them; the
  // launch grid is kXcds times the longest run.
  std::vector<int> xcd_bounds() const {
    std::vector<int> b(kXcds + 1, 0);
    int64_t total = 0;
    for (int k : ksteps) total += k;
    int64_t run = 0;
    size_t i = 0;
    for (int x = 1; x < kXcds; ++x) {
      const int64_t target = total * x / kXcds;
      while (i < ksteps.size() && run + ksteps[i] <= target) run += ksteps[i++];
      b[x] = (int)i;
    }
    b[kXcds] = (int)tiles.size();
    return b;
  }
};

struct GroupCollector {
  std::vector<GroupBatch> batches;
  template <class Prob>
  void add(const Prob& pr, TileF64 tc, int nblocks, int BM, int BN, int ksteps, GroupedFn fn, const char* type) {
    const std::string key = std::string(type) + "/" + std::to_string(tc.fm) + "x" + std::to_string(tc.fn);
    GroupBatch* b = nullptr;
    for (GroupBatch& x : batches)
      if (x.key == key) b = &x;
    if (!b) {
      batches.emplace_back();
      b = &batches.back();
      b->key = key;
      b->prob_size = sizeof(Prob);
      b->launch = fn;
    }
    const int slot = (int)(b->probs.size() / sizeof(Prob));
    const char* raw = reinterpret_cast<const char*>(&pr);
    b->probs.insert(b->probs.end(), raw, raw + sizeof(Prob));
    for (int id = 0; id < nblocks; ++id) {
      Tile t;
      if (pr.tile(id, BM, BN, t)) {
        b->tiles.push_back(make_int2(slot, id));
        b->ksteps.push_back(ksteps);
      }
    }
  }
};

// Lays out n pieces of bytes[i] bytes (0: piece not needed, ptr[i] = null) one after the
// other from base, each at a multiple of 256 bytes; returns the bytes used. With base null
// only the size is computed (ptr untouched).
inline size_t ws_carve(const size_t* bytes, int n, char* base, void** ptr) {
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    if (base) ptr[i] = bytes[i] ? base + off : nullptr;
    if (bytes[i]) off += (bytes[i] + 255) / 256 * 256;
  }
  return off;
}

// Geometry of one build. Row strides are in elements.
struct BuildGeom {
  int nbf;        // basis functions
  int naux;       // auxiliary functions
  int naux_p;     // padded aux stride (multiple of 64), zero-filled tail
  int nbf5;       // active orbitals
  int64_t ntri;   // nbf*(nbf+1)/2 packed pair rows
  int ldcp;       // row stride of the padded C block Cp[nbf_p][ldcp]; the
                  // allocation carries slack rows so whole tiles stay in bounds
  int nbf_p;      // rows of Cp / columns of Hp (multiple of 32), zero-filled tail
};

// X[q][nu*naux_p + P] = sum_mu Cp[mu][q] b_tri[tri(mu,nu)][P]   (half-transform)
hipError_t launch_ht_f64(const BuildGeom& g, TileF64 tc, const double* Cp, const double* btri, double* X,
                         hipStream_t s, GroupCollector* col = nullptr);

// Y_p[q][P] = sum_nu C[nu][q] X[p][nu][P] over the cyclic window of rows q (qt_window);
// writes the row sums of Y_p[q]^2 per 64 aux columns of the pairs batch p owns into
// kpart[(max(p,q)*nbf5 + min(p,q))*nt + tile] and d[p][P] = Y_p[p][P]. Cw: the two window
// copies of C (launch_window_c), ldw >= qt_window_cols(nbf5). Returns the number of
// 64-column partials in *ntiles. window = 1 computes window row 0 only (d, and the diagonal
// pair's partials): each output element's MFMA sequence does not depend on the other rows,
// so d has the bits of the whole window's; 0 is the whole window.
hipError_t launch_qt_f64(const BuildGeom& g, TileF64 tc, const double* Cw, int ldw, const double* X, double* kpart,
                         double* d, int* ntiles, hipStream_t s, GroupCollector* col = nullptr, int window = 0);
// The refinement's FP64 MO integrals from the symmetric b = L + L^T (L: b below the diagonal,
// b / 2 on it). T[q][mu][P] = sum_{nu <= mu} C[nu][q] L[mu][nu][P] (T as X: rows of naux_p,
// nbf rows per orbital); dhalf (nbf x naux_p) and zeros (naux_p) from launch_half_diag.
hipError_t launch_htl_f64(const BuildGeom& g, TileF64 tc, const double* Cp, const double* btri, const double* dhalf,
                          const double* zeros, double* T, hipStream_t s);
hipError_t launch_half_diag(const double* btri, int nbf, int naux_p, double* dhalf, double* zeros, hipStream_t s);
// W (nbf5 x nbf5 rows of naux_p): row (x, y) = W_xy = sum_mu C[mu][x] T[y][mu][P], i.e. C^T L C.
hipError_t launch_w_f64(const BuildGeom& g, TileF64 tc, const double* Cp, const double* T, double* W,
                        hipStream_t s);
// From W: K_MO (both triangles) = sum_{P < naux} (W_pq + W_qp)^2 and d[p][P] = 2 W_pp[P];
// with jmo (else NULL) also J_MO = d d^T, in the same pass.
hipError_t launch_ymo_f64(const BuildGeom& g, const double* W, double* kmo, double* d, double* jmo, hipStream_t s);
// The same from the split route's FP32 W (sums in FP64).
hipError_t launch_ymo_f32(const BuildGeom& g, const float* W, double* kmo, double* d, double* jmo, hipStream_t s);
// Cw (2 x rows x ldw): Cw0[nu][j] = Cp[nu][j mod n], Cw1[nu][j] = Cp[nu][(j + 1) mod n].
hipError_t launch_window_c(const double* Cp, int ldcp, int rows, int n, int ldw, double* Cw, hipStream_t s);

// K_MO[p][q] = K_MO[q][p] = sum_t kpart[(p*nbf5+q)*nt + t]  (q <= p)
// Orbital-gradient MO integrals: b_MO[p][q][P] (p < nbf, q < nbf5, P < naux) from X
// and the full-width C (Cf, ldcf >= nbf rounded up to the tile rows); H_mat = Cf^T T.
hipError_t launch_mo_f64(const BuildGeom& g, TileF64 tc, const double* Cf, int ldcf, const double* X, double* b_mo,
                         int64_t ldrow, hipStream_t s);
// Orbital-gradient Lagrangian from b_MO (rows of stride ldrow = nchunks * kc, zero past
// nbf5 * naux), Hm = H_mat, occupations and zero-diagonal C^J, C^K (device memory):
// elag (nbf x nbf). G (nbf5 x naux), W (nbf5 x ldrow), part (nchunks x nbf x nbf5) are workspace.
hipError_t launch_lag_f64(const BuildGeom& g, const double* B, int64_t ldrow, int64_t kc, int nchunks,
                          const double* nocc, const double* cj, const double* ck, int nbeta, int nalpha,
                          const double* Hm, double* G, double* W, double* part, double* elag, bool exchange,
                          hipStream_t s);
// BF16x3 exchange partials of the Lagrangian (split form): converts B to planes (bh, bl:
// nbf x ldrow) and forms W16 (wh, wl: nbf5 x ldrow); part as in launch_lag_f64.
hipError_t launch_lag_x3(int nbf, int nbf5, int naux, const double* B, int64_t ldrow, int64_t kc, int nchunks,
                         const double* ck, uint16_t* bh, uint16_t* bl, uint16_t* wh, uint16_t* wl, double* part,
                         bool planes_ready, hipStream_t s);
hipError_t launch_lag_coef(const double* src, double* dst, int n, hipStream_t s);
// Generalised Fock matrices F (nbf5 x nbf x nbf), closed shell or (open_shell) restricted open
// shell, from J^n, K^n rows of stride ld (even, >= nbf^2), device H (nbf x nbf), occupations and
// coefficients; A (nbf5 x roundup(2 nbf5, 16)) and hc (nbf5) are workspace.
hipError_t launch_fock_f64(int nbf, int n5, const double* J, const double* K, int64_t ld, const double* Hd,
                           const double* nocc, const double* cj, const double* ck, int ini, int nbeta, int nalpha,
                           bool open_shell, double* A, double* hc, double* F, hipStream_t s);
// Lagrange multipliers from F (nbf5 x nbf x nbf): G (nbf x ldcp, + 64 slack, columns past
// nbf5 finite) and tmp (nbf x nbf5) workspace; Cf the full-width C of ldcf columns (as for
// the MO integrals); elag nbf x nbf.
hipError_t launch_lagrange_fock_f64(const BuildGeom& g, const double* F, const double* Cf, int ldcf, double* G,
                                    double* tmp, int noptorb, double* elag, hipStream_t s);
hipError_t launch_fill_cols(double* p, int rows, int64_t ld, int64_t c0, int64_t c1, hipStream_t s);
hipError_t launch_hmat_f64(const BuildGeom& g, const double* Cf, int ldcf, const double* T, double* hmat,
                           hipStream_t s);
hipError_t launch_kmo_reduce(int nbf5, int nt, const double* kpart, double* kmo, hipStream_t s);

// O[b][i][j] = sum_k A[b][i][k] A[b][j][k] for i, j < M, computed on the lower
// triangle of tiles: each value with i >= j is written at (i, j) and, for i > j,
// also at (j, i), so O is exactly symmetric with no separate pass.
hipError_t launch_syrk_f64(int batch, int M, int K, const double* A, int64_t lda, int64_t sa, double* O,
                           int64_t ldo, int64_t so, TileF64 tc, hipStream_t s, GroupCollector* col = nullptr);

// J[q][mu][nu] = sum_P d[q][P] b_tri[tri(mu,nu)][P] for mu >= nu (lower triangle only).
hipError_t launch_jn_f64(const BuildGeom& g, TileF64 tc, const double* d, const double* btri, double* J,
                         hipStream_t s, GroupCollector* col = nullptr);

// O[b][j][i] = O[b][i][j] for i > j, b < batch (n x n matrices, row stride ld).
hipError_t launch_symmetrize(int batch, int n, double* O, int64_t ld, int64_t stride, hipStream_t s);
// launch_symmetrize over many batches of matrices in one launch: job j covers workgroups
// [first, first + count * pairs(n)) of the grid, jobs in increasing `first`.
struct SymJob {
  double* O;
  int64_t ld, stride, first;
  int n, count;
};
hipError_t launch_symmetrize_grouped(const SymJob* jobs, int njobs, int64_t blocks, hipStream_t s);
__host__ __device__ inline int64_t sym_blocks(int n) {  // 32 x 32 lower tile pairs of an n x n matrix
  const int64_t t = (n + 31) / 32;
  return t * (t + 1) / 2;
}

// T[mu][i] = sum_nu Hp[mu][nu] Cp[nu][i] (mu < nbf, i < ldcp), then
// hcore[i] = sum_mu Cp[mu][i] T[mu][i] for i < nbf5.
hipError_t launch_hcore_f64(const BuildGeom& g, const double* Hp, int ldh, const double* Cp, double* T,
                            double* hcore, hipStream_t s);

// Packs rows mu in [mu0, mu1) of a full (nbf, nbf, naux) b (b_chunk holds
// b[mu0:mu1, :, :]) into btri rows t = tri(mu, nu), nu <= mu, stride naux_p with
// a zero tail. Chunks must be packed in descending mu order on one stream.
// Folds max |b[mu,nu,P] - b[nu,mu,P]| into *asym_max and counts non-finite
// values into *nonfinite (both device scalars).
// Rows mu0 <= mu < mu1 of b in the (nbf, nbf, naux) layout from the packed rows.
hipError_t launch_unpack_b(int nbf, int naux, int naux_p, int mu0, int mu1, const double* btri, double* out,
                           hipStream_t s);
hipError_t launch_pack_b(int nbf, int naux, int naux_p, int mu0, int mu1, const double* b_chunk, double* btri,
                         double* asym_max, unsigned long long* nonfinite, hipStream_t s);

// ---------------------------------------------------------------- split (BF16x3)
struct SplitGeom {
  int nbf;     // basis functions
  int naux_p;  // padded aux stride (multiple of 64)
  int nbf5;    // active orbitals
  int ldk;     // contraction stride over mu/nu of Ct16 and bht16 (nbf rounded up to 32, zero tail)
};

// X16[q] holds the (nu, P) matrix of orbital q in 16-row x 32-column blocks (blk_index in
// jk_split.hip) over nbf rounded up to 16 rows: this many elements per orbital.
__host__ __device__ inline int64_t x16_stride(int nbf, int naux_p) { return (int64_t)((nbf + 15) / 16 * 16) * naux_p; }
// Elements past the last orbital of X16 that a 128-column quarter-transform tile may read (64
// columns past naux_p of the last row block land up to two 512-element blocks further on).
constexpr int64_t kX16Slack = 1024;

// Split planes of a row-major rows x cols matrix (cols a multiple of 32) in 16-row x 32-column
// blocks (blk_index in jk_split.hip); rows rounded up to 16 with zeros.
hipError_t launch_split_blocked(const double* src, int64_t rows, int64_t cols, uint16_t* hi, uint16_t* lo,
                                hipStream_t s);
// X16 planes (x16_stride layout, zero pad rows) of FP64 X[q][nu][P] (row stride naux_p, nbf rows
// per orbital) for nbf5 orbitals.
hipError_t launch_split_x16(const double* X, const SplitGeom& g, uint16_t* hi, uint16_t* lo, hipStream_t s);
// bht planes: row n = blk_index(nu, P, naux_p / 32) of the (nu, P) matrix (nu < nbf rounded up to
// 16, zero rows past nbf) holds b[mu][nu][P] over mu < ldk (zero for mu >= nbf); the rows are
// stored in 16 x 32 blocks over (n, mu). The half-transform's output row X16[q][n] is then in
// the x16_stride layout.
hipError_t launch_build_bht(const double* btri, int nbf, int naux_p, int ldk, uint16_t* hi, uint16_t* lo,
                            hipStream_t s);
// Ct[q][mu] planes (rows q < rows rounded up to 16, in 16 x 32 blocks, blk_index) from Cp[mu][q].
hipError_t launch_split_ct(const double* Cp, int ldcp, int nbf, int nbf5, int rows, int ldk, uint16_t* hi,
                           uint16_t* lo, hipStream_t s);
// X16[q] = split(sum_mu C[mu][q] b[mu][nu][P]) in the x16_stride layout.
hipError_t launch_ht_x3(const SplitGeom& g, TileF64 tc, const uint16_t* ct_hi, const uint16_t* ct_lo,
                        const uint16_t* bht_hi, const uint16_t* bht_lo, uint16_t* x_hi, uint16_t* x_lo,
                        hipStream_t s, GroupCollector* col = nullptr);
// The split triangular MO route: T16[mu][q][P] = split(sum_{nu <= mu} Ct[q][nu] L[mu][nu][P]) from
// bt16 with the split halved diagonal rows of b (and a zero row) stored from row hd on; then
// W[p][(q, P)] = sum_mu Ct[p][mu] T16[mu][(q, P)] in FP32 (nbf5 x nbf5 naux_p).
hipError_t launch_tl_x3(const SplitGeom& g, const uint16_t* ct_hi, const uint16_t* ct_lo, const uint16_t* bt_hi,
                        const uint16_t* bt_lo, int64_t hd, uint16_t* t_hi, uint16_t* t_lo, hipStream_t s);
hipError_t launch_ws_x3(const SplitGeom& g, const uint16_t* ct_hi, const uint16_t* ct_lo, const uint16_t* t_hi,
                        const uint16_t* t_lo, float* W, hipStream_t s);
// Quarter transform from X16: K_MO row-sum partials, d (FP64 and split planes).
// BF16x3 orbital-gradient MO integrals from X16 and the split planes of C^T over
// all nbf columns (rows clamped at nbf); FP64 output b_MO[p][q][P].
// With bmo_hi/bmo_lo, also writes the bf16 split planes of b_MO (same layout, valid columns).
hipError_t launch_mo_x3(const SplitGeom& g, int naux, TileF64 tc, const uint16_t* ct_hi, const uint16_t* ct_lo,
                        const uint16_t* x_hi, const uint16_t* x_lo, double* b_mo, int64_t ldrow, hipStream_t s,
                        uint16_t* bmo_hi = nullptr, uint16_t* bmo_lo = nullptr);
// Zeroes columns [c0, c1) of two bf16 planes (rows of stride ld).
hipError_t launch_fill_cols16(uint16_t* h, uint16_t* l, int rows, int64_t ld, int64_t c0, int64_t c1, hipStream_t s);
// The BF16x3 quarter transform, as launch_qt_f64 (window likewise).
hipError_t launch_qt_x3(const SplitGeom& g, TileF64 tc, const uint16_t* ct_hi, const uint16_t* ct_lo,
                        const uint16_t* x_hi, const uint16_t* x_lo, double* kpart, double* d, uint16_t* d_hi,
                        uint16_t* d_lo, int* ntiles, hipStream_t s, GroupCollector* col = nullptr, int window = 0);
// J^n from split d and btri planes (FP64 output, mu >= nu only).
hipError_t launch_jn_x3(const SplitGeom& g, TileF64 tc, const uint16_t* d_hi, const uint16_t* d_lo,
                        const uint16_t* bt_hi, const uint16_t* bt_lo, double* J, hipStream_t s,
                        GroupCollector* col = nullptr);
// K^n from split X16 (x16_stride layout; FP64 output, lower triangle of tiles only).
hipError_t launch_kn_x3(const SplitGeom& g, TileF64 tc, const uint16_t* x_hi, const uint16_t* x_lo, double* K,
                        hipStream_t s, GroupCollector* col = nullptr);

// out[0] = sum |e - e^T|, out[1] = max |e - e^T| of an n x n matrix (device out[2]).
hipError_t launch_lag_conv(const double* elag, int n, double* out, hipStream_t s);

// ---------------------------------------------------------------- b from 3-centre integrals
// Vp, Vs (ld x ld, zero-padded) from the metric's eigenpairs; Vs rows scaled by D^-1/2 with
// the relative eigenvalue cut.
hipError_t launch_metric_scale(const double* D, const double* M, int n, int ld, double cutoff, double* Vp, double* Vs,
                               hipStream_t s);
// out = A^T B, K x K (K a multiple of 64).
hipError_t launch_atb_f64(const double* A, const double* B, int K, double* out, hipStream_t s);
// out (rows x K) = A (rows x K) W (K x K, + 128 doubles of slack); K a multiple of 64.
hipError_t launch_rowgemm_f64(const double* A, int64_t rows, const double* W, int K, double* out, hipStream_t s);

// ---------------------------------------------------------------- occupation step
// PNOF functional and occupation parametrisation (the param fields of parameters.py:46-110).
struct PnofDev {
  int ipnof;      // 5, 7 or 8 (GNOF)
  int ista;       // PNOF7 Phi: 0 sqrt(n(1-n)), else 2n(1-n)
  int method;     // 0 softmax, 1 trigonometric
  int high_spin;  // singly occupied n = 1 instead of 0.5
  int mspin;      // 0 closed shell; otherwise the open-shell energy and gradient
  int no1, ndoc, nsoc, ncwo, nbf5, nv, nbeta, nalpha, ndns;
};

// n (nbf5) and dn (nbf5 x nv, orbital-major) from the parameters x (nv).
hipError_t launch_occupation(const PnofDev& p, const double* x, double* n, double* dn, hipStream_t s);
// Per-orbital Phi, n_d, sqrt(n_d); with dn, also their derivatives (nbf5 x nv each).
hipError_t launch_pnof_orbitals(const PnofDev& p, const double* n, const double* dn, double* fi, double* nd,
                                double* nd12, double* dfi, double* dnd, double* dnd12, hipStream_t s);
// C^J, C^K (nbf5 x nbf5); either may be null.
hipError_t launch_pnof_coef(const PnofDev& p, const double* n, const double* fi, const double* nd, const double* nd12,
                            double* cj, double* ck, hipStream_t s);
// dst[i][j] = src[j][i], n x n.
hipError_t launch_occ_transpose(const double* src, int n, double* dst, hipStream_t s);
// Closed-shell energy into E[0]; C^J, C^K from cjm/ckm when given, else from the PNOF.
// part: ceil(nbf5 / 4) doubles of workspace.
hipError_t launch_occ_energy(const PnofDev& p, const double* n, const double* fi, const double* nd, const double* nd12,
                             const double* cjm, const double* ckm, const double* JT, const double* KT,
                             const double* hc, double* part, double* E, hipStream_t s);
// Occupation gradient (nv).
hipError_t launch_occ_grad(const PnofDev& p, const double* n, const double* fi, const double* nd, const double* nd12,
                           const double* dn, const double* dfi, const double* dnd, const double* dnd12,
                           const double* JT, const double* KT, const double* hc, double* grad, hipStream_t s);

}  // namespace pnjk
