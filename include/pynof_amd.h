/* pynof_amd — MI355X-native density-fitted JK build: the C ABI.
 *
 * The entry points replace the RI branch of PyNOF's JK builds (paths relative
 * to the PyNOF source tree):
 *   computeJKj(C, I, b_mnl, p)        src/pynof/integrals.py:172-191  -> J^n, K^n
 *     JKj_RI        integrals.py:208-218 (numpy), JKj_RI_GPU :234-245 (cupy),
 *     JKj_RI_jit    integrals.py:282-301 (numba)
 *   computeJKH_MO(C, H, I, b_mnl, p)  src/pynof/integrals.py:307-326  -> J_MO, K_MO, H_core
 *     JKH_MO_RI     integrals.py:348-362, JKH_MO_RI_GPU :383-399, JKH_MO_RI_jit :457-484
 *   the device upload of b          src/pynof/integrals.py:51-53  (cp.array(b_mnl))
 *
 * Conventions: plain pointers and sizes; arrays are C-order FP64. A pointer may
 * be host or device memory (HIP unified addressing decides) unless stated.
 * Every call returns PNJK_OK (0) or an error code; pnjk_last_error() gives the
 * calling thread's message. Calls on one engine must not overlap.
 */
#ifndef PYNOF_AMD_H
#define PYNOF_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PNJK_OK 0
#define PNJK_EINVAL 1   /* bad argument: sizes, null pointers, non-finite or asymmetric b */
#define PNJK_EHIP 2     /* a HIP runtime call failed */
#define PNJK_ENOMEM 3   /* device allocation failed */
#define PNJK_ESTATE 4   /* engine not ready (no b uploaded) */
#define PNJK_EUNSUPPORTED 5

/* Arithmetic of a build. */
#define PNJK_FP64 0          /* every contraction in FP64 (MFMA f64) */
#define PNJK_SPLIT 1         /* bulk contractions as BF16x3 split products, FP32 accumulation */
#define PNJK_SPLIT_REFINE 2  /* SPLIT for J^n/K^n; MO quantities refined to FP64 accuracy */

/* Output selection bits for pnjk_jk_build. */
#define PNJK_OUT_J 1u
#define PNJK_OUT_K 2u
#define PNJK_OUT_JMO 4u
#define PNJK_OUT_KMO 8u
#define PNJK_OUT_HCORE 16u
#define PNJK_OUT_ALL 31u

typedef struct pnjk_engine pnjk_engine;

/* Last error message of the calling thread ("" if none). */
const char* pnjk_last_error(void);

/* Library version string. */
const char* pnjk_version(void);

/* Creates an engine bound to HIP device `device`. Replaces the implicit cupy
 * device of integrals.py:8-11/:51-53. */
int pnjk_create(int device, pnjk_engine** out);
int pnjk_destroy(pnjk_engine* e);

/* Uploads b_mnl (nbf, nbf, naux), C-order, host or device memory, and keeps it
 * resident in the engine's packed layout (lower triangle mu >= nu, aux stride
 * padded). Replaces cp.array(b_mnl) at integrals.py:51-53. Rejects non-finite
 * values, and b whose max |b[mu,nu,P] - b[nu,mu,P]| exceeds `sym_tol` (b is
 * symmetric in mu,nu by construction, integrals.py:35-49). Writes the resident
 * footprint in bytes to *footprint (may be NULL) and the measured asymmetry to
 * *asym (may be NULL). */
int pnjk_upload_b(pnjk_engine* e, const double* b_mnl, int64_t nbf, int64_t naux, double sym_tol,
                  uint64_t* footprint, double* asym);

/* Forms b on the device from the 3-centre integrals and the metric, as the
 * reference's b producer does (integrals.py:35-49): mnk[mu][nu][P] = (mu nu|P)
 * (nbf, nbf, naux), C-order, symmetric in (mu, nu), host or device;
 * metric[P][Q] = (P|Q) (naux, naux), host or device. The metric's inverse square
 * root is taken as psi4's Matrix::power(-0.5, cutoff) does (integrals.py:41-42;
 * psi4 is not pinned by the reference): eigenpairs (rocSOLVER dsyevd), eigenvalues
 * with |a| < cutoff * max|a| dropped, V diag(a^-1/2) V^T. Then
 * b[mu][nu][P] = sum_Q (mu nu|Q) [(P|Q)^-1/2][Q][P] (integrals.py:49) as one FP64
 * GEMM over the packed mu >= nu rows, left resident as pnjk_upload_b leaves b.
 * mnk is checked like b in pnjk_upload_b (finiteness, symmetry within sym_tol,
 * *asym). The reference passes cutoff = 1e-14. */
int pnjk_upload_b_3c(pnjk_engine* e, const double* mnk, int64_t nbf, int64_t naux, const double* metric,
                     double cutoff, double sym_tol, uint64_t* footprint, double* asym);

/* Writes the resident b back in the reference layout (nbf, nbf, naux), C-order,
 * to host or device memory (host memory, e.g. a memory-mapped .npy, is filled in
 * mu-row chunks). The inverse of pnjk_upload_b; no counterpart in the reference,
 * which keeps b as a numpy/cupy array (integrals.py:49-53). */
int pnjk_download_b(pnjk_engine* e, double* b_mnl);

/* Aux stride of the packed layout the engine uses for (nbf, naux). */
int64_t pnjk_packed_ld(int64_t nbf, int64_t naux);

/* Adopts an already packed b: device rows t = mu*(mu+1)/2 + nu (mu >= nu), each
 * `ld` doubles with b[mu,nu,0:naux] first (ld = pnjk_packed_ld(nbf, naux), zero
 * tail). The data is copied. */
int pnjk_upload_b_packed(pnjk_engine* e, const double* btri_dev, int64_t nbf, int64_t naux, int64_t ld,
                         uint64_t* footprint);

/* One fused JK build for orbitals C (element (mu, q) at C[mu*ldc + q], mu < nbf,
 * q < nbf5; only the first nbf5 columns are read, integrals.py:211,:353) and,
 * when PNJK_OUT_HCORE is set, H (element (mu, nu) at H[mu*ldh + nu]).
 * Outputs (C-order, FP64), each written when its bit is set in `outputs`:
 *   J, K     (nbf5, nbf, nbf)   J^n, K^n           integrals.py:214, :216
 *   J_MO     (nbf5, nbf5)                          integrals.py:356
 *   K_MO     (nbf5, nbf5)                          integrals.py:358
 *   H_core   (nbf5)                                integrals.py:360
 * A build that asks for neither J nor K forms J_MO and K_MO by the triangular route
 * b = L + L^T (T = L C, W = C^T T, Y_pq = W_pq + W_qp) when that is the faster route:
 * nbf5 <= 2 nbf / 3 in PNJK_FP64 and nbf5 <= 2 nbf / 5 in PNJK_SPLIT. Its J_MO and K_MO
 * differ from the full route's in rounding only. PNJK_SPLIT_REFINE forms its FP64 J_MO and
 * K_MO by that route up to nbf5 <= 2 nbf / 3 (beside a split half-transform for K^n), and
 * past it by the FP64 form's half- and quarter transforms (bitwise the FP64 form's J_MO and
 * K_MO), splitting their FP64 X for K^n.
 * Output pointers may be host or device memory. `stream` is a hipStream_t
 * (NULL: the null stream, which is also torch's default stream). The call
 * returns after the outputs are complete when any output is host memory;
 * otherwise it is asynchronous on `stream`. */
int pnjk_jk_build(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H, int64_t ldh,
                  int precision, uint32_t outputs, double* J, double* K, double* J_MO, double* K_MO,
                  double* H_core, void* stream);

/* pnjk_jk_build without allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, outputs and bits as pnjk_jk_build,
 * always asynchronous on `stream`, and it calls no hipMalloc, hipFree or allocating
 * library routine: the build's workspace (C staging, the half-transform X, d, the K_MO
 * partials, ...) lives in the library's static device array of 64 MiB per GPU (held until
 * the process ends). C, H and every output must be device memory; for the split forms the
 * split planes of b must exist already (the first pnjk_jk_build in a split form after an
 * upload makes them), and so must the halved diagonal rows of b for an FP64 build that
 * takes the triangular MO route (the first such pnjk_jk_build, or any split one, makes
 * them). A build whose workspace exceeds 64 MiB (FP64 form, nbf5 = nbf: about
 * nbf = 125 with naux = 3.6 nbf; split_refine about 109) fails with PNJK_EINVAL and names the bytes it needs. All
 * no-allocation builds on one GPU share that array: calls that overlap in time on one GPU
 * (different streams, threads or engines) race; calls issued on one stream are ordered. */
int pnjk_jk_build_noalloc(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H,
                          int64_t ldh, int precision, uint32_t outputs, double* J, double* K, double* J_MO,
                          double* K_MO, double* H_core, void* stream);

/* Orbital-gradient MO integrals of the RI path: replaces JKH_MO_tmp with
 * p.RI (integrals.py:665-682), i.e. Integrals_MO_RI_CPU / _GPU
 * (integrals.py:703-716), which computeLagrange2 (utils.py:262-265) calls for
 * every orbital gradient:
 *   H_mat[i][j]   = sum_mn C[m][i] H[m][n] C[n][j]       i < nbf, j < nbf5
 *   b_MO[p][q][P] = sum_mn C[m][p] C[n][q] b[m][n][P]    p < nbf, q < nbf5
 * b_MO is C-order (nbf, nbf5, naux), H_mat (nbf, nbf5). C is (nbf, ldc >= nbf),
 * row-major; every column is used. precision: PNJK_FP64 or PNJK_SPLIT (BF16x3,
 * FP32 accumulation; b_MO then carries ~1e-5 relative error). Either output may
 * be NULL (not computed); H_mat needs H (ldh >= nbf). Outputs may be host or
 * device memory; with a host output the call returns after the copy back. */
int pnjk_mo_integrals(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H, int64_t ldh,
                      int precision, double* b_MO, double* H_mat, void* stream);

/* Orbital-gradient Lagrangian of the RI path, closed shell: replaces
 * computeLagrange2 with p.RI and p.MSpin == 0 (utils.py:262-270, :328-355),
 * which calcorbg / calcorbeg (pnof.py:1114-1154) call on every orbital step.
 * With B = b_MO and Hm = H_mat of pnjk_mo_integrals (kept on the device):
 *   elag[a][b] = n[b] Hm[a][b]
 *              + w_b n[b] sum_P B[a][b][P] B[b][b][P]   (w_b = 1 for b < nbeta or b >= nalpha)
 *              + sum_P B[a][b][P] sum_q cj12[b][q] B[q][q][P]
 *              - sum_q ck12[b][q] sum_P B[a][q][P] B[b][q][P]       for b < nbf5,
 *   elag[a][b] = 0 for b >= nbf5;  a < nbf.
 * n (nbf5), cj12 and ck12 (nbf5 x nbf5, row-major) are read with their
 * diagonals taken as zero (the reference zeroes the caller's arrays in place,
 * utils.py:269-270; this call leaves its inputs unchanged). elag is (nbf, nbf),
 * H_mat (nbf, nbf5) optional; host or device memory, as in pnjk_jk_build.
 * precision (PNJK_FP64 or PNJK_SPLIT) applies to b_MO; the contractions are FP64.
 * The orbital gradient is 4 (elag - elag^T) (pnof.py:1120). */
int pnjk_lagrangian(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H, int64_t ldh,
                    const double* n, const double* cj12, const double* ck12, int64_t nbeta, int64_t nalpha,
                    int precision, double* elag, double* H_mat, void* stream);

/* Generalised Fock matrices, closed shell: replaces computeF_RC_driver /
 * computeF_RC_CPU / computeF_RC_GPU (utils.py:15-54, :58-84), which ENERGY1r
 * (utils.py:415-427) and the orbital optimisers call when p.no1 > 0:
 *   F[i] = n[i] H + w_i n[i] J[i] + sum_j cj12'[i][j] J[j] - sum_j ck12'[i][j] K[j]
 * for i < nbf5, with w_i = 1 for ini <= i < nbeta or nalpha <= i < nbf5,
 * ini = no1 if no1 > 1 else 0 (utils.py:31-33), and cj12', ck12' the
 * coefficients with the diagonal of their [ini:, ini:] block taken as zero (the
 * reference zeroes it in the caller's arrays, utils.py:45, :50; this call leaves
 * its inputs unchanged). J, K, F are (nbf5, nbf, nbf) C-order; H (nbf, nbf) of
 * row stride ldh; n (nbf5); cj12, ck12 (nbf5, nbf5). Arrays may be host or
 * device memory; the engine needs no b. One FP64 GEMM over the stacked J and K
 * (MFMA f64) with n[i] H added in its epilogue. Needs nbf^2 <= 2^30. */
int pnjk_fock_rc(pnjk_engine* e, const double* J, const double* K, int64_t nbf, int64_t nbf5, const double* n,
                 const double* H, int64_t ldh, const double* cj12, const double* ck12, int64_t no1, int64_t nbeta,
                 int64_t nalpha, double* F, void* stream);

/* Generalised Fock matrices, restricted open shell: replaces computeF_RO_driver /
 * computeF_RO_CPU / computeF_RO_GPU (utils.py:87-93, :96-171, :174-246), which
 * ENERGY1r calls when p.no1 > 0 and p.MSpin != 0 (utils.py:422-425). With
 * D = [0, nbeta), S = [nbeta, nalpha) (singly occupied), W = [nalpha, nbf5):
 *   i in D or W: F[i] = n[i] H + w_i n[i] J[i] + sum_{j in D, W} (cj12'[i][j] J[j]
 *                - ck12'[i][j] K[j]) + n[i] sum_{s in S} (J[s] - K[s] / 2),
 *   i in S:      F[i] = H / 2 + sum_{s in S, s != i} (J[s] - K[s]) / 2
 *                + sum_{j in D, W} n[j] (J[j] - K[j] / 2),
 * with w_i, ini and the zero diagonals as for pnjk_fock_rc. Same arguments, memory
 * rules and limits as pnjk_fock_rc; the same GEMM with these coefficients. */
int pnjk_fock_ro(pnjk_engine* e, const double* J, const double* K, int64_t nbf, int64_t nbf5, const double* n,
                 const double* H, int64_t ldh, const double* cj12, const double* ck12, int64_t no1, int64_t nbeta,
                 int64_t nalpha, double* F, void* stream);

/* Lagrange multipliers from the Fock matrices: replaces computeLagrange
 * (utils.py:249-259):
 *   G[m][i]    = sum_n F[i][m][n] C[n][i]                 i < nbf5
 *   elag[i][j] = sum_m C[m][i] G[m][j]   for i < noptorb, j < nbf5; 0 elsewhere.
 * F (nbf5, nbf, nbf) C-order, C (nbf, ldc >= nbf) row-major, elag (nbf, nbf).
 * Host or device memory; FP64 throughout; the engine needs no b. */
int pnjk_lagrange_fock(pnjk_engine* e, const double* F, int64_t nbf, int64_t nbf5, const double* C, int64_t ldc,
                       int64_t noptorb, double* elag, void* stream);

/* Convergence measures of the Lagrange multipliers: replaces
 * computeLagrangeConvergency (utils.py:405-411): *sumdiff = sum |elag - elag^T|,
 * *maxdiff = max |elag - elag^T| over the (nbf, nbf) matrix elag (host or
 * device). Waits for the result. */
int pnjk_lagrange_convergence(pnjk_engine* e, const double* elag, int64_t nbf, double* sumdiff, double* maxdiff,
                              void* stream);

/* ---- occupation step (pnof.py) ----
 * The PNOF functional and occupation parametrisation, the fields of PyNOF's param
 * (parameters.py:46-110) that the occupation step reads. nbf5 must equal
 * no1 + ndoc (1 + ncwo) + nsoc (parameters.py:65); nbeta = no1 + ndoc,
 * nalpha = nbeta + nsoc, nv = ndoc ncwo are derived. */
#define PNJK_OCC_SOFTMAX 0         /* p.occ_method == "Softmax" (pnof.py:770-810) */
#define PNJK_OCC_TRIGONOMETRIC 1   /* p.occ_method == "Trigonometric" (pnof.py:704-768) */
typedef struct pnjk_pnof {
  int32_t ipnof;      /* p.ipnof: 5 (PNOF5), 7 (PNOF7) or 8 (GNOF) */
  int32_t ista;       /* p.ista: PNOF7 Phi = sqrt(n(1-n)) for 0, else 2n(1-n) */
  int32_t occ_method; /* PNJK_OCC_SOFTMAX or PNJK_OCC_TRIGONOMETRIC */
  int32_t high_spin;  /* p.HighSpin: singly occupied n = 1 instead of 0.5 */
  int32_t mspin;      /* p.MSpin: 0 closed shell; otherwise the open-shell energy and gradient */
  int32_t no1, ndoc, nsoc, ncwo, nbf5;
} pnjk_pnof;

/* One occupation step from the parameters gamma (nv): replaces ocupacion
 * (pnof.py:692-810), PNOFi_selector (:657-667, CJCKD5/7/8 :155-422), calcocce
 * (:994-1001, calce :905-990), der_PNOFi_selector (:670-689, der_CJCKD5/7/8
 * :206-525) and calcoccg (:1004-1098), closed (MSpin = 0) and open shell, which the
 * occupation optimiser (minimization.py:86-88) calls on every iteration.
 * Outputs, each written when not NULL: E (1) the energy of calcocce; grad (nv)
 * of calcoccg; n (nbf5) and dn_dgamma (nbf5, nv) of ocupacion; cj12, ck12
 * (nbf5, nbf5) of PNOFi_selector. E and grad need J_MO, K_MO (nbf5, nbf5) and
 * H_core (nbf5), e.g. the device outputs of pnjk_jk_build. Arrays may be host
 * or device memory; with a host output the call returns after the copy back.
 * The reference's softmax writes the strong occupations at index i < ndoc
 * (pnof.py:792, :800), not no1 + i; this is reproduced as it is. */
int pnjk_occupation_step(pnjk_engine* e, const pnjk_pnof* p, const double* gamma, const double* J_MO,
                         const double* K_MO, const double* H_core, double* E, double* grad, double* n,
                         double* dn_dgamma, double* cj12, double* ck12, void* stream);

/* C^J, C^K (nbf5, nbf5) for given occupations n (nbf5): replaces PNOFi_selector
 * (pnof.py:657-667) for PNOF5, PNOF7 and GNOF. Either output may be NULL. */
int pnjk_pnof_coefficients(pnjk_engine* e, const pnjk_pnof* p, const double* n, double* cj12, double* ck12,
                           void* stream);

/* Energy for given n, cj12, ck12: replaces calce (pnof.py:905-990). MSpin = 0:
 * 2 sum n H_core + sum n_i J_ii over i < nbeta and nalpha <= i < nbf5
 * + sum' cj12 J_MO - sum' ck12 K_MO, diagonals excluded (the reference zeroes them
 * in the caller's arrays, :913, :917; this call leaves its inputs unchanged).
 * MSpin != 0: the cj12/ck12 sums skip the singly occupied block
 * [nbeta, nalpha), which enters through the n-weighted J_MO/K_MO terms of
 * :924-988. E is one double, host or device. */
int pnjk_energy(pnjk_engine* e, int64_t nbf5, int64_t nbeta, int64_t nalpha, int32_t mspin, const double* n,
                const double* cj12, const double* ck12, const double* J_MO, const double* K_MO, const double* H_core,
                double* E, void* stream);

/* Grouped builds: many engines (one b each, different sizes allowed) built in
 * one pass. Each GEMM stage of all instances runs as one grouped launch (one
 * workgroup per tile) per problem type and tile shape (tile shapes per instance from the
 * engines' tile settings), so small instances share the device instead of
 * each underfilling it. The reference has no batched form; it builds one
 * system per call (integrals.py:172, :307). Arrays hold one entry per engine,
 * in the order given to pnjk_group_create; NULL output arrays skip that output.
 * Outputs must be device memory. use_graph = 1 captures the launch sequence
 * in a hipGraph on the first call (and again whenever any pointer, size, tile
 * or precision changes) and replays it afterwards; it needs device-resident C
 * and H. The launch plan (problem structs and tile tables on the device) is
 * kept across calls whose engines, outputs, sizes, precision and tiles are
 * unchanged; new C or H arrays reuse it. Results are bitwise those of
 * pnjk_jk_build on each engine alone. split_refine, and builds without J and K in which
 * an instance takes pnjk_jk_build's triangular MO route, are not grouped: those calls
 * build the instances one after another on the stream, with no graph. An engine must not be used elsewhere while its group builds. */
typedef struct pnjk_group pnjk_group;
int pnjk_group_create(pnjk_engine* const* engines, int count, pnjk_group** out);
int pnjk_group_destroy(pnjk_group* g);
int pnjk_group_build(pnjk_group* g, const double* const* C, const int64_t* ldc, const int64_t* nbf5,
                     const double* const* H, const int64_t* ldh, int precision, uint32_t outputs,
                     double* const* J, double* const* K, double* const* J_MO, double* const* K_MO,
                     double* const* H_core, int use_graph, void* stream);

/* Per-stage timing with HIP events recorded on the build stream (no
 * counterpart in the reference, which has wall-clock prints only,
 * energy.py:29, 223-224). When on, every build records one event per stage;
 * pnjk_stage_times writes the last profiled build's stage durations in
 * milliseconds to ms[0 .. pnjk_stage_count()-1], waiting for that build. */
int pnjk_set_profiling(pnjk_engine* e, int on);
int pnjk_stage_count(void);
const char* pnjk_stage_name(int i);
int pnjk_stage_times(pnjk_engine* e, double* ms);

/* Tile shape override for the kernels of one arithmetic (0: FP64 MFMA, 1: BF16x3)
 * and one stage (index as in pnjk_stage_name: 1 half_transform, 2
 * quarter_transform, 5 coulomb_jn, 6 exchange_kn), used by the autotuner;
 * BM = 32*fm rows, BN = 32*fn columns per workgroup. fm = fn = 0 restores the
 * built-in choice. No counterpart in the reference. A tile shape changes no
 * result bit: each output's summation order is fixed by the MFMA sequence. */
int pnjk_set_tile(pnjk_engine* e, int arith, int stage, int fm, int fn);

/* Device bytes the engine holds now (b plus workspace). */
uint64_t pnjk_footprint(const pnjk_engine* e);

/* Device bytes a build of this size needs beyond b (workspace), for admission
 * control. */
uint64_t pnjk_workspace_bytes(int64_t nbf, int64_t naux, int64_t nbf5, int precision);

#ifdef __cplusplus
}
#endif

#endif /* PYNOF_AMD_H */
