// The C ABI of the MI355X RI JK build: engine state, b upload, build orchestration.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "jk_common.h"
#include "jk_kernels.h"
#include "pynof_amd.h"

// Workspace of pnjk_jk_build_noalloc: one array per GPU, held by the process from the
// library's first use of that GPU to its end. Every no-allocation build on a GPU works in
// this one array, so such builds must not overlap in time on one GPU (they are ordered when
// issued on one stream); its contents persist between calls and are never assumed zero.
// Host code must not pass `scratch` itself to a kernel or a copy (on the host the name is a
// stand-in, not the device address): scratch_ptr() gives each GPU's address, taken once per
// GPU with hipGetSymbolAddress in pnjk_create.
static constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace pnjk {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {

inline int64_t roundup(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// Device address of `scratch` per GPU, taken once per GPU (the first pnjk_create on it).
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};

// With `device` current: takes the address of this GPU's scratch array if not done yet.
int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return set_error(PNJK_EINVAL, "scratch: device id out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return PNJK_OK;
  void* p = nullptr;
  const hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (he != hipSuccess || !p)
    return set_error(PNJK_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(he));
  g_scratch[device] = static_cast<char*>(p);
  return PNJK_OK;
}

char* scratch_ptr(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want) {
    if (want <= bytes) return PNJK_OK;
    if (p) {
      (void)hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    if (want == 0) return PNJK_OK;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return set_error(PNJK_ENOMEM, "hipMalloc(" + std::to_string(want) + " bytes) failed: " + hipGetErrorString(e));
    }
    bytes = want;
    return PNJK_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

__global__ void k_fill(double* p, int64_t n, double v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Graph-safe clear (a kernel, not a memset node).
hipError_t fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

__global__ void k_copy2d(const double* __restrict__ src, int64_t lds, double* __restrict__ dst, int64_t ldd,
                         int rows, int cols) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)rows * cols) return;
  const int r = (int)(i / cols), c = (int)(i % cols);
  dst[(int64_t)r * ldd + c] = src[(int64_t)r * lds + c];
}

bool is_device_ptr(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Quarter-transform tile for the window of h = nbf5 / 2 + 1 rows (rules read off the tuner's
// timings at nbf5 = 160 ... 1000 on MI355X). FP64: the row tile (160, 128 or 64) that pads
// the window least, ties to 128, then 64 rows. BF16x3 (copy-bound, so larger tiles pay): the
// largest row tile whose rows are at least 75% live.
TileF64 qt_tile(int arith, int nbf5) {
  const int h = nbf5 / 2 + 1;
  auto padded = [&](int bm) { return (h + bm - 1) / bm * bm; };
  if (arith == 0) {
    TileF64 best{4, 2};
    int rows = padded(128);
    if (padded(64) < rows) best = TileF64{2, 4}, rows = padded(64);
    if (padded(160) < rows) best = TileF64{5, 2};
    return best;
  }
  for (int fm : {5, 4}) {
    if (4 * h >= 3 * padded(32 * fm)) return TileF64{fm, 4};
  }
  return TileF64{2, 4};
}

// Built-in tile shapes (from autotuning at nbf=1000, naux=3600, nbf5=300 on MI355X;
// small problems keep 64 x 64 tiles so the grid fills the chip).
TileF64 default_tile(int arith, int stage_kind, int nbf, int nbf5) {
  // stage_kind: 0 HT, 1 QT, 2 JN, 3 K
  const bool big = nbf5 > 128 && nbf >= 256;
  if (stage_kind == 1 && nbf5 > 128) return qt_tile(arith, nbf5);
  if (arith == 0) {
    switch (stage_kind) {
      case 0: return big ? TileF64{5, 4} : TileF64{2, 2};
      case 1: return TileF64{2, 2};
      case 2: return big ? TileF64{5, 4} : TileF64{2, 2};
      default: return big ? TileF64{4, 4} : TileF64{2, 2};
    }
  }
  switch (stage_kind) {
    case 0: return big ? TileF64{5, 5} : TileF64{2, 2};
    case 1: return TileF64{2, 2};
    case 2: return big ? TileF64{5, 4} : TileF64{2, 2};
    default: return nbf >= 256 ? TileF64{4, 4} : TileF64{2, 2};
  }
}

}  // namespace
}  // namespace pnjk

using namespace pnjk;

// Stages of one build, timed with events on the build stream when profiling is on.
enum Stage { ST_IN = 0, ST_HT, ST_QT, ST_KMO, ST_JMO, ST_JN, ST_K, ST_HC, ST_OUT, ST_COUNT };
static const char* kStageNames[ST_COUNT] = {"copy_in", "half_transform", "quarter_transform", "kmo_reduce",
                                            "jmo_syrk", "coulomb_jn", "exchange_kn", "hcore", "copy_out"};

struct pnjk_engine {
  int device = 0;
  hipStream_t own = nullptr;
  // Side stream of a single build (default priority) and its fork/join events: the
  // latency-bound J_MO and H_core steps overlap the MFMA-bound J^n GEMM.
  hipStream_t aux = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  std::mutex mu;
  bool prof = false;
  bool prof_recorded = false;
  hipEvent_t ev[ST_COUNT + 1] = {};
  int nbf = 0, naux = 0, naux_p = 0;
  int64_t ntri = 0;
  DevBuf btri;
  DevBuf bt16h, bt16l, bht16h, bht16l;  // split planes of b, built on the first split build
  DevBuf bdh;                            // b's diagonal rows halved, then one zero row (with the planes)
  DevBuf Wmo;                            // split_refine: W = C^T L C (nbf5^2 rows of naux_p)
  DevBuf Cp, Hp, X, d, kpart, T, outs, stage, scal;
  DevBuf Cw;                                     // window copies of C for the FP64 quarter transform
  DevBuf ct16h, ct16l, x16h, x16l, d16h, d16l;  // split workspace
  DevBuf Cf, ctf16h, ctf16l, hc;                 // MO integrals: all columns of C; H_core scratch
  DevBuf bmo, hm, lagw, lagg, lagp, lagc;        // Lagrangian workspace: b_MO, H_mat, W, G, partials, n/C^J/C^K
  DevBuf bm16h, bm16l, w16h, w16l;               // split form of the Lagrangian's exchange operands
  DevBuf fjk, fh, fco, ff;                       // Fock: staged J/K, H, coefficients, F
  DevBuf occ;                                    // occupation step: every array of the step
  rocblas_handle blas = nullptr;                 // for the metric eigensolver (b from 3-centre integrals)
  bool split_ready = false;
  bool hd_ready = false;  // bdh holds the halved diagonal rows of the resident b
  // Tile override per arithmetic (0: FP64 kernels, 1: BF16x3 kernels) and stage;
  // {0, 0} selects the built-in choice.
  TileF64 tiles[2][ST_COUNT] = {};
  DevBuf* all[44];
  int nall = 0;
  pnjk_engine() {
    for (DevBuf* b : {&btri, &bt16h, &bt16l, &bht16h, &bht16l, &Cp, &Hp, &X, &d, &kpart, &T, &outs, &stage, &scal,
                      &ct16h, &ct16l, &x16h, &x16l, &d16h, &d16l, &Cf, &ctf16h, &ctf16l, &hc,
                      &bmo, &hm, &lagw, &lagg, &lagp, &lagc, &bm16h, &bm16l, &w16h, &w16l, &fjk, &fh, &fco, &ff, &occ, &Cw,
                      &bdh, &Wmo})
      all[nall++] = b;
  }
};

extern "C" {

const char* pnjk_last_error(void) { return g_last_error.c_str(); }

const char* pnjk_version(void) { return "pynof_amd 0.1 (gfx950)"; }

int pnjk_create(int device, pnjk_engine** out) {
  if (!out) return set_error(PNJK_EINVAL, "pnjk_create: out is NULL");
  *out = nullptr;
  int n = 0;
  PNJK_HIP_CHECK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return set_error(PNJK_EINVAL, "pnjk_create: device " + std::to_string(device) + " of " + std::to_string(n));
  PNJK_HIP_CHECK(hipSetDevice(device));
  int rc = scratch_init(device);
  if (rc) return rc;
  pnjk_engine* e = new pnjk_engine();
  e->device = device;
  hipError_t he = hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipStreamCreateWithFlags(&e->aux, hipStreamNonBlocking);
  for (hipEvent_t* ev : {&e->fork_ev, &e->join_ev})
    if (he == hipSuccess) he = hipEventCreateWithFlags(ev, hipEventDisableTiming);
  if (he != hipSuccess) {
    for (hipEvent_t ev : {e->fork_ev, e->join_ev})
      if (ev) (void)hipEventDestroy(ev);
    if (e->aux) (void)hipStreamDestroy(e->aux);
    if (e->own) (void)hipStreamDestroy(e->own);
    delete e;
    return set_error(PNJK_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  }
  *out = e;
  return PNJK_OK;
}

int pnjk_destroy(pnjk_engine* e) {
  if (!e) return PNJK_OK;
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->own);
  if (e->blas) (void)rocblas_destroy_handle(e->blas);
  for (int i = 0; i < e->nall; ++i) e->all[i]->release();
  for (hipEvent_t& ev : e->ev)
    if (ev) (void)hipEventDestroy(ev);
  (void)hipStreamSynchronize(e->aux);
  for (hipEvent_t ev : {e->fork_ev, e->join_ev})
    if (ev) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(e->aux);
  (void)hipStreamDestroy(e->own);
  delete e;
  return PNJK_OK;
}

int64_t pnjk_packed_ld(int64_t nbf, int64_t naux) {
  (void)nbf;
  return roundup(naux, 64);
}

uint64_t pnjk_footprint(const pnjk_engine* e) {
  if (!e) return 0;
  uint64_t s = 0;
  for (int i = 0; i < e->nall; ++i) s += e->all[i]->bytes;
  return s;
}

// Whether a build without J and K forms its MO integrals by the triangular route (T = L C,
// W = C^T T). It needs fewer flops whenever nbf5 < nbf (nbf^2 nbf5 + 2 nbf nbf5^2 against
// 2 nbf5 nbf^2 + nbf5^2 nbf), but its triangular GEMM (batches of K = mu + 1) runs below the
// half-transform's rate, so it is the faster route only up to about nbf5 / nbf = 2/3 in FP64
// and 2/5 in BF16x3 (measured over nbf = 200..1000; beyond, up to 1.4x slower).
static bool mo_tri_route(int precision, int64_t nbf, int64_t nbf5) {
  if (precision == PNJK_FP64) return 3 * nbf5 <= 2 * nbf;
  if (precision == PNJK_SPLIT) return 5 * nbf5 <= 2 * nbf;
  return false;
}

uint64_t pnjk_workspace_bytes(int64_t nbf, int64_t naux, int64_t nbf5, int precision) {
  const int64_t naux_p = pnjk_packed_ld(nbf, naux);
  const int64_t nbf_p = roundup(nbf, 32), ldcp = roundup(nbf5, 32);
  uint64_t s = 0;
  s += (uint64_t)((nbf_p + 8) * ldcp + 1024) * 8;           // Cp
  s += (uint64_t)(nbf * nbf_p) * 8;                          // Hp
  s += (uint64_t)(nbf5 * nbf * naux_p + 64) * 8;             // X (+ slack)
  s += (uint64_t)(roundup(nbf5, 160) * naux_p) * 8;          // d
  s += (uint64_t)(nbf5 * nbf5 * (naux_p / 64)) * 8;          // kpart
  s += (uint64_t)(nbf * ldcp) * 8;                           // T
  if (precision != PNJK_SPLIT) s += (uint64_t)(2 * nbf_p * qt_window_cols((int)nbf5)) * 8;  // Cw
  // W: split_refine, and FP64 builds without J and K that take the triangular route (this
  // total bounds both routes of an FP64 build)
  const bool tri_fp64 = precision == PNJK_FP64 && mo_tri_route(precision, nbf, nbf5);
  if (precision == PNJK_SPLIT_REFINE || tri_fp64) s += (uint64_t)(nbf5 * nbf5 * naux_p) * 8;
  if (tri_fp64) s += (uint64_t)((nbf + 1) * naux_p) * 8;   // halved diagonal rows of b and a zero row
  if (precision != PNJK_FP64) {
    const int64_t ntri = nbf * (nbf + 1) / 2;
    s += (uint64_t)((roundup(ntri, 16) + roundup(nbf + 1, 16)) * naux_p) * 4;  // bt16 hi/lo + halved diagonal rows
    if (mo_tri_route(precision, nbf, nbf5)) s += (uint64_t)(nbf5 * nbf5 * naux_p) * 4;  // FP32 W (triangular route)
    s += (uint64_t)(roundup(nbf, 16) * naux_p * nbf_p) * 4;  // bht16 hi/lo (rows nu < nbf rounded up to 16)
    s += (uint64_t)(roundup(nbf5, 16) * nbf_p) * 4;          // Ct16 hi/lo (whole 16-row blocks)
    s += (uint64_t)(nbf5 * x16_stride((int)nbf, (int)naux_p) + kX16Slack) * 4;  // X16 hi/lo (+ slack)
    s += (uint64_t)(roundup(nbf5, 16) * naux_p) * 4;         // d16 hi/lo (whole 16-row blocks)
    s += (uint64_t)((nbf + 1) * naux_p) * 8;                 // halved diagonal rows of b and a zero row
  }
  return s;
}

// The halved diagonal rows of b and a zero row, which the triangular half-transform reads
// (once per uploaded b).
static int ensure_half_diag(pnjk_engine* e, hipStream_t s) {
  if (e->hd_ready) return PNJK_OK;
  const int64_t naux_p = e->naux_p, nbf = e->nbf;
  int rc;
  if ((rc = e->bdh.ensure((size_t)((nbf + 1) * naux_p) * 8))) return rc;
  PNJK_HIP_CHECK(launch_half_diag(e->btri.as<double>(), (int)nbf, (int)naux_p, e->bdh.as<double>(),
                                  e->bdh.as<double>() + nbf * naux_p, s));
  e->hd_ready = true;
  return PNJK_OK;
}

// Builds the bf16 planes of b the split kernels read, and the halved diagonal rows the
// refinement's triangular half-transform reads (once per uploaded b).
static int ensure_split_planes(pnjk_engine* e, hipStream_t s) {
  if (e->split_ready) return PNJK_OK;
  const int64_t ntri = e->ntri, naux_p = e->naux_p, nbf = e->nbf, ldk = roundup(nbf, 32);
  int rc;
  // bt16 in whole 16-row blocks (blk_index): the packed rows, then from row roundup(ntri, 16)
  // the halved diagonal rows and a zero row (the split triangular route's L diagonal)
  const int64_t hd = roundup(ntri, 16);
  const size_t tri_bytes = (size_t)((hd + roundup(nbf + 1, 16)) * naux_p) * 2,
               ht_bytes = (size_t)(roundup(nbf, 16) * naux_p * ldk) * 2;
  if ((rc = e->bt16h.ensure(tri_bytes)) || (rc = e->bt16l.ensure(tri_bytes))) return rc;
  if ((rc = e->bht16h.ensure(ht_bytes)) || (rc = e->bht16l.ensure(ht_bytes))) return rc;
  PNJK_HIP_CHECK(launch_split_blocked(e->btri.as<double>(), ntri, naux_p, e->bt16h.as<uint16_t>(),
                                      e->bt16l.as<uint16_t>(), s));
  if ((rc = ensure_half_diag(e, s))) return rc;
  PNJK_HIP_CHECK(launch_split_blocked(e->bdh.as<double>(), nbf + 1, naux_p, e->bt16h.as<uint16_t>() + hd * naux_p,
                                      e->bt16l.as<uint16_t>() + hd * naux_p, s));
  PNJK_HIP_CHECK(launch_build_bht(e->btri.as<double>(), (int)nbf, (int)naux_p, (int)ldk, e->bht16h.as<uint16_t>(),
                                  e->bht16l.as<uint16_t>(), s));
  e->split_ready = true;
  return PNJK_OK;
}

static int upload_common(pnjk_engine* e, int64_t nbf, int64_t naux) {
  e->split_ready = false;
  e->hd_ready = false;
  if (nbf < 1 || naux < 1) return set_error(PNJK_EINVAL, "upload_b: nbf and naux must be >= 1");
  if (nbf > 46000) return set_error(PNJK_EINVAL, "upload_b: nbf too large");
  e->nbf = (int)nbf;
  e->naux = (int)naux;
  e->naux_p = (int)pnjk_packed_ld(nbf, naux);
  e->ntri = nbf * (nbf + 1) / 2;
  if (e->ntri * (int64_t)e->naux_p > (int64_t)1 << 40) return set_error(PNJK_EINVAL, "upload_b: b too large");
  return e->btri.ensure((size_t)(e->ntri * e->naux_p) * sizeof(double));
}

}  // extern "C"

// pnjk_upload_b with the engine's mutex held.
static int upload_b_locked(pnjk_engine* e, const double* b, int64_t nbf, int64_t naux, double sym_tol,
                           double* asym) {
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  e->nbf = 0;
  int rc = upload_common(e, nbf, naux);
  if (rc) return rc;
  if ((rc = e->scal.ensure(256))) return rc;
  hipStream_t s = e->own;
  double* amax = e->scal.as<double>();
  unsigned long long* bad = reinterpret_cast<unsigned long long*>(amax + 1);
  PNJK_HIP_CHECK(hipMemsetAsync(e->scal.p, 0, 16, s));
  const bool on_dev = is_device_ptr(b);
  const int64_t row_bytes = nbf * naux * (int64_t)sizeof(double);  // one mu row of b
  if (on_dev) {
    PNJK_HIP_CHECK(launch_pack_b((int)nbf, (int)naux, e->naux_p, 0, (int)nbf, b, e->btri.as<double>(), amax, bad, s));
  } else {
    // Host b: stream mu-row chunks (~512 MiB) through a staging buffer, highest mu first.
    const int64_t rows = std::max<int64_t>(1, std::min<int64_t>(nbf, ((int64_t)512 << 20) / row_bytes));
    if ((rc = e->stage.ensure((size_t)(rows * row_bytes)))) return rc;
    for (int64_t mu1 = nbf; mu1 > 0; mu1 -= rows) {
      const int64_t mu0 = std::max<int64_t>(0, mu1 - rows);
      PNJK_HIP_CHECK(hipMemcpyAsync(e->stage.p, b + mu0 * nbf * naux, (size_t)((mu1 - mu0) * row_bytes),
                                    hipMemcpyHostToDevice, s));
      PNJK_HIP_CHECK(launch_pack_b((int)nbf, (int)naux, e->naux_p, (int)mu0, (int)mu1, e->stage.as<double>(),
                                   e->btri.as<double>(), amax, bad, s));
    }
  }
  double host[2];
  PNJK_HIP_CHECK(hipMemcpyAsync(host, e->scal.p, 16, hipMemcpyDeviceToHost, s));
  PNJK_HIP_CHECK(hipStreamSynchronize(s));
  e->stage.release();
  unsigned long long nbad;
  std::memcpy(&nbad, &host[1], sizeof(nbad));
  if (asym) *asym = host[0];
  if (nbad) {
    e->nbf = 0;
    return set_error(PNJK_EINVAL, "upload_b: b holds " + std::to_string(nbad) + " non-finite values");
  }
  if (!(host[0] <= sym_tol)) {
    e->nbf = 0;
    return set_error(PNJK_EINVAL, "upload_b: b is not symmetric in (mu, nu): max |b - b^T| = " +
                                      std::to_string(host[0]) + " > tolerance " + std::to_string(sym_tol));
  }
  return PNJK_OK;
}

extern "C" {

int pnjk_upload_b(pnjk_engine* e, const double* b, int64_t nbf, int64_t naux, double sym_tol, uint64_t* footprint,
                  double* asym) {
  if (!e || !b) return set_error(PNJK_EINVAL, "pnjk_upload_b: null argument");
  std::lock_guard<std::mutex> lk(e->mu);
  const int rc = upload_b_locked(e, b, nbf, naux, sym_tol, asym);
  if (rc) return rc;
  if (footprint) *footprint = pnjk_footprint(e);
  return PNJK_OK;
}

int pnjk_download_b(pnjk_engine* e, double* b_mnl) {
  if (!e || !b_mnl) return set_error(PNJK_EINVAL, "pnjk_download_b: null argument");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  if (e->nbf == 0) return set_error(PNJK_ESTATE, "pnjk_download_b: no b uploaded");
  hipStream_t s = e->own;
  const int64_t nbf = e->nbf, naux = e->naux, row_bytes = nbf * naux * 8;
  if (is_device_ptr(b_mnl)) {
    PNJK_HIP_CHECK(launch_unpack_b((int)nbf, (int)naux, e->naux_p, 0, (int)nbf, e->btri.as<double>(), b_mnl, s));
  } else {  // stream mu-row chunks (~512 MiB) through a staging buffer
    const int64_t rows = std::max<int64_t>(1, std::min<int64_t>(nbf, ((int64_t)512 << 20) / row_bytes));
    int rc;
    if ((rc = e->stage.ensure((size_t)(rows * row_bytes)))) return rc;
    for (int64_t mu0 = 0; mu0 < nbf; mu0 += rows) {
      const int64_t mu1 = std::min(nbf, mu0 + rows);
      PNJK_HIP_CHECK(launch_unpack_b((int)nbf, (int)naux, e->naux_p, (int)mu0, (int)mu1, e->btri.as<double>(),
                                     e->stage.as<double>(), s));
      PNJK_HIP_CHECK(hipMemcpyAsync(b_mnl + mu0 * nbf * naux, e->stage.p, (size_t)((mu1 - mu0) * row_bytes),
                                    hipMemcpyDeviceToHost, s));
      PNJK_HIP_CHECK(hipStreamSynchronize(s));
    }
    e->stage.release();
  }
  PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
}

int pnjk_upload_b_3c(pnjk_engine* e, const double* mnk, int64_t nbf, int64_t naux, const double* metric,
                     double cutoff, double sym_tol, uint64_t* footprint, double* asym) {
  if (!e || !mnk || !metric) return set_error(PNJK_EINVAL, "pnjk_upload_b_3c: null argument");
  if (!(cutoff >= 0.0)) return set_error(PNJK_EINVAL, "pnjk_upload_b_3c: cutoff must be >= 0");
  std::lock_guard<std::mutex> lk(e->mu);
  // The packed 3-centre integrals (mu nu|P), checked like b, land where b will live.
  int rc = upload_b_locked(e, mnk, nbf, naux, sym_tol, asym);
  if (rc) return rc;
  // Not a b until the transform below completes: any early return leaves no b uploaded.
  const int nbf_ready = e->nbf;
  e->nbf = 0;
  hipStream_t s = e->own;
  const int n = (int)naux, ld = e->naux_p;
  const int64_t ldld = (int64_t)ld * ld;
  DevBuf M, D, E, info, Vp, Vs, W, out;
  struct Release {  // frees the local buffers on every return
    DevBuf* b[8];
    ~Release() {
      for (DevBuf* x : b) x->release();
    }
  } release_all{{&M, &D, &E, &info, &Vp, &Vs, &W, &out}};
  if ((rc = M.ensure((size_t)naux * naux * 8)) || (rc = D.ensure((size_t)naux * 8)) ||
      (rc = E.ensure((size_t)naux * 8)) || (rc = info.ensure(64)) || (rc = Vp.ensure((size_t)ldld * 8)) ||
      (rc = Vs.ensure((size_t)ldld * 8)) || (rc = W.ensure((size_t)(ldld + 128) * 8)) ||
      (rc = out.ensure((size_t)(e->ntri * ld) * 8))) {
    e->nbf = 0;
    return rc;
  }
  auto fail = [&](int code, const std::string& msg) {
    e->nbf = 0;
    return set_error(code, msg);
  };
  if (hipMemcpyAsync(M.p, metric, (size_t)naux * naux * 8,
                     is_device_ptr(metric) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s) != hipSuccess)
    return fail(PNJK_EHIP, "pnjk_upload_b_3c: metric copy failed");
  // (P|Q)^-1/2 as psi4's Matrix::power(-0.5, cutoff) forms it (integrals.py:41-42):
  // eigenpairs, the relative eigenvalue cut, V diag(a^-1/2) V^T.
  if (!e->blas) {
    if (rocblas_create_handle(&e->blas) != rocblas_status_success) {
      e->blas = nullptr;
      return fail(PNJK_EHIP, "pnjk_upload_b_3c: rocblas_create_handle failed");
    }
  }
  if (rocblas_set_stream(e->blas, s) != rocblas_status_success)
    return fail(PNJK_EHIP, "pnjk_upload_b_3c: rocblas_set_stream failed");
  if (rocsolver_dsyevd(e->blas, rocblas_evect_original, rocblas_fill_lower, n, M.as<double>(), n, D.as<double>(),
                       E.as<double>(), reinterpret_cast<rocblas_int*>(info.p)) != rocblas_status_success)
    return fail(PNJK_EHIP, "pnjk_upload_b_3c: rocsolver_dsyevd failed");
  rocblas_int hinfo = 0;
  PNJK_HIP_CHECK(hipMemcpyAsync(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, s));
  PNJK_HIP_CHECK(hipStreamSynchronize(s));
  if (hinfo != 0)
    return fail(PNJK_EINVAL, "pnjk_upload_b_3c: the metric eigensolver did not converge (info " +
                                 std::to_string(hinfo) + ")");
  PNJK_HIP_CHECK(launch_metric_scale(D.as<double>(), M.as<double>(), n, ld, cutoff, Vp.as<double>(), Vs.as<double>(),
                                     s));
  PNJK_HIP_CHECK(fill(W.as<double>() + ldld, 128, 0.0, s));
  PNJK_HIP_CHECK(launch_atb_f64(Vp.as<double>(), Vs.as<double>(), ld, W.as<double>(), s));
  // b[t][P] = sum_Q (mu nu|Q) W[Q][P] over the packed rows (integrals.py:49)
  PNJK_HIP_CHECK(launch_rowgemm_f64(e->btri.as<double>(), e->ntri, W.as<double>(), ld, out.as<double>(), s));
  PNJK_HIP_CHECK(hipStreamSynchronize(s));
  std::swap(e->btri.p, out.p);  // `out` now holds the 3-centre integrals and is freed on return
  std::swap(e->btri.bytes, out.bytes);
  e->split_ready = false;
  e->hd_ready = false;
  e->nbf = nbf_ready;
  if (footprint) *footprint = pnjk_footprint(e);
  return PNJK_OK;
}

int pnjk_upload_b_packed(pnjk_engine* e, const double* btri, int64_t nbf, int64_t naux, int64_t ld,
                         uint64_t* footprint) {
  if (!e || !btri) return set_error(PNJK_EINVAL, "pnjk_upload_b_packed: null argument");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  if (ld != pnjk_packed_ld(nbf, naux)) return set_error(PNJK_EINVAL, "upload_b_packed: ld must be pnjk_packed_ld");
  if (!is_device_ptr(btri)) return set_error(PNJK_EINVAL, "upload_b_packed: btri must be device memory");
  e->nbf = 0;
  int rc = upload_common(e, nbf, naux);
  if (rc) return rc;
  PNJK_HIP_CHECK(hipMemcpyAsync(e->btri.p, btri, (size_t)(e->ntri * e->naux_p) * sizeof(double),
                                hipMemcpyDeviceToDevice, e->own));
  PNJK_HIP_CHECK(hipStreamSynchronize(e->own));
  if (footprint) *footprint = pnjk_footprint(e);
  return PNJK_OK;
}

}  // extern "C"

namespace {

// Arguments of one build, as given at the C ABI.
struct BuildArgs {
  const double* C;
  int64_t ldc, nbf5;
  const double* H;
  int64_t ldh;
  int precision;
  uint32_t outputs;
  double *J, *K, *J_MO, *K_MO, *H_core;
};

// Per-build state between stages: geometry, tiles, device output pointers.
struct BuildState {
  BuildArgs a;
  BuildGeom g;
  SplitGeom sg;
  TileF64 ht, qt, jt, kt;
  TileF64 ht16;  // split_refine: the BF16x3 half-transform (X16 for K^n) beside the FP64 MO path
  bool want_j, want_k, want_jmo, want_kmo, want_h;
  bool hJ, hK, hJMO, hKMO, hH;  // output goes to host memory (staged)
  bool c_dev = false, h_dev = false;  // C / H are device memory
  double *dJ, *dK, *dJMO, *dKMO, *dH;
  int nt = 0;  // P-tiles of the quarter transform
  // Device workspace of this build: the engine's grow-only buffers, or pieces of the
  // library's static scratch array (pnjk_jk_build_noalloc).
  double *Cp = nullptr, *Cw = nullptr, *X = nullptr, *d = nullptr, *kpart = nullptr, *Hp = nullptr, *T = nullptr;
  uint16_t *ct16h = nullptr, *ct16l = nullptr, *x16h = nullptr, *x16l = nullptr, *d16h = nullptr, *d16l = nullptr;
  double* W = nullptr;  // triangular route: C^T L C
  // J_MO, K_MO by the triangular route (T = L C, W = C^T T, Y_pq = W_pq + W_qp; the
  // refinement's FP64 MO path) instead of the half- and quarter transforms.
  bool tri = false;
  // The split form's triangular MO route (builds without J and K, mo_tri_route): T16 = L C
  // and W = C^T T16 in BF16x3, W in FP32, K_MO and d reduced from it.
  bool tri16 = false;
  // Quarter-transform window rows: 0 for the whole window; 1 when neither J_MO nor K_MO is
  // requested and only d (window row 0, for J^n) is needed, with the bits of the whole
  // window's d.
  int qt_rows = 0;
};

// The workspace pieces of one build, in bytes (0: not needed), in a fixed order.
enum WsPiece { W_CP = 0, W_CW, W_X, W_D, W_KPART, W_HP, W_T, W_CT16H, W_CT16L, W_X16H, W_X16L, W_D16H, W_D16L,
               W_STAGE, W_W, W_COUNT };
struct WsPlan {
  size_t bytes[W_COUNT] = {};
};

// A stage refuses to enqueue work on a workspace pointer its build did not prepare (a
// kernel would otherwise write through a null address).
#define PNJK_NEED(ptr)                                                                               \
  do {                                                                                               \
    if (!(ptr)) return set_error(PNJK_ESTATE, "internal: workspace " #ptr " was not prepared"); \
  } while (0)

// Validates the arguments, sizes the workspace, resolves output staging and
// tiles. Enqueues nothing. With `scratch` (scratch_bytes long) the workspace and the
// staging of unrequested J_MO/K_MO are carved from it and nothing is allocated: every
// output must then be device memory (host outputs would need staging buffers).
int build_prepare(pnjk_engine* e, const BuildArgs& a, BuildState& st, char* scratch = nullptr,
                  size_t scratch_bytes = 0, bool tri_routes = false) {
  if (e->nbf == 0) return set_error(PNJK_ESTATE, "pnjk_jk_build: no b uploaded");
  if (!a.C) return set_error(PNJK_EINVAL, "pnjk_jk_build: C is NULL");
  if (a.nbf5 < 1 || a.nbf5 > e->nbf)
    return set_error(PNJK_EINVAL, "pnjk_jk_build: nbf5 must be in [1, nbf]; got " + std::to_string(a.nbf5));
  if (a.ldc < a.nbf5) return set_error(PNJK_EINVAL, "pnjk_jk_build: ldc < nbf5");
  if (a.outputs & ~PNJK_OUT_ALL) return set_error(PNJK_EINVAL, "pnjk_jk_build: unknown output bits");
  if ((a.outputs & PNJK_OUT_J) && !a.J) return set_error(PNJK_EINVAL, "pnjk_jk_build: J requested but NULL");
  if ((a.outputs & PNJK_OUT_K) && !a.K) return set_error(PNJK_EINVAL, "pnjk_jk_build: K requested but NULL");
  if ((a.outputs & PNJK_OUT_JMO) && !a.J_MO)
    return set_error(PNJK_EINVAL, "pnjk_jk_build: J_MO requested but NULL");
  if ((a.outputs & PNJK_OUT_KMO) && !a.K_MO)
    return set_error(PNJK_EINVAL, "pnjk_jk_build: K_MO requested but NULL");
  if ((a.outputs & PNJK_OUT_HCORE) && (!a.H_core || !a.H))
    return set_error(PNJK_EINVAL, "pnjk_jk_build: H_core requested but H or H_core is NULL");
  if ((a.outputs & PNJK_OUT_HCORE) && a.ldh < e->nbf) return set_error(PNJK_EINVAL, "pnjk_jk_build: ldh < nbf");
  if (a.precision != PNJK_FP64 && a.precision != PNJK_SPLIT && a.precision != PNJK_SPLIT_REFINE)
    return set_error(PNJK_EUNSUPPORTED, "pnjk_jk_build: unknown precision " + std::to_string(a.precision));
  st.a = a;
  st.c_dev = is_device_ptr(a.C);
  st.h_dev = a.H && is_device_ptr(a.H);
  BuildGeom& g = st.g;
  g.nbf = e->nbf;
  g.naux = e->naux;
  g.naux_p = e->naux_p;
  g.nbf5 = (int)a.nbf5;
  g.ntri = e->ntri;
  g.nbf_p = (int)roundup(e->nbf, 32);
  g.ldcp = (int)roundup(a.nbf5, 32);
  st.sg.nbf = g.nbf;
  st.sg.naux_p = g.naux_p;
  st.sg.nbf5 = g.nbf5;
  st.sg.ldk = g.nbf_p;
  const int64_t nbf = g.nbf, n5 = a.nbf5;
  const int precision = a.precision;

  int rc;
  st.want_j = a.outputs & PNJK_OUT_J;
  st.want_k = a.outputs & PNJK_OUT_K;
  st.want_jmo = a.outputs & PNJK_OUT_JMO;
  st.want_kmo = a.outputs & PNJK_OUT_KMO;
  st.want_h = a.outputs & PNJK_OUT_HCORE;
  // The triangular route (mo_tri_route): split_refine always takes it; an FP64 or split build
  // takes it (in FP64 or BF16x3) when its caller allows it, asks for neither J nor K (which
  // need the whole half-transformed X) and nbf5 / nbf is small enough for it to be faster.
  const bool mo_only = tri_routes && !st.want_j && !st.want_k && mo_tri_route(precision, nbf, n5);
  // split_refine's FP64 MO integrals follow the FP64 rule; past it the FP64 half- and quarter
  // transforms run, and X16 for the split K^n is the split of their FP64 X
  st.tri = (precision == PNJK_SPLIT_REFINE && mo_tri_route(PNJK_FP64, nbf, n5)) ||
           (mo_only && precision == PNJK_FP64);
  st.tri16 = mo_only && precision == PNJK_SPLIT;
  st.qt_rows = (st.want_jmo || st.want_kmo) ? 0 : 1;
  const int64_t jk_elems = n5 * nbf * nbf;
  st.dJ = a.J;
  st.dK = a.K;
  st.dJMO = a.J_MO;
  st.dKMO = a.K_MO;
  st.dH = a.H_core;
  st.hJ = st.want_j && !is_device_ptr(a.J);
  st.hK = st.want_k && !is_device_ptr(a.K);
  st.hJMO = !st.want_jmo || !is_device_ptr(a.J_MO);  // staged when not requested (the triangular
  st.hKMO = !st.want_kmo || !is_device_ptr(a.K_MO);  // route's reduction writes K_MO regardless)
  st.hH = st.want_h && !is_device_ptr(a.H_core);
  if (scratch && (st.hJ || st.hK || (st.want_jmo && st.hJMO) || (st.want_kmo && st.hKMO) || st.hH))
    return set_error(PNJK_EINVAL, "pnjk_jk_build_noalloc: every output must be device memory");
  int64_t stage_elems = 0;
  const int64_t off_J = stage_elems;
  if (st.hJ) stage_elems += jk_elems;
  const int64_t off_K = stage_elems;
  if (st.hK) stage_elems += jk_elems;
  const int64_t off_JMO = stage_elems;
  if (st.hJMO) stage_elems += n5 * n5;
  const int64_t off_KMO = stage_elems;
  if (st.hKMO) stage_elems += n5 * n5;
  const int64_t off_H = stage_elems;
  if (st.hH) stage_elems += n5;

  WsPlan w;
  w.bytes[W_CP] = (size_t)((int64_t)(g.nbf_p + 8) * g.ldcp + 1024) * 8;
  if (precision != PNJK_SPLIT) {
    // X slack: a 128-column quarter-transform tile reads up to 64 columns past naux_p.
    // (split_refine keeps its triangular half-transform T here.)
    w.bytes[W_X] = (size_t)(n5 * nbf * g.naux_p + 64) * 8;
  }
  if (precision != PNJK_SPLIT && !st.tri) w.bytes[W_CW] = (size_t)(2 * g.nbf_p * (int64_t)qt_window_cols(g.nbf5)) * 8;
  if (st.tri) w.bytes[W_W] = (size_t)(n5 * n5 * g.naux_p) * 8;
  if (st.tri16) w.bytes[W_W] = (size_t)(n5 * n5 * g.naux_p) * 4;  // FP32 W
  w.bytes[W_D] = (size_t)(roundup(n5, 160) * g.naux_p) * 8;
  if (!st.tri && !st.tri16) w.bytes[W_KPART] = (size_t)(n5 * n5 * (g.naux_p / 64)) * 8;
  if (st.want_h) {
    w.bytes[W_HP] = (size_t)(nbf * g.nbf_p) * 8;
    w.bytes[W_T] = (size_t)(nbf * g.ldcp) * 8;
  }
  if (precision != PNJK_FP64) {
    w.bytes[W_CT16H] = w.bytes[W_CT16L] = (size_t)(roundup(n5, 16) * st.sg.ldk) * 2;  // whole 16-row blocks
    w.bytes[W_X16H] = w.bytes[W_X16L] = (size_t)(n5 * x16_stride(nbf, g.naux_p) + kX16Slack) * 2;
    w.bytes[W_D16H] = w.bytes[W_D16L] = (size_t)(roundup(n5, 16) * g.naux_p) * 2;  // whole 16-row blocks
  }
  w.bytes[W_STAGE] = (size_t)std::max<int64_t>(stage_elems, 1) * 8;
  void* ptr[W_COUNT] = {};
  if (scratch) {
    const size_t off = ws_carve(w.bytes, W_COUNT, nullptr, ptr);
    if (off > scratch_bytes)
      return set_error(PNJK_EINVAL, "pnjk_jk_build_noalloc: this build needs " + std::to_string(off) +
                                        " bytes of workspace; the library's scratch array holds " +
                                        std::to_string(scratch_bytes));
    ws_carve(w.bytes, W_COUNT, scratch, ptr);
  } else {
    DevBuf* bufs[W_COUNT] = {&e->Cp, &e->Cw, &e->X, &e->d, &e->kpart, &e->Hp, &e->T, &e->ct16h, &e->ct16l,
                             &e->x16h, &e->x16l, &e->d16h, &e->d16l, &e->outs, &e->Wmo};
    for (int i = 0; i < W_COUNT; ++i) {
      if (!w.bytes[i]) continue;
      if ((rc = bufs[i]->ensure(w.bytes[i]))) return rc;
      ptr[i] = bufs[i]->p;
    }
  }
  st.Cp = static_cast<double*>(ptr[W_CP]);
  st.Cw = static_cast<double*>(ptr[W_CW]);
  st.X = static_cast<double*>(ptr[W_X]);
  st.d = static_cast<double*>(ptr[W_D]);
  st.kpart = static_cast<double*>(ptr[W_KPART]);
  st.Hp = static_cast<double*>(ptr[W_HP]);
  st.T = static_cast<double*>(ptr[W_T]);
  st.ct16h = static_cast<uint16_t*>(ptr[W_CT16H]);
  st.ct16l = static_cast<uint16_t*>(ptr[W_CT16L]);
  st.x16h = static_cast<uint16_t*>(ptr[W_X16H]);
  st.x16l = static_cast<uint16_t*>(ptr[W_X16L]);
  st.d16h = static_cast<uint16_t*>(ptr[W_D16H]);
  st.d16l = static_cast<uint16_t*>(ptr[W_D16L]);
  st.W = static_cast<double*>(ptr[W_W]);
  double* ob = static_cast<double*>(ptr[W_STAGE]);
  if (st.hJ) st.dJ = ob + off_J;
  if (st.hK) st.dK = ob + off_K;
  if (st.hJMO) st.dJMO = ob + off_JMO;
  if (st.hKMO) st.dKMO = ob + off_KMO;
  if (st.hH) st.dH = ob + off_H;

  // FP64 kernels run HT/QT in the fp64 and split_refine forms, J^n/K^n in fp64 only.
  const int a_mo = precision == PNJK_SPLIT ? 1 : 0, a_jk = precision == PNJK_FP64 ? 0 : 1;
  auto chosen = [&](int arith, int stg, int kind) {
    return e->tiles[arith][stg].fm ? e->tiles[arith][stg] : default_tile(arith, kind, g.nbf, g.nbf5);
  };
  st.ht = chosen(a_mo, ST_HT, 0);
  st.qt = chosen(a_mo, ST_QT, 1);
  st.jt = chosen(a_jk, ST_JN, 2);
  st.kt = chosen(a_jk, ST_K, 3);
  st.ht16 = chosen(1, ST_HT, 0);
  return PNJK_OK;
}

// C (and H) into the padded device buffers; split planes of C^T.
int stage_in(pnjk_engine* e, BuildState& st, hipStream_t s) {
  PNJK_NEED(st.Cp);
  if (st.a.precision != PNJK_FP64) {
    int rc = ensure_split_planes(e, s);
    if (rc) return rc;
  } else if (st.tri) {
    int rc = ensure_half_diag(e, s);
    if (rc) return rc;
  }
  const BuildGeom& g = st.g;
  const int64_t nbf = g.nbf, n5 = g.nbf5;
  double* Cp = st.Cp;
  PNJK_HIP_CHECK(fill(Cp, (int64_t)(g.nbf_p + 8) * g.ldcp + 1024, 0.0, s));
  if (st.c_dev) {
    const int64_t n = nbf * n5;
    hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, st.a.C, st.a.ldc, Cp,
                       (int64_t)g.ldcp, (int)nbf, (int)n5);
    PNJK_HIP_CHECK(hipGetLastError());
  } else {
    PNJK_HIP_CHECK(hipMemcpy2DAsync(Cp, (size_t)g.ldcp * 8, st.a.C, (size_t)st.a.ldc * 8, (size_t)n5 * 8,
                                    (size_t)nbf, hipMemcpyHostToDevice, s));
  }
  if (st.a.precision != PNJK_FP64) {  // split: C^T planes; split_refine: for its split half-transform
    PNJK_NEED(st.ct16h);
    PNJK_NEED(st.ct16l);
    PNJK_HIP_CHECK(launch_split_ct(Cp, g.ldcp, g.nbf, g.nbf5, g.nbf5, st.sg.ldk, st.ct16h, st.ct16l, s));
  }
  if (st.a.precision != PNJK_SPLIT && !st.tri) {  // the FP64 quarter transform's window copies of C
    PNJK_NEED(st.Cw);
    PNJK_HIP_CHECK(launch_window_c(Cp, g.ldcp, g.nbf_p, g.nbf5, qt_window_cols(g.nbf5), st.Cw, s));
  }
  return PNJK_OK;
}

int stage_ht(pnjk_engine* e, BuildState& st, hipStream_t s, GroupCollector* col) {
  if (st.a.precision == PNJK_SPLIT) {
    PNJK_NEED(st.ct16h);
    PNJK_NEED(st.ct16l);
    PNJK_NEED(st.x16h);
    PNJK_NEED(st.x16l);
    PNJK_HIP_CHECK(launch_ht_x3(st.sg, st.ht, st.ct16h, st.ct16l, e->bht16h.as<uint16_t>(), e->bht16l.as<uint16_t>(),
                                st.x16h, st.x16l, s, col));
  } else {
    PNJK_NEED(st.Cp);
    PNJK_NEED(st.X);
    PNJK_HIP_CHECK(launch_ht_f64(st.g, st.ht, st.Cp, e->btri.as<double>(), st.X, s, col));
  }
  return PNJK_OK;
}

int stage_qt(pnjk_engine* e, BuildState& st, hipStream_t s, GroupCollector* col) {
  PNJK_NEED(st.kpart);
  PNJK_NEED(st.d);
  if (!st.want_j && !st.want_jmo && !st.want_kmo) return PNJK_OK;  // nothing reads d or Y
  // d alone (window row 0): the smallest row tile, 128 aux columns
  const TileF64 qt = st.qt_rows == 1 ? TileF64{2, 4} : st.qt;
  if (st.a.precision == PNJK_SPLIT) {
    PNJK_NEED(st.ct16h);
    PNJK_NEED(st.ct16l);
    PNJK_NEED(st.x16h);
    PNJK_NEED(st.x16l);
    PNJK_NEED(st.d16h);
    PNJK_NEED(st.d16l);
    PNJK_HIP_CHECK(launch_qt_x3(st.sg, qt, st.ct16h, st.ct16l, st.x16h, st.x16l, st.kpart, st.d, st.d16h, st.d16l,
                                &st.nt, s, col, st.qt_rows));
  } else {
    PNJK_NEED(st.Cw);
    PNJK_NEED(st.X);
    PNJK_HIP_CHECK(launch_qt_f64(st.g, qt, st.Cw, qt_window_cols(st.g.nbf5), st.X, st.kpart, st.d, &st.nt, s, col,
                                 st.qt_rows));
  }
  return PNJK_OK;
}

int stage_kmo(pnjk_engine* e, BuildState& st, hipStream_t s) {
  if (!st.want_kmo) return PNJK_OK;
  PNJK_NEED(st.kpart);
  PNJK_NEED(st.dKMO);
  PNJK_HIP_CHECK(launch_kmo_reduce(st.g.nbf5, st.nt, st.kpart, st.dKMO, s));
  return PNJK_OK;
}

int stage_jmo(pnjk_engine* e, BuildState& st, hipStream_t s, GroupCollector* col) {
  if (!st.want_jmo) return PNJK_OK;
  PNJK_NEED(st.d);
  PNJK_NEED(st.dJMO);
  PNJK_HIP_CHECK(launch_syrk_f64(1, st.g.nbf5, st.g.naux_p, st.d, st.g.naux_p, 0, st.dJMO, st.g.nbf5, 0,
                                 TileF64{2, 2}, s, col));
  return PNJK_OK;
}

int stage_jn(pnjk_engine* e, BuildState& st, hipStream_t s, GroupCollector* col) {
  if (!st.want_j) return PNJK_OK;
  PNJK_NEED(st.dJ);
  if (st.a.precision == PNJK_FP64) {
    PNJK_NEED(st.d);
    PNJK_HIP_CHECK(launch_jn_f64(st.g, st.jt, st.d, e->btri.as<double>(), st.dJ, s, col));
  } else {
    PNJK_NEED(st.d16h);
    PNJK_NEED(st.d16l);
    PNJK_HIP_CHECK(launch_jn_x3(st.sg, st.jt, st.d16h, st.d16l, e->bt16h.as<uint16_t>(), e->bt16l.as<uint16_t>(),
                                st.dJ, s, col));
  }
  return PNJK_OK;
}

int stage_k(pnjk_engine* e, BuildState& st, hipStream_t s, GroupCollector* col) {
  if (!st.want_k) return PNJK_OK;
  PNJK_NEED(st.dK);
  const int64_t nbf = st.g.nbf;
  if (st.a.precision == PNJK_FP64) {
    PNJK_NEED(st.X);
    PNJK_HIP_CHECK(launch_syrk_f64(st.g.nbf5, st.g.nbf, st.g.naux_p, st.X, st.g.naux_p, nbf * st.g.naux_p, st.dK,
                                   nbf, nbf * nbf, st.kt, s, col));
  } else {
    PNJK_NEED(st.x16h);
    PNJK_NEED(st.x16l);
    PNJK_HIP_CHECK(launch_kn_x3(st.sg, st.kt, st.x16h, st.x16l, st.dK, s, col));
  }
  return PNJK_OK;
}

int stage_sym(BuildState& st, double* O, hipStream_t s) {
  const int64_t nbf = st.g.nbf;
  PNJK_HIP_CHECK(launch_symmetrize(st.g.nbf5, st.g.nbf, O, nbf, nbf * nbf, s));
  return PNJK_OK;
}

int stage_hcore(pnjk_engine* e, BuildState& st, hipStream_t s) {
  if (!st.want_h) return PNJK_OK;
  PNJK_NEED(st.Hp);
  PNJK_NEED(st.T);
  PNJK_NEED(st.Cp);
  PNJK_NEED(st.dH);
  const BuildGeom& g = st.g;
  const int64_t nbf = g.nbf;
  double* Hp = st.Hp;
  PNJK_HIP_CHECK(fill(Hp, nbf * g.nbf_p, 0.0, s));
  if (st.h_dev) {
    const int64_t n = nbf * nbf;
    hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, st.a.H, st.a.ldh, Hp,
                       (int64_t)g.nbf_p, (int)nbf, (int)nbf);
    PNJK_HIP_CHECK(hipGetLastError());
  } else {
    PNJK_HIP_CHECK(hipMemcpy2DAsync(Hp, (size_t)g.nbf_p * 8, st.a.H, (size_t)st.a.ldh * 8, (size_t)nbf * 8,
                                    (size_t)nbf, hipMemcpyHostToDevice, s));
  }
  PNJK_HIP_CHECK(launch_hcore_f64(g, Hp, g.nbf_p, st.Cp, st.T, st.dH, s));
  return PNJK_OK;
}

// Copies host-destined outputs back; returns whether a stream sync is needed.
int stage_out(BuildState& st, hipStream_t s, bool* sync) {
  const int64_t n5 = st.g.nbf5, nbf = st.g.nbf, jk = n5 * nbf * nbf;
  if (st.hJ) PNJK_HIP_CHECK(hipMemcpyAsync(st.a.J, st.dJ, (size_t)jk * 8, hipMemcpyDeviceToHost, s));
  if (st.hK) PNJK_HIP_CHECK(hipMemcpyAsync(st.a.K, st.dK, (size_t)jk * 8, hipMemcpyDeviceToHost, s));
  if (st.want_jmo && st.hJMO)
    PNJK_HIP_CHECK(hipMemcpyAsync(st.a.J_MO, st.dJMO, (size_t)(n5 * n5) * 8, hipMemcpyDeviceToHost, s));
  if (st.want_kmo && st.hKMO)
    PNJK_HIP_CHECK(hipMemcpyAsync(st.a.K_MO, st.dKMO, (size_t)(n5 * n5) * 8, hipMemcpyDeviceToHost, s));
  if (st.hH) PNJK_HIP_CHECK(hipMemcpyAsync(st.a.H_core, st.dH, (size_t)n5 * 8, hipMemcpyDeviceToHost, s));
  *sync = *sync || st.hJ || st.hK || (st.want_jmo && st.hJMO) || (st.want_kmo && st.hKMO) || st.hH;
  return PNJK_OK;
}

}  // namespace

// One build on stream s with the engine's mutex held and its device current; the workspace
// comes from the engine's buffers, or from `scratch` when given (nothing allocated).
static int jk_build_locked(pnjk_engine* e, const BuildArgs& a, hipStream_t s, char* scratch, size_t scratch_bytes) {
  auto mark = [&](int st) {
    if (e->prof) (void)hipEventRecord(e->ev[st], s);
  };
  BuildState st;
  int rc;
  // (the no-allocation build takes the same routes: its entry point checks that the halved
  // diagonal rows the triangular route reads, part of the resident b, exist already)
  if ((rc = build_prepare(e, a, st, scratch, scratch_bytes, true))) return rc;
  mark(ST_IN);
  if ((rc = stage_in(e, st, s))) return rc;
  const bool refine = st.a.precision == PNJK_SPLIT_REFINE;
  // The forms other than split_refine run J_MO and H_core on the side stream beside J^n
  // (about 0.3 ms hidden per build at nbf=1000); J^n's mirror stays in line: beside K^n it
  // slowed K^n by more than it saved.
  const bool side = !st.tri && !st.tri16, k_side = refine && st.want_k;
  hipStream_t sa = (side || k_side) ? e->aux : s;
  bool forked = false;
  auto fork = [&]() -> int {
    PNJK_HIP_CHECK(hipEventRecord(e->fork_ev, s));
    PNJK_HIP_CHECK(hipStreamWaitEvent(sa, e->fork_ev, 0));
    forked = true;
    return PNJK_OK;
  };
  // Everything from the fork to the join; an error part-way still joins the side stream
  // below, so no work of this build is left running past the call on the side stream.
  auto body = [&]() -> int {
    int r;
    mark(ST_HT);
    if ((r = stage_ht(e, st, s, nullptr))) return r;
    mark(ST_QT);
    if ((r = stage_qt(e, st, s, nullptr))) return r;
    mark(ST_KMO);
    if ((r = stage_kmo(e, st, s))) return r;
    mark(ST_JMO);
    if ((r = fork())) return r;
    if ((r = stage_jmo(e, st, sa, nullptr)) || (r = stage_hcore(e, st, sa))) return r;
    if (refine) {  // the split J^n and K^n of split_refine read the split planes of d and X
      const BuildGeom& g = st.g;
      if (st.want_j) {
        PNJK_NEED(st.d16h);
        PNJK_NEED(st.d16l);
        PNJK_HIP_CHECK(launch_split_blocked(st.d, g.nbf5, g.naux_p, st.d16h, st.d16l, s));
      }
      if (st.want_k) {
        PNJK_NEED(st.X);
        PNJK_NEED(st.x16h);
        PNJK_NEED(st.x16l);
        PNJK_HIP_CHECK(launch_split_x16(st.X, st.sg, st.x16h, st.x16l, s));
      }
    }
    mark(ST_JN);
    if ((r = stage_jn(e, st, s, nullptr))) return r;
    if (st.want_j && (r = stage_sym(st, st.dJ, s))) return r;
    mark(ST_K);
    if ((r = stage_k(e, st, s, nullptr))) return r;  // K^n comes out symmetric
    mark(ST_HC);
    return PNJK_OK;
  };
  // The triangular route (split_refine; FP64 builds without J and K): FP64 MO integrals from
  // b = L + L^T (L below the diagonal, b / 2 on it): T = L C (the triangular half-transform,
  // half the half-transform's work), W = C^T T and Y_pq = W_pq + W_qp, so K_MO and d come
  // from W. In split_refine the split half-transform of C^T b and K^n run on the side stream
  // beside them, and J^n (split, from the FP64 d) after W.
  auto refine_body = [&]() -> int {
    int r;
    const BuildGeom& g = st.g;
    PNJK_NEED(st.X);
    PNJK_NEED(st.W);
    PNJK_NEED(st.d);
    PNJK_NEED(st.dKMO);
    if (k_side) {
      PNJK_NEED(st.ct16h);
      PNJK_NEED(st.ct16l);
      PNJK_NEED(st.x16h);
      PNJK_NEED(st.x16l);
      if ((r = fork())) return r;
      PNJK_HIP_CHECK(launch_ht_x3(st.sg, st.ht16, st.ct16h, st.ct16l, e->bht16h.as<uint16_t>(),
                                  e->bht16l.as<uint16_t>(), st.x16h, st.x16l, sa, nullptr));
      if ((r = stage_k(e, st, sa, nullptr))) return r;
    }
    mark(ST_HT);
    TileF64 tt = st.ht;  // a tile spans at most two row slots: at most naux_p columns
    if (32 * tt.fn > g.naux_p) tt.fn = 2;
    const double* dh = e->bdh.as<double>();
    PNJK_HIP_CHECK(
        launch_htl_f64(g, tt, st.Cp, e->btri.as<double>(), dh, dh + (int64_t)g.nbf * g.naux_p, st.X, s));
    mark(ST_QT);
    PNJK_HIP_CHECK(launch_w_f64(g, g.nbf5 > 64 ? TileF64{5, 2} : TileF64{2, 2}, st.Cp, st.X, st.W, s));
    mark(ST_KMO);
    PNJK_HIP_CHECK(launch_ymo_f64(g, st.W, st.dKMO, st.d, st.want_jmo ? st.dJMO : nullptr, s));
    mark(ST_JMO);  // J_MO came with K_MO
    if (st.want_j) {
      PNJK_NEED(st.d16h);
      PNJK_NEED(st.d16l);
      PNJK_HIP_CHECK(launch_split_blocked(st.d, g.nbf5, g.naux_p, st.d16h, st.d16l, s));
    }
    mark(ST_JN);
    if ((r = stage_jn(e, st, s, nullptr))) return r;
    if (st.want_j && (r = stage_sym(st, st.dJ, s))) return r;
    mark(ST_K);  // K^n ran on the side stream
    mark(ST_HC);
    return stage_hcore(e, st, s);
  };
  // The split triangular route: T16 = L C (the HT stage's events), W = C^T T16 (QT), the
  // K_MO and d reduction (KMO), J_MO and H_core in line.
  auto tri16_body = [&]() -> int {
    int r;
    const BuildGeom& g = st.g;
    PNJK_NEED(st.ct16h);
    PNJK_NEED(st.ct16l);
    PNJK_NEED(st.x16h);
    PNJK_NEED(st.x16l);
    PNJK_NEED(st.W);
    PNJK_NEED(st.d);
    PNJK_NEED(st.dKMO);
    mark(ST_HT);
    PNJK_HIP_CHECK(launch_tl_x3(st.sg, st.ct16h, st.ct16l, e->bt16h.as<uint16_t>(), e->bt16l.as<uint16_t>(),
                                roundup(g.ntri, 16), st.x16h, st.x16l, s));
    mark(ST_QT);
    PNJK_HIP_CHECK(launch_ws_x3(st.sg, st.ct16h, st.ct16l, st.x16h, st.x16l, reinterpret_cast<float*>(st.W), s));
    mark(ST_KMO);
    PNJK_HIP_CHECK(
        launch_ymo_f32(g, reinterpret_cast<const float*>(st.W), st.dKMO, st.d, st.want_jmo ? st.dJMO : nullptr, s));
    mark(ST_JMO);  // J_MO came with K_MO
    mark(ST_JN);
    mark(ST_K);
    mark(ST_HC);
    return stage_hcore(e, st, s);
  };
  rc = st.tri ? refine_body() : st.tri16 ? tri16_body() : body();
  if (forked) {  // join the side stream
    const hipError_t j1 = hipEventRecord(e->join_ev, sa);
    const hipError_t j2 = j1 == hipSuccess ? hipStreamWaitEvent(s, e->join_ev, 0) : j1;
    if (j2 != hipSuccess && !rc) PNJK_HIP_CHECK(j2);
  }
  if (rc) return rc;
  mark(ST_OUT);
  bool sync = false;
  if ((rc = stage_out(st, s, &sync))) return rc;
  mark(ST_COUNT);
  if (e->prof) e->prof_recorded = true;
  if (sync) PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
}


extern "C" {

int pnjk_jk_build(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H, int64_t ldh,
                  int precision, uint32_t outputs, double* J, double* K, double* J_MO, double* K_MO, double* H_core,
                  void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_jk_build: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  const BuildArgs a{C, ldc, nbf5, H, ldh, precision, outputs, J, K, J_MO, K_MO, H_core};
  return jk_build_locked(e, a, static_cast<hipStream_t>(stream), nullptr, 0);  // NULL stream: the null stream
}

int pnjk_jk_build_noalloc(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H, int64_t ldh,
                          int precision, uint32_t outputs, double* J, double* K, double* J_MO, double* K_MO,
                          double* H_core, void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_jk_build_noalloc: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  char* ws = scratch_ptr(e->device);
  if (!ws) return set_error(PNJK_ESTATE, "pnjk_jk_build_noalloc: no scratch address for device " +
                                             std::to_string(e->device) + " (pnjk_create initialises it)");
  if (e->nbf == 0) return set_error(PNJK_ESTATE, "pnjk_jk_build_noalloc: no b uploaded");
  if (precision != PNJK_FP64 && !e->split_ready)
    return set_error(PNJK_ESTATE, "pnjk_jk_build_noalloc: the split planes of b are built by the first split build "
                                  "through pnjk_jk_build (they are part of the resident b, not of the workspace)");
  if (precision == PNJK_FP64 && !(outputs & (PNJK_OUT_J | PNJK_OUT_K)) && nbf5 >= 1 &&
      mo_tri_route(precision, e->nbf, nbf5) && !e->hd_ready)
    return set_error(PNJK_ESTATE, "pnjk_jk_build_noalloc: an FP64 build without J and K at nbf5 <= 2 nbf / 3 takes "
                                  "the triangular MO route, whose halved diagonal rows of b are built by the first "
                                  "such build (or any split build) through pnjk_jk_build");
  if (!C || !is_device_ptr(C) || ((outputs & PNJK_OUT_HCORE) && (!H || !is_device_ptr(H))))
    return set_error(PNJK_EINVAL, "pnjk_jk_build_noalloc: C (and H for H_core) must be device memory");
  const BuildArgs a{C, ldc, nbf5, H, ldh, precision, outputs, J, K, J_MO, K_MO, H_core};
  return jk_build_locked(e, a, static_cast<hipStream_t>(stream), ws, kScratchBytes);
}

}  // extern "C"

// Full-width C, k-major: Cf[nu][p] for nu < nbf, p < nbf (from C, row stride ldc, host or
// device); zero rows to nbf_p + 8 and columns to *ldcf, which covers every M-tile (at most
// 160 rows) past nbf.
static int load_cf(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf, int64_t nbf_p, hipStream_t s,
                   int* ldcf_out) {
  const int ldcf = (int)roundup(nbf, 32) + 160;
  const int64_t cf_elems = (nbf_p + 8) * ldcf;
  int rc;
  if ((rc = e->Cf.ensure((size_t)cf_elems * 8))) return rc;
  double* Cf = e->Cf.as<double>();
  PNJK_HIP_CHECK(fill(Cf, cf_elems, 0.0, s));
  if (is_device_ptr(C)) {
    const int64_t n = nbf * nbf;
    hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, C, ldc, Cf, (int64_t)ldcf,
                       (int)nbf, (int)nbf);
    PNJK_HIP_CHECK(hipGetLastError());
  } else {
    PNJK_HIP_CHECK(hipMemcpy2DAsync(Cf, (size_t)ldcf * 8, C, (size_t)ldc * 8, (size_t)nbf * 8, (size_t)nbf,
                                    hipMemcpyHostToDevice, s));
  }
  *ldcf_out = ldcf;
  return PNJK_OK;
}

// The MO integrals into device buffers: b_MO rows of stride ldrow (dB, may be
// NULL) and H_mat (dHm, may be NULL). `st` comes from build_prepare with no outputs.
static int mo_compute(pnjk_engine* e, BuildState& st, double* dB, int64_t ldrow, double* dHm, hipStream_t s,
                      uint16_t* bh = nullptr, uint16_t* bl = nullptr) {
  const BuildGeom& g = st.g;
  const int64_t nbf = g.nbf, n5 = g.nbf5;
  const int precision = st.a.precision;
  int rc;
  int ldcf = 0;
  if (dHm) {  // T = H C[:, :nbf5] (through the H_core stage), then H_mat = Cf^T T
    st.want_h = true;
    if ((rc = e->Hp.ensure((size_t)(nbf * g.nbf_p) * 8)) || (rc = e->T.ensure((size_t)(nbf * g.ldcp + 64) * 8)) ||
        (rc = e->hc.ensure((size_t)n5 * 8)))
      return rc;
    st.dH = e->hc.as<double>();
    st.hH = false;
    st.Hp = e->Hp.as<double>();
    st.T = e->T.as<double>();
  }
  if ((rc = stage_in(e, st, s))) return rc;
  if ((rc = load_cf(e, st.a.C, st.a.ldc, nbf, g.nbf_p, s, &ldcf))) return rc;
  double* Cf = e->Cf.as<double>();
  if (dB) {
    if ((rc = stage_ht(e, st, s, nullptr))) return rc;
    const bool big = nbf >= 256;
    if (precision == PNJK_SPLIT) {
      const size_t ct_bytes = (size_t)(roundup(nbf, 16) * st.sg.ldk) * 2;  // whole 16-row blocks
      if ((rc = e->ctf16h.ensure(ct_bytes)) || (rc = e->ctf16l.ensure(ct_bytes))) return rc;
      PNJK_HIP_CHECK(launch_split_ct(Cf, ldcf, (int)nbf, (int)nbf, (int)nbf, st.sg.ldk, e->ctf16h.as<uint16_t>(),
                                     e->ctf16l.as<uint16_t>(), s));
      PNJK_HIP_CHECK(launch_mo_x3(st.sg, g.naux, big ? TileF64{6, 4} : TileF64{2, 2}, e->ctf16h.as<uint16_t>(),
                                  e->ctf16l.as<uint16_t>(), st.x16h, st.x16l, dB, ldrow,
                                  s, bh, bl));
    } else {  // 128 x 64 tiles (160 x 64: the FP64 Lagrangian 3.5 ms slower at nbf = 1000)
      PNJK_HIP_CHECK(
          launch_mo_f64(g, big ? TileF64{4, 2} : TileF64{2, 2}, Cf, ldcf, st.X, dB, ldrow, s));
    }
  }
  if (dHm) {
    if ((rc = stage_hcore(e, st, s))) return rc;
    PNJK_HIP_CHECK(launch_hmat_f64(g, Cf, ldcf, st.T, dHm, s));
  }
  return PNJK_OK;
}

// Copies n doubles from host or device memory into a device buffer.
static int to_device(const double* src, double* dst, int64_t n, hipStream_t s) {
  PNJK_HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)n * 8, is_device_ptr(src) ? hipMemcpyDeviceToDevice
                                                                            : hipMemcpyHostToDevice, s));
  return PNJK_OK;
}

extern "C" {

int pnjk_mo_integrals(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H, int64_t ldh,
                      int precision, double* b_MO, double* H_mat, void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_mo_integrals: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  if (e->nbf == 0) return set_error(PNJK_ESTATE, "pnjk_mo_integrals: no b uploaded");
  if (!C) return set_error(PNJK_EINVAL, "pnjk_mo_integrals: C is NULL");
  if (ldc < e->nbf) return set_error(PNJK_EINVAL, "pnjk_mo_integrals: ldc < nbf (every column of C is used)");
  if (precision != PNJK_FP64 && precision != PNJK_SPLIT)
    return set_error(PNJK_EUNSUPPORTED, "pnjk_mo_integrals: precision must be PNJK_FP64 or PNJK_SPLIT");
  if (!b_MO && !H_mat) return set_error(PNJK_EINVAL, "pnjk_mo_integrals: neither b_MO nor H_mat requested");
  if (H_mat && (!H || ldh < e->nbf))
    return set_error(PNJK_EINVAL, "pnjk_mo_integrals: H_mat needs H with ldh >= nbf");
  BuildArgs a{C, ldc, nbf5, H, ldh, precision, 0u, nullptr, nullptr, nullptr, nullptr, nullptr};
  BuildState st;
  int rc;
  if ((rc = build_prepare(e, a, st))) return rc;
  const int64_t nbf = st.g.nbf, n5 = st.g.nbf5, naux = st.g.naux;
  const bool host_b = b_MO && !is_device_ptr(b_MO), host_h = H_mat && !is_device_ptr(H_mat);
  const int64_t b_elems = nbf * n5 * naux, h_elems = nbf * n5;
  if ((host_b || host_h) &&
      (rc = e->outs.ensure((size_t)((host_b ? b_elems : 0) + (host_h ? h_elems : 0)) * 8)))
    return rc;
  double* dB = host_b ? e->outs.as<double>() : b_MO;
  double* dHm = host_h ? e->outs.as<double>() + (host_b ? b_elems : 0) : H_mat;
  if ((rc = mo_compute(e, st, dB, n5 * naux, dHm, s))) return rc;
  if (host_b) PNJK_HIP_CHECK(hipMemcpyAsync(b_MO, dB, (size_t)b_elems * 8, hipMemcpyDeviceToHost, s));
  if (host_h) PNJK_HIP_CHECK(hipMemcpyAsync(H_mat, dHm, (size_t)h_elems * 8, hipMemcpyDeviceToHost, s));
  if (host_b || host_h) PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
}

int pnjk_lagrangian(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H, int64_t ldh,
                    const double* n, const double* cj12, const double* ck12, int64_t nbeta, int64_t nalpha,
                    int precision, double* elag, double* H_mat, void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_lagrangian: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  if (e->nbf == 0) return set_error(PNJK_ESTATE, "pnjk_lagrangian: no b uploaded");
  if (!C || !H || !n || !cj12 || !ck12 || !elag)
    return set_error(PNJK_EINVAL, "pnjk_lagrangian: C, H, n, cj12, ck12 and elag are required");
  if (ldc < e->nbf || ldh < e->nbf) return set_error(PNJK_EINVAL, "pnjk_lagrangian: ldc and ldh must be >= nbf");
  if (precision != PNJK_FP64 && precision != PNJK_SPLIT)
    return set_error(PNJK_EUNSUPPORTED, "pnjk_lagrangian: precision must be PNJK_FP64 or PNJK_SPLIT");
  if (nbeta < 0 || nalpha < nbeta || nalpha > nbf5)
    return set_error(PNJK_EINVAL, "pnjk_lagrangian: need 0 <= nbeta <= nalpha <= nbf5");
  BuildArgs a{C, ldc, nbf5, H, ldh, precision, 0u, nullptr, nullptr, nullptr, nullptr, nullptr};
  BuildState st;
  int rc;
  if ((rc = build_prepare(e, a, st))) return rc;
  const BuildGeom& g = st.g;
  const int64_t nbf = g.nbf, n5 = g.nbf5, naux = g.naux;
  // Split-K chunking of k' = (q, P) for the exchange GEMM: about 2048 tiles of 128 x 64.
  const int64_t kfull = n5 * naux, tiles = ((nbf + 127) / 128) * ((n5 + 63) / 64);
  const int64_t want = std::max<int64_t>(1, (2048 + tiles - 1) / tiles);
  const int64_t kc = roundup(std::max<int64_t>(1, (kfull + want - 1) / want), 64);
  const int64_t nchunks = (kfull + kc - 1) / kc, ldrow = nchunks * kc;
  if ((rc = e->bmo.ensure((size_t)(nbf * ldrow) * 8)) || (rc = e->hm.ensure((size_t)(nbf * n5) * 8)) ||
      // lagw: the FP64 W of the exchange (FP64 form), and staging for C^J, C^K in both forms
      (rc = e->lagw.ensure((size_t)(precision == PNJK_SPLIT ? n5 * n5 : n5 * ldrow) * 8)) ||
      (rc = e->lagg.ensure((size_t)(n5 * naux) * 8)) ||
      (rc = e->lagp.ensure((size_t)(nchunks * nbf * n5) * 8)) || (rc = e->lagc.ensure((size_t)(3 * n5 * n5) * 8)))
    return rc;
  const bool host_e = !is_device_ptr(elag), host_h = H_mat && !is_device_ptr(H_mat);
  if (host_e && (rc = e->outs.ensure((size_t)(nbf * nbf) * 8))) return rc;
  double* dE = host_e ? e->outs.as<double>() : elag;
  double* B = e->bmo.as<double>();
  double* Hm = e->hm.as<double>();
  double* cf = e->lagc.as<double>();  // [n5 * n5 C^J | n5 * n5 C^K | n5 occupations]
  double* cjd = cf;
  double* ckd = cf + n5 * n5;
  double* nd = cf + 2 * n5 * n5;
  const bool split = precision == PNJK_SPLIT;
  if (split) {  // exchange in BF16x3: b_MO's split planes come from the MO GEMM's epilogue
    const size_t bb = (size_t)(nbf * ldrow) * 2, wb = (size_t)(n5 * ldrow) * 2;
    if ((rc = e->bm16h.ensure(bb)) || (rc = e->bm16l.ensure(bb)) || (rc = e->w16h.ensure(wb)) ||
        (rc = e->w16l.ensure(wb)))
      return rc;
  }
  if ((rc = mo_compute(e, st, B, ldrow, Hm, s, split ? e->bm16h.as<uint16_t>() : nullptr,
                       split ? e->bm16l.as<uint16_t>() : nullptr)))
    return rc;
  PNJK_HIP_CHECK(launch_fill_cols(B, (int)nbf, ldrow, kfull, ldrow, s));
  if (split)
    PNJK_HIP_CHECK(launch_fill_cols16(e->bm16h.as<uint16_t>(), e->bm16l.as<uint16_t>(), (int)nbf, ldrow, kfull, ldrow,
                                      s));
  // Coefficients: copy, then clear the diagonal (utils.py:269-270) on the device.
  if ((rc = to_device(cj12, e->lagw.as<double>(), n5 * n5, s))) return rc;
  PNJK_HIP_CHECK(launch_lag_coef(e->lagw.as<double>(), cjd, (int)n5, s));
  if ((rc = to_device(ck12, e->lagw.as<double>(), n5 * n5, s))) return rc;
  PNJK_HIP_CHECK(launch_lag_coef(e->lagw.as<double>(), ckd, (int)n5, s));
  if ((rc = to_device(n, nd, n5, s))) return rc;
  if (split)
    PNJK_HIP_CHECK(launch_lag_x3((int)nbf, (int)n5, (int)naux, B, ldrow, kc, (int)nchunks, ckd,
                                 e->bm16h.as<uint16_t>(), e->bm16l.as<uint16_t>(), e->w16h.as<uint16_t>(),
                                 e->w16l.as<uint16_t>(), e->lagp.as<double>(), true, s));
  PNJK_HIP_CHECK(launch_lag_f64(g, B, ldrow, kc, (int)nchunks, nd, cjd, ckd, (int)nbeta, (int)nalpha, Hm,
                                e->lagg.as<double>(), e->lagw.as<double>(), e->lagp.as<double>(), dE, !split, s));
  if (host_e) PNJK_HIP_CHECK(hipMemcpyAsync(elag, dE, (size_t)(nbf * nbf) * 8, hipMemcpyDeviceToHost, s));
  if (H_mat)
    PNJK_HIP_CHECK(hipMemcpyAsync(H_mat, Hm, (size_t)(nbf * n5) * 8,
                                  host_h ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
  if (host_e || host_h) PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
}

// Copies a rows x cols block (source row stride lds, host or device) to device memory of
// row stride ldd.
static int to_device_2d(const double* src, int64_t lds, double* dst, int64_t ldd, int64_t rows, int64_t cols,
                        hipStream_t s) {
  PNJK_HIP_CHECK(hipMemcpy2DAsync(dst, (size_t)ldd * 8, src, (size_t)lds * 8, (size_t)cols * 8, (size_t)rows,
                                  is_device_ptr(src) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  return PNJK_OK;
}

// pnjk_fock_rc and pnjk_fock_ro: one coefficient kernel and one FP64 GEMM, the shell chosen by
// open_shell; `fn` names the entry point in error messages.
static int fock_common(pnjk_engine* e, const char* fn, bool open_shell, const double* J, const double* K,
                       int64_t nbf, int64_t nbf5, const double* n, const double* H, int64_t ldh,
                       const double* cj12, const double* ck12, int64_t no1, int64_t nbeta, int64_t nalpha, double* F,
                       void* stream) {
  if (!e) return set_error(PNJK_EINVAL, std::string(fn) + ": engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  if (!J || !K || !n || !H || !cj12 || !ck12 || !F)
    return set_error(PNJK_EINVAL, std::string(fn) + ": J, K, n, H, cj12, ck12 and F are required");
  if (nbf < 1 || nbf5 < 1 || nbf5 > nbf || nbf * nbf > (int64_t)1 << 30)
    return set_error(PNJK_EINVAL, std::string(fn) + ": need 1 <= nbf5 <= nbf and nbf^2 <= 2^30");
  if (ldh < nbf) return set_error(PNJK_EINVAL, std::string(fn) + ": ldh < nbf");
  if (no1 < 0 || nbeta < 0 || nalpha < nbeta || nalpha > nbf5)
    return set_error(PNJK_EINVAL, std::string(fn) + ": need no1 >= 0 and 0 <= nbeta <= nalpha <= nbf5");
  const int64_t n5 = nbf5, nn = nbf * nbf;
  const int ini = no1 > 1 ? (int)no1 : 0;  // utils.py:31-33, :101-103
  // J and K feed the GEMM in place when they are device rows with 16-byte aligned pairs;
  // otherwise they are staged with an even row stride.
  const bool direct = is_device_ptr(J) && is_device_ptr(K) && nn % 2 == 0 && (uintptr_t)J % 16 == 0 &&
                      (uintptr_t)K % 16 == 0;
  const int64_t ld = direct ? nn : roundup(nn, 2);
  const int64_t kp = roundup(2 * n5, 16);
  int rc;
  if ((!direct && (rc = e->fjk.ensure((size_t)(2 * n5 * ld) * 8))) || (rc = e->fh.ensure((size_t)nn * 8)) ||
      (rc = e->fco.ensure((size_t)(n5 * kp + 2 * n5 * n5 + 2 * n5) * 8)))
    return rc;
  const bool host_f = !is_device_ptr(F);
  if (host_f && (rc = e->ff.ensure((size_t)(n5 * nn) * 8))) return rc;
  const double* dJ = J;
  const double* dK = K;
  if (!direct) {
    double* jk = e->fjk.as<double>();
    if ((rc = to_device_2d(J, nn, jk, ld, n5, nn, s)) || (rc = to_device_2d(K, nn, jk + n5 * ld, ld, n5, nn, s)))
      return rc;
    dJ = jk;
    dK = jk + n5 * ld;
  }
  double* A = e->fco.as<double>();
  double* cjd = A + n5 * kp;
  double* ckd = cjd + n5 * n5;
  double* nd = ckd + n5 * n5;
  double* hcd = nd + n5;
  if ((rc = to_device_2d(H, ldh, e->fh.as<double>(), nbf, nbf, nbf, s)) || (rc = to_device(cj12, cjd, n5 * n5, s)) ||
      (rc = to_device(ck12, ckd, n5 * n5, s)) || (rc = to_device(n, nd, n5, s)))
    return rc;
  double* dF = host_f ? e->ff.as<double>() : F;
  PNJK_HIP_CHECK(launch_fock_f64((int)nbf, (int)n5, dJ, dK, ld, e->fh.as<double>(), nd, cjd, ckd, ini, (int)nbeta,
                                 (int)nalpha, open_shell, A, hcd, dF, s));
  if (host_f) {
    PNJK_HIP_CHECK(hipMemcpyAsync(F, dF, (size_t)(n5 * nn) * 8, hipMemcpyDeviceToHost, s));
    PNJK_HIP_CHECK(hipStreamSynchronize(s));
  } else if (!is_device_ptr(J) || !is_device_ptr(K) || !is_device_ptr(H) || !is_device_ptr(n) ||
             !is_device_ptr(cj12) || !is_device_ptr(ck12)) {
    // Host inputs were copied asynchronously from pageable memory the caller may reuse.
    PNJK_HIP_CHECK(hipStreamSynchronize(s));
  }
  return PNJK_OK;
}

int pnjk_fock_rc(pnjk_engine* e, const double* J, const double* K, int64_t nbf, int64_t nbf5, const double* n,
                 const double* H, int64_t ldh, const double* cj12, const double* ck12, int64_t no1, int64_t nbeta,
                 int64_t nalpha, double* F, void* stream) {
  return fock_common(e, "pnjk_fock_rc", false, J, K, nbf, nbf5, n, H, ldh, cj12, ck12, no1, nbeta, nalpha, F, stream);
}

int pnjk_fock_ro(pnjk_engine* e, const double* J, const double* K, int64_t nbf, int64_t nbf5, const double* n,
                 const double* H, int64_t ldh, const double* cj12, const double* ck12, int64_t no1, int64_t nbeta,
                 int64_t nalpha, double* F, void* stream) {
  return fock_common(e, "pnjk_fock_ro", true, J, K, nbf, nbf5, n, H, ldh, cj12, ck12, no1, nbeta, nalpha, F, stream);
}

int pnjk_lagrange_fock(pnjk_engine* e, const double* F, int64_t nbf, int64_t nbf5, const double* C, int64_t ldc,
                       int64_t noptorb, double* elag, void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_lagrange_fock: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  if (!F || !C || !elag) return set_error(PNJK_EINVAL, "pnjk_lagrange_fock: F, C and elag are required");
  if (nbf < 1 || nbf5 < 1 || nbf5 > nbf || nbf5 > 65535 || nbf * nbf > (int64_t)1 << 30)
    return set_error(PNJK_EINVAL, "pnjk_lagrange_fock: need 1 <= nbf5 <= min(nbf, 65535) and nbf^2 <= 2^30");
  if (ldc < nbf) return set_error(PNJK_EINVAL, "pnjk_lagrange_fock: ldc < nbf (every column of C is used)");
  if (noptorb < 0 || noptorb > nbf) return set_error(PNJK_EINVAL, "pnjk_lagrange_fock: need 0 <= noptorb <= nbf");
  const int64_t n5 = nbf5, nn = nbf * nbf;
  BuildGeom g{};
  g.nbf = (int)nbf;
  g.nbf5 = (int)n5;
  g.nbf_p = (int)roundup(nbf, 32);
  g.ldcp = (int)roundup(n5, 32);
  int rc;
  const bool host_f = !is_device_ptr(F), host_e = !is_device_ptr(elag);
  if ((host_f && (rc = e->ff.ensure((size_t)(n5 * nn) * 8))) ||
      (rc = e->T.ensure((size_t)(nbf * g.ldcp + 64) * 8)) || (rc = e->hm.ensure((size_t)(nbf * n5) * 8)) ||
      (host_e && (rc = e->outs.ensure((size_t)nn * 8))))
    return rc;
  const double* dF = F;
  if (host_f) {
    if ((rc = to_device(F, e->ff.as<double>(), n5 * nn, s))) return rc;
    dF = e->ff.as<double>();
  }
  int ldcf = 0;
  if ((rc = load_cf(e, C, ldc, nbf, g.nbf_p, s, &ldcf))) return rc;
  double* G = e->T.as<double>();
  PNJK_HIP_CHECK(fill(G, nbf * g.ldcp + 64, 0.0, s));
  double* dE = host_e ? e->outs.as<double>() : elag;
  PNJK_HIP_CHECK(launch_lagrange_fock_f64(g, dF, e->Cf.as<double>(), ldcf, G, e->hm.as<double>(), (int)noptorb, dE, s));
  if (host_e) PNJK_HIP_CHECK(hipMemcpyAsync(elag, dE, (size_t)nn * 8, hipMemcpyDeviceToHost, s));
  if (host_e || host_f || !is_device_ptr(C)) PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
}

int pnjk_lagrange_convergence(pnjk_engine* e, const double* elag, int64_t nbf, double* sumdiff, double* maxdiff,
                              void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_lagrange_convergence: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  if (!elag || !sumdiff || !maxdiff) return set_error(PNJK_EINVAL, "pnjk_lagrange_convergence: null argument");
  if (nbf < 1 || nbf * nbf > (int64_t)1 << 31) return set_error(PNJK_EINVAL, "pnjk_lagrange_convergence: bad nbf");
  int rc;
  const bool host_e = !is_device_ptr(elag);
  if ((rc = e->scal.ensure(256)) || (host_e && (rc = e->outs.ensure((size_t)(nbf * nbf) * 8)))) return rc;
  const double* dE = elag;
  if (host_e) {
    if ((rc = to_device(elag, e->outs.as<double>(), nbf * nbf, s))) return rc;
    dE = e->outs.as<double>();
  }
  double* out = e->scal.as<double>() + 2;
  PNJK_HIP_CHECK(launch_lag_conv(dE, (int)nbf, out, s));
  double h[2];
  PNJK_HIP_CHECK(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, s));
  PNJK_HIP_CHECK(hipStreamSynchronize(s));
  *sumdiff = h[0];
  *maxdiff = h[1];
  return PNJK_OK;
}

// ---------------------------------------------------------------- occupation step
// Checks a pnjk_pnof against the relations of parameters.py:46-65, 94 and fills the
// kernel-side copy.
static int pnof_from(const pnjk_pnof* p, PnofDev* d, const char* who) {
  const std::string w(who);
  if (!p) return set_error(PNJK_EINVAL, w + ": pnof parameters are NULL");
  if (p->ipnof != 5 && p->ipnof != 7 && p->ipnof != 8)
    return set_error(PNJK_EUNSUPPORTED, w + ": ipnof must be 5, 7 or 8");
  if (p->occ_method != PNJK_OCC_SOFTMAX && p->occ_method != PNJK_OCC_TRIGONOMETRIC)
    return set_error(PNJK_EUNSUPPORTED, w + ": occ_method must be PNJK_OCC_SOFTMAX or PNJK_OCC_TRIGONOMETRIC");
  if (p->no1 < 0 || p->ndoc < 0 || p->nsoc < 0 || p->ncwo < 0)
    return set_error(PNJK_EINVAL, w + ": no1, ndoc, nsoc, ncwo must be >= 0");
  const int64_t nbf5 = (int64_t)p->no1 + (int64_t)p->ndoc * (1 + (int64_t)p->ncwo) + p->nsoc;
  if (p->nbf5 != nbf5 || nbf5 < 1 || nbf5 > 16384)
    return set_error(PNJK_EINVAL, w + ": nbf5 must equal no1 + ndoc (1 + ncwo) + nsoc, in [1, 16384]");
  d->ipnof = p->ipnof;
  d->ista = p->ista;
  d->method = p->occ_method;
  d->high_spin = p->high_spin ? 1 : 0;
  d->mspin = p->mspin;
  d->no1 = p->no1;
  d->ndoc = p->ndoc;
  d->nsoc = p->nsoc;
  d->ncwo = p->ncwo;
  d->nbf5 = (int)nbf5;
  d->nv = p->ndoc * p->ncwo;
  d->nbeta = p->no1 + p->ndoc;
  d->nalpha = d->nbeta + p->nsoc;
  d->ndns = p->ndoc + p->nsoc;
  return PNJK_OK;
}

// Device arrays of one occupation step, carved from the engine's `occ` buffer.
struct OccBufs {
  double *x, *n, *dn, *fi, *nd, *nd12, *dfi, *dnd, *dnd12, *JT, *KT, *tmp, *hc, *E, *grad, *cj, *ck, *part;
};

static int occ_bufs(pnjk_engine* e, int64_t n5, int64_t nv, OccBufs& b) {
  const int64_t mm = n5 * n5, mv = n5 * nv;
  const int64_t sizes[18] = {nv, n5, mv, n5, n5, n5, mv, mv, mv, mm, mm, mm, n5, 1, nv, mm, mm, (n5 + 3) / 4};
  int64_t total = 0;
  for (int64_t z : sizes) total += (z + 31) / 32 * 32;  // 256-byte aligned pieces
  int rc = e->occ.ensure((size_t)std::max<int64_t>(total, 32) * 8);
  if (rc) return rc;
  double* base = e->occ.as<double>();
  double** slots[18] = {&b.x, &b.n, &b.dn, &b.fi, &b.nd, &b.nd12, &b.dfi, &b.dnd, &b.dnd12,
                        &b.JT, &b.KT, &b.tmp, &b.hc, &b.E, &b.grad, &b.cj, &b.ck, &b.part};
  int64_t off = 0;
  for (int i = 0; i < 18; ++i) {
    *slots[i] = base + off;
    off += (sizes[i] + 31) / 32 * 32;
  }
  return PNJK_OK;
}

// Stages M (n x n, host or device) transposed into dst.
static int stage_transposed(const double* M, int64_t n, double* tmp, double* dst, hipStream_t s) {
  const double* src = M;
  if (!is_device_ptr(M)) {
    PNJK_HIP_CHECK(hipMemcpyAsync(tmp, M, (size_t)(n * n) * 8, hipMemcpyHostToDevice, s));
    src = tmp;
  }
  PNJK_HIP_CHECK(launch_occ_transpose(src, (int)n, dst, s));
  return PNJK_OK;
}

// Copies n doubles from device memory to host or device memory.
static int from_device(const double* src, double* dst, int64_t n, hipStream_t s, bool* host) {
  const bool h = !is_device_ptr(dst);
  PNJK_HIP_CHECK(hipMemcpyAsync(dst, src, (size_t)n * 8, h ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, s));
  *host = *host || h;
  return PNJK_OK;
}

int pnjk_occupation_step(pnjk_engine* e, const pnjk_pnof* p, const double* gamma, const double* J_MO,
                         const double* K_MO, const double* H_core, double* E, double* grad, double* n,
                         double* dn_dgamma, double* cj12, double* ck12, void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_occupation_step: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  PnofDev d;
  int rc;
  if ((rc = pnof_from(p, &d, "pnjk_occupation_step"))) return rc;
  if (!E && !grad && !n && !dn_dgamma && !cj12 && !ck12)
    return set_error(PNJK_EINVAL, "pnjk_occupation_step: no output requested");
  if (d.nv > 0 && !gamma) return set_error(PNJK_EINVAL, "pnjk_occupation_step: gamma is NULL");
  if ((E || grad) && (!J_MO || !K_MO || !H_core))
    return set_error(PNJK_EINVAL, "pnjk_occupation_step: E and grad need J_MO, K_MO and H_core");
  const int64_t n5 = d.nbf5, nv = d.nv;
  OccBufs b;
  if ((rc = occ_bufs(e, n5, nv, b))) return rc;
  if (nv > 0 && (rc = to_device(gamma, b.x, nv, s))) return rc;
  PNJK_HIP_CHECK(launch_occupation(d, b.x, b.n, b.dn, s));
  const bool want_der = grad != nullptr;
  PNJK_HIP_CHECK(launch_pnof_orbitals(d, b.n, want_der ? b.dn : nullptr, b.fi, b.nd, b.nd12, b.dfi, b.dnd, b.dnd12, s));
  if (E || grad) {
    if ((rc = stage_transposed(J_MO, n5, b.tmp, b.JT, s)) || (rc = stage_transposed(K_MO, n5, b.tmp, b.KT, s)) ||
        (rc = to_device(H_core, b.hc, n5, s)))
      return rc;
  }
  if (E) PNJK_HIP_CHECK(launch_occ_energy(d, b.n, b.fi, b.nd, b.nd12, nullptr, nullptr, b.JT, b.KT, b.hc, b.part, b.E, s));
  if (grad)
    PNJK_HIP_CHECK(launch_occ_grad(d, b.n, b.fi, b.nd, b.nd12, b.dn, b.dfi, b.dnd, b.dnd12, b.JT, b.KT, b.hc, b.grad,
                                   s));
  if (cj12 || ck12) PNJK_HIP_CHECK(launch_pnof_coef(d, b.n, b.fi, b.nd, b.nd12, b.cj, b.ck, s));
  bool host = false;
  if ((E && (rc = from_device(b.E, E, 1, s, &host))) || (grad && nv > 0 && (rc = from_device(b.grad, grad, nv, s, &host))) ||
      (n && (rc = from_device(b.n, n, n5, s, &host))) ||
      (dn_dgamma && nv > 0 && (rc = from_device(b.dn, dn_dgamma, n5 * nv, s, &host))) ||
      (cj12 && (rc = from_device(b.cj, cj12, n5 * n5, s, &host))) ||
      (ck12 && (rc = from_device(b.ck, ck12, n5 * n5, s, &host))))
    return rc;
  if (host || (gamma && !is_device_ptr(gamma))) PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
}

int pnjk_pnof_coefficients(pnjk_engine* e, const pnjk_pnof* p, const double* n, double* cj12, double* ck12,
                           void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_pnof_coefficients: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  PnofDev d;
  int rc;
  if ((rc = pnof_from(p, &d, "pnjk_pnof_coefficients"))) return rc;
  if (!n || (!cj12 && !ck12)) return set_error(PNJK_EINVAL, "pnjk_pnof_coefficients: n and an output are required");
  const int64_t n5 = d.nbf5;
  OccBufs b;
  if ((rc = occ_bufs(e, n5, d.nv, b)) || (rc = to_device(n, b.n, n5, s))) return rc;
  PNJK_HIP_CHECK(launch_pnof_orbitals(d, b.n, nullptr, b.fi, b.nd, b.nd12, nullptr, nullptr, nullptr, s));
  PNJK_HIP_CHECK(launch_pnof_coef(d, b.n, b.fi, b.nd, b.nd12, b.cj, b.ck, s));
  bool host = false;
  if ((cj12 && (rc = from_device(b.cj, cj12, n5 * n5, s, &host))) ||
      (ck12 && (rc = from_device(b.ck, ck12, n5 * n5, s, &host))))
    return rc;
  if (host || !is_device_ptr(n)) PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
}

int pnjk_energy(pnjk_engine* e, int64_t nbf5, int64_t nbeta, int64_t nalpha, int32_t mspin, const double* n,
                const double* cj12, const double* ck12, const double* J_MO, const double* K_MO, const double* H_core,
                double* E, void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_energy: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  if (!n || !cj12 || !ck12 || !J_MO || !K_MO || !H_core || !E)
    return set_error(PNJK_EINVAL, "pnjk_energy: every array is required");
  if (nbf5 < 1 || nbf5 > 16384 || nbeta < 0 || nalpha < nbeta || nalpha > nbf5)
    return set_error(PNJK_EINVAL, "pnjk_energy: need 1 <= nbf5 <= 16384 and 0 <= nbeta <= nalpha <= nbf5");
  PnofDev d{};
  d.nbf5 = (int)nbf5;
  d.nbeta = (int)nbeta;
  d.nalpha = (int)nalpha;
  d.nsoc = (int)(nalpha - nbeta);
  d.mspin = mspin;
  OccBufs b;
  int rc;
  if ((rc = occ_bufs(e, nbf5, 0, b))) return rc;
  if ((rc = to_device(n, b.n, nbf5, s)) || (rc = to_device(cj12, b.cj, nbf5 * nbf5, s)) ||
      (rc = to_device(ck12, b.ck, nbf5 * nbf5, s)) || (rc = stage_transposed(J_MO, nbf5, b.tmp, b.JT, s)) ||
      (rc = stage_transposed(K_MO, nbf5, b.tmp, b.KT, s)) || (rc = to_device(H_core, b.hc, nbf5, s)))
    return rc;
  PNJK_HIP_CHECK(launch_occ_energy(d, b.n, nullptr, nullptr, nullptr, b.cj, b.ck, b.JT, b.KT, b.hc, b.part, b.E, s));
  bool host = false;
  if ((rc = from_device(b.E, E, 1, s, &host))) return rc;
  if (host || !is_device_ptr(n) || !is_device_ptr(cj12) || !is_device_ptr(ck12) || !is_device_ptr(J_MO))
    PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
}

}  // extern "C"

struct pnjk_group {
  std::vector<pnjk_engine*> eng;
  int device = 0;
  std::mutex mu;
  hipStream_t cap = nullptr;            // capture stream (made on the first graph-mode build)
  std::vector<BuildState> st;
  std::vector<char> sig;                // signature of the cached plan
  std::vector<char> in_sig;             // C and H pointers the captured graph reads
  std::vector<GroupCollector> plan;     // grouped stages: HT, QT, J_MO, J^n, K^n
  std::vector<size_t> solo_ht;          // instances whose half-transform runs as its own launch
  std::vector<SymJob> sym;              // symmetrize jobs of the J^n outputs
  size_t sym_off = 0;                   // their offset in the plan buffer
  int64_t sym_grid = 0;
  DevBuf plan_dev;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_valid = false;
  hipEvent_t done = nullptr;  // recorded after each build's last launch
  ~pnjk_group() {
    if (done) (void)hipEventDestroy(done);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    plan_dev.release();
  }
};

// Plan buffers of destroyed groups, per device, handed to new groups: freeing device memory
// synchronises with the whole device, and a serving loop destroys a group whenever its set of
// requests changes. A buffer enters the pool only after its group's last build has run (the
// event recorded behind it), so no launch still reads it when a new group overwrites it.
namespace {
constexpr size_t kPlanPoolMax = 16;
std::mutex g_plan_mu;
std::vector<DevBuf> g_plan_pool[64];
}  // namespace

namespace {

enum { GS_HT = 0, GS_QT, GS_JMO, GS_JN, GS_K, GS_COUNT };

int launch_plan(pnjk_group* g, int stage, hipStream_t s) {
  const char* base = static_cast<const char*>(g->plan_dev.p);
  for (const GroupBatch& b : g->plan[stage].batches) {
    if (b.tiles.empty()) continue;
    PNJK_HIP_CHECK(b.launch(base + b.probs_off, reinterpret_cast<const int2*>(base + b.tiles_off),
                            reinterpret_cast<const int*>(base + b.bounds_off), b.grid, s));
  }
  return PNJK_OK;
}

// The whole grouped build, enqueued on s (capturable: kernels only).
int group_enqueue(pnjk_group* g, hipStream_t s) {
  const size_t n = g->eng.size();
  int rc;
  for (size_t i = 0; i < n; ++i)
    if ((rc = stage_in(g->eng[i], g->st[i], s))) return rc;
  for (size_t i : g->solo_ht)
    if ((rc = stage_ht(g->eng[i], g->st[i], s, nullptr))) return rc;
  if ((rc = launch_plan(g, GS_HT, s)) || (rc = launch_plan(g, GS_QT, s))) return rc;
  for (size_t i = 0; i < n; ++i)
    if ((rc = stage_kmo(g->eng[i], g->st[i], s))) return rc;
  if ((rc = launch_plan(g, GS_JMO, s))) return rc;
  const char* base = static_cast<const char*>(g->plan_dev.p);
  if ((rc = launch_plan(g, GS_JN, s))) return rc;
  PNJK_HIP_CHECK(launch_symmetrize_grouped(reinterpret_cast<const SymJob*>(base + g->sym_off),
                                           (int)g->sym.size(), g->sym_grid, s));
  if ((rc = launch_plan(g, GS_K, s))) return rc;  // K^n comes out symmetric
  for (size_t i = 0; i < n; ++i)
    if ((rc = stage_hcore(g->eng[i], g->st[i], s))) return rc;
  return PNJK_OK;
}

template <class T>
void sig_put(std::vector<char>& v, const T& x) {
  const char* r = reinterpret_cast<const char*>(&x);
  v.insert(v.end(), r, r + sizeof(T));
}

}  // namespace

extern "C" {

int pnjk_group_create(pnjk_engine* const* engines, int count, pnjk_group** out) {
  if (!out || !engines || count < 1) return set_error(PNJK_EINVAL, "pnjk_group_create: bad arguments");
  *out = nullptr;
  for (int i = 0; i < count; ++i) {
    if (!engines[i]) return set_error(PNJK_EINVAL, "pnjk_group_create: NULL engine");
    if (engines[i]->device != engines[0]->device)
      return set_error(PNJK_EINVAL, "pnjk_group_create: engines must share one device");
    for (int j = 0; j < i; ++j)
      if (engines[j] == engines[i])
        return set_error(PNJK_EINVAL, "pnjk_group_create: an engine appears twice (each instance needs its own)");
  }
  pnjk_group* g = new pnjk_group();
  g->eng.assign(engines, engines + count);
  g->device = engines[0]->device;
  g->plan.resize(GS_COUNT);
  if (g->device >= 0 && g->device < 64) {
    std::lock_guard<std::mutex> pl(g_plan_mu);
    std::vector<DevBuf>& pool = g_plan_pool[g->device];
    if (!pool.empty()) {
      g->plan_dev = pool.back();
      pool.pop_back();
    }
  }
  // The capture stream is made on the first graph-mode build: creating a stream synchronises
  // with the device, and groups that never capture (a serving loop's) need none.
  const hipError_t he = hipSetDevice(g->device);
  if (he != hipSuccess) {
    delete g;
    return set_error(PNJK_EHIP, std::string("pnjk_group_create: ") + hipGetErrorString(he));
  }
  *out = g;
  return PNJK_OK;
}

int pnjk_group_destroy(pnjk_group* g) {
  if (!g) return PNJK_OK;
  (void)hipSetDevice(g->device);
  if (g->cap) (void)hipStreamSynchronize(g->cap);
  if (g->plan_dev.p && g->device >= 0 && g->device < 64 && (!g->done || hipEventSynchronize(g->done) == hipSuccess)) {
    std::lock_guard<std::mutex> pl(g_plan_mu);
    std::vector<DevBuf>& pool = g_plan_pool[g->device];
    if (pool.size() < kPlanPoolMax) {
      pool.push_back(g->plan_dev);
      g->plan_dev = DevBuf{};
    }
  }
  delete g;
  return PNJK_OK;
}

int pnjk_group_build(pnjk_group* g, const double* const* C, const int64_t* ldc, const int64_t* nbf5,
                     const double* const* H, const int64_t* ldh, int precision, uint32_t outputs,
                     double* const* J, double* const* K, double* const* J_MO, double* const* K_MO,
                     double* const* H_core, int use_graph, void* stream) {
  if (!g || !C || !ldc || !nbf5) return set_error(PNJK_EINVAL, "pnjk_group_build: null argument");
  std::lock_guard<std::mutex> lk(g->mu);
  const size_t n = g->eng.size();
  std::vector<std::unique_lock<std::mutex>> locks;
  locks.reserve(n);
  for (pnjk_engine* e : g->eng) locks.emplace_back(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(g->device));
  hipStream_t s = static_cast<hipStream_t>(stream);  // NULL: the null stream
  auto at = [](double* const* a, size_t i) { return a ? a[i] : nullptr; };
  int rc;
  // The triangular MO routes (a triangular half-transform, W and its reduction, in FP64 or
  // BF16x3) are not grouped: split_refine, and FP64 or split builds without J and K when an
  // instance takes such a route alone (mo_tri_route), build each instance on its own, in
  // order, on this stream (a captured graph is not used for them), so every instance takes
  // the route it takes alone.
  bool per_instance = precision == PNJK_SPLIT_REFINE;
  if (precision != PNJK_SPLIT_REFINE && !(outputs & (PNJK_OUT_J | PNJK_OUT_K)))
    for (size_t i = 0; i < n; ++i) per_instance = per_instance || mo_tri_route(precision, g->eng[i]->nbf, nbf5[i]);
  if (per_instance) {
    for (size_t i = 0; i < n; ++i) {
      const BuildArgs a{C[i], ldc[i], nbf5[i], H ? H[i] : nullptr, ldh ? ldh[i] : 0, precision, outputs,
                        at(J, i), at(K, i), at(J_MO, i), at(K_MO, i), at(H_core, i)};
      for (const double* o : {a.J, a.K, a.J_MO, a.K_MO, a.H_core})
        if (o && !is_device_ptr(o))
          return set_error(PNJK_EINVAL, "pnjk_group_build: outputs of a grouped build must be device memory");
      if ((rc = jk_build_locked(g->eng[i], a, s, nullptr, 0))) return rc;
    }
    return PNJK_OK;
  }
  g->st.resize(n);
  // The plan (problem structs, tile tables) refers to the engines' buffers and the outputs;
  // C and H are only read by the per-call staging, so new C or H tensors (an optimiser's
  // next orbitals) reuse the plan and only invalidate a captured graph.
  std::vector<char> sig, in_sig;
  sig_put(sig, precision);
  sig_put(sig, outputs);
  for (size_t i = 0; i < n; ++i) {
    BuildArgs a{C[i], ldc[i], nbf5[i], H ? H[i] : nullptr, ldh ? ldh[i] : 0, precision, outputs,
                at(J, i), at(K, i), at(J_MO, i), at(K_MO, i), at(H_core, i)};
    if ((rc = build_prepare(g->eng[i], a, g->st[i]))) return rc;
    const BuildState& st = g->st[i];
    if (st.hJ || st.hK || (st.want_jmo && st.hJMO) || (st.want_kmo && st.hKMO) || st.hH)
      return set_error(PNJK_EINVAL, "pnjk_group_build: outputs of a grouped build must be device memory");
    if (use_graph && (!st.c_dev || (st.want_h && !st.h_dev)))
      return set_error(PNJK_EINVAL, "pnjk_group_build: graph mode needs device-resident C and H");
    if (precision != PNJK_FP64 && (rc = ensure_split_planes(g->eng[i], s))) return rc;
    for (const void* ptr : {(const void*)st.a.J, (const void*)st.a.K, (const void*)st.a.J_MO,
                            (const void*)st.a.K_MO, (const void*)st.a.H_core})
      sig_put(sig, ptr);
    sig_put(sig, st.a.nbf5);
    // The problem structs carry the instance's geometry: a b re-uploaded with another naux
    // into the same (grow-only) buffers keeps every pointer but needs a new plan.
    for (int64_t v : {(int64_t)st.g.nbf, (int64_t)st.g.naux, (int64_t)st.g.naux_p, (int64_t)st.g.nbf5,
                      st.g.ntri, (int64_t)st.g.ldcp, (int64_t)st.g.nbf_p, (int64_t)st.sg.ldk})
      sig_put(sig, v);
    sig_put(in_sig, st.a.C);
    sig_put(in_sig, st.a.ldc);
    sig_put(in_sig, st.a.H);
    sig_put(in_sig, st.a.ldh);
    sig_put(sig, st.ht);
    sig_put(sig, st.qt);
    sig_put(sig, st.jt);
    sig_put(sig, st.kt);
    for (int k = 0; k < g->eng[i]->nall; ++k) sig_put(sig, g->eng[i]->all[k]->p);
  }
  if (sig != g->sig) {
    // New plan: collect every instance's tiles per grouped stage, upload once.
    for (GroupCollector& c : g->plan) c.batches.clear();
    // Largest instances first: each XCD's run then dispatches big tiles before
    // small ones (longest-first balance); an instance's tiles stay adjacent.
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
      const BuildGeom &gx = g->st[x].g, &gy = g->st[y].g;
      return (double)gx.nbf * gx.nbf * gx.naux > (double)gy.nbf * gy.nbf * gy.naux;
    });
    // An instance whose half-transform alone has kSoloHtTiles tiles or more (four waves of
    // two workgroups on every CU) fills the GPU by itself; its HT runs as its own launch,
    // which measured faster than inside the grouped HT launch (the 64-instance batch, split
    // form: grouped HT 18.2 ms per pass against 15.8 ms for the same tiles launched per
    // instance). Every other stage stays grouped.
    constexpr int64_t kSoloHtTiles = 4 * 2 * 256;
    g->solo_ht.clear();
    for (size_t i : order) {
      pnjk_engine* e = g->eng[i];
      BuildState& st = g->st[i];
      const BuildGeom& bg = st.g;
      const int64_t ht_tiles = ((int64_t)bg.nbf5 + 32 * st.ht.fm - 1) / (32 * st.ht.fm) *
                               (((int64_t)bg.nbf * bg.naux_p + 32 * st.ht.fn - 1) / (32 * st.ht.fn));
      const bool solo = ht_tiles >= kSoloHtTiles;
      if (solo) g->solo_ht.push_back(i);
      if ((!solo && (rc = stage_ht(e, st, s, &g->plan[GS_HT]))) || (rc = stage_qt(e, st, s, &g->plan[GS_QT])) ||
          (rc = stage_jmo(e, st, s, &g->plan[GS_JMO])) || (rc = stage_jn(e, st, s, &g->plan[GS_JN])) ||
          (rc = stage_k(e, st, s, &g->plan[GS_K])))
        return rc;
    }
    // One symmetrize launch over every instance's J^n matrices.
    g->sym.clear();
    int64_t first = 0;
    for (size_t i = 0; i < n; ++i) {
      const BuildState& st = g->st[i];
      if (!st.want_j || st.g.nbf <= 1) continue;
      const int64_t nbf = st.g.nbf;
      g->sym.push_back(SymJob{st.dJ, nbf, nbf * nbf, first, (int)nbf, st.g.nbf5});
      first += (int64_t)st.g.nbf5 * sym_blocks((int)nbf);
    }
    g->sym_grid = first;
    size_t bytes = 0;
    g->sym_off = bytes;
    bytes = roundup((int64_t)(bytes + g->sym.size() * sizeof(SymJob)), 256);
    for (GroupCollector& c : g->plan)
      for (GroupBatch& b : c.batches) {
        b.probs_off = bytes;
        bytes = roundup((int64_t)(bytes + b.probs.size()), 256);
        b.tiles_off = bytes;
        bytes = roundup((int64_t)(bytes + b.tiles.size() * sizeof(int2)), 256);
        b.bounds_off = bytes;
        bytes = roundup((int64_t)(bytes + (kXcds + 1) * sizeof(int)), 256);
      }
    // This group's earlier builds may still read the plan buffer (a pooled buffer's previous
    // group had finished when it entered the pool); other work on the stream is no concern.
    if (g->done) PNJK_HIP_CHECK(hipEventSynchronize(g->done));
    if ((rc = g->plan_dev.ensure(std::max<size_t>(bytes, 256)))) return rc;
    char* base = static_cast<char*>(g->plan_dev.p);
    if (!g->sym.empty())
      PNJK_HIP_CHECK(hipMemcpy(base + g->sym_off, g->sym.data(), g->sym.size() * sizeof(SymJob),
                               hipMemcpyHostToDevice));
    for (GroupCollector& c : g->plan)
      for (GroupBatch& b : c.batches) {
        PNJK_HIP_CHECK(hipMemcpy(base + b.probs_off, b.probs.data(), b.probs.size(), hipMemcpyHostToDevice));
        PNJK_HIP_CHECK(hipMemcpy(base + b.tiles_off, b.tiles.data(), b.tiles.size() * sizeof(int2),
                                 hipMemcpyHostToDevice));
        const std::vector<int> xb = b.xcd_bounds();
        int longest = 0;
        for (int x = 0; x < kXcds; ++x) longest = std::max(longest, xb[x + 1] - xb[x]);
        b.grid = kXcds * longest;
        PNJK_HIP_CHECK(hipMemcpy(base + b.bounds_off, xb.data(), xb.size() * sizeof(int), hipMemcpyHostToDevice));
      }
    g->sig = sig;
    g->graph_valid = false;
  } else {
    for (size_t i = 0; i < n; ++i)  // K_MO partials per row pair: one per 64 aux columns, as
      g->st[i].nt = g->st[i].g.naux_p / 64;  // the quarter transform's launch sets it (any tile width)
  }
  if (in_sig != g->in_sig) {
    g->in_sig = in_sig;
    g->graph_valid = false;
  }
  // The event behind the build's launches tells pnjk_group_destroy when the plan buffer is free.
  auto mark = [&]() -> int {
    if (!g->done) PNJK_HIP_CHECK(hipEventCreateWithFlags(&g->done, hipEventDisableTiming));
    PNJK_HIP_CHECK(hipEventRecord(g->done, s));
    return PNJK_OK;
  };
  if (!use_graph) {
    if ((rc = group_enqueue(g, s))) return rc;
    return mark();
  }
  if (!g->graph_valid) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    g->exec = nullptr;
    g->graph = nullptr;
    PNJK_HIP_CHECK(hipStreamSynchronize(s));
    if (!g->cap) PNJK_HIP_CHECK(hipStreamCreateWithFlags(&g->cap, hipStreamNonBlocking));
    PNJK_HIP_CHECK(hipStreamBeginCapture(g->cap, hipStreamCaptureModeThreadLocal));
    rc = group_enqueue(g, g->cap);
    hipGraph_t graph = nullptr;
    const hipError_t he = hipStreamEndCapture(g->cap, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    PNJK_HIP_CHECK(he);
    g->graph = graph;
    PNJK_HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    g->graph_valid = true;
  }
  PNJK_HIP_CHECK(hipGraphLaunch(g->exec, s));
  return mark();
}

int pnjk_set_profiling(pnjk_engine* e, int on) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_set_profiling: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  if (on)
    for (hipEvent_t& ev : e->ev)
      if (!ev) PNJK_HIP_CHECK(hipEventCreate(&ev));
  e->prof = on != 0;
  e->prof_recorded = false;
  return PNJK_OK;
}

int pnjk_stage_count(void) { return ST_COUNT; }

int pnjk_set_tile(pnjk_engine* e, int arith, int stage, int fm, int fn) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_set_tile: engine is NULL");
  if (arith != 0 && arith != 1) return set_error(PNJK_EINVAL, "pnjk_set_tile: arith must be 0 (FP64) or 1 (BF16x3)");
  if (stage != ST_HT && stage != ST_QT && stage != ST_JN && stage != ST_K)
    return set_error(PNJK_EINVAL, "pnjk_set_tile: only half_transform, quarter_transform, coulomb_jn and "
                                  "exchange_kn take a tile shape");
  const bool reset = fm == 0 && fn == 0;
  bool known = false;
  for (const TileF64 t : {TileF64{2, 2}, TileF64{4, 4}, TileF64{5, 4}, TileF64{4, 2}, TileF64{2, 4}, TileF64{5, 2}})
    known = known || (t.fm == fm && t.fn == fn);
  // The eight-wave 320 x 128 tile exists for the BF16x3 half-transform and J^n only, the
  // 160 x 160 one for those and the BF16x3 exchange SYRK.
  if (arith == 1 && (stage == ST_HT || stage == ST_JN) && fm == 10 && fn == 4) known = true;
  if (arith == 1 && (stage == ST_HT || stage == ST_JN || stage == ST_K) && fm == 5 && fn == 5) known = true;
  if (!reset && !known) return set_error(PNJK_EINVAL, "pnjk_set_tile: unsupported tile shape");
  if (!reset && stage == ST_QT && fn != 2 && fn != 4)
    return set_error(PNJK_EINVAL, "pnjk_set_tile: the quarter transform needs fn = 2 or 4 (64 or 128 aux columns)");
  if (!reset && stage == ST_K && fm != fn)
    return set_error(PNJK_EINVAL, "pnjk_set_tile: the exchange SYRK needs square tiles (fm = fn)");
  std::lock_guard<std::mutex> lk(e->mu);
  e->tiles[arith][stage] = reset ? TileF64{0, 0} : TileF64{fm, fn};
  return PNJK_OK;
}

const char* pnjk_stage_name(int i) { return (i >= 0 && i < ST_COUNT) ? kStageNames[i] : ""; }

int pnjk_stage_times(pnjk_engine* e, double* ms) {
  if (!e || !ms) return set_error(PNJK_EINVAL, "pnjk_stage_times: null argument");
  std::lock_guard<std::mutex> lk(e->mu);
  if (!e->prof_recorded) return set_error(PNJK_ESTATE, "pnjk_stage_times: no profiled build recorded");
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  PNJK_HIP_CHECK(hipEventSynchronize(e->ev[ST_COUNT]));
  for (int i = 0; i < ST_COUNT; ++i) {
    float f = 0.f;
    PNJK_HIP_CHECK(hipEventElapsedTime(&f, e->ev[i], e->ev[i + 1]));
    ms[i] = f;
  }
  return PNJK_OK;
}

}  // extern "C"
