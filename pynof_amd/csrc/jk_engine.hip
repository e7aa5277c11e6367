// The C ABI of the MI355X RI JK build: engine state, b upload, build orchestration.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "jk_common.h"
#include "jk_kernels.h"
#include "pynof_amd.h"

namespace pnjk {

static thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {

inline int64_t roundup(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// Grow-only device buffer.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want) {
    if (want <= bytes) return PNJK_OK;
    if (p) {
      hipFree(p);
      p = nullptr;
      bytes = 0;
    }
    if (want == 0) return PNJK_OK;
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
      hipGetLastError();
      p = nullptr;
      return set_error(PNJK_ENOMEM, "hipMalloc(" + std::to_string(want) + " bytes) failed: " + hipGetErrorString(e));
    }
    bytes = want;
    return PNJK_OK;
  }
  void release() {
    if (p) hipFree(p);
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
    hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Built-in tile shapes (from autotuning at nbf=1000, naux=3600, nbf5=300 on MI355X;
// small problems keep 64 x 64 tiles so the grid fills the chip).
TileF64 default_tile(int arith, int stage_kind, int nbf, int nbf5) {
  // stage_kind: 0 HT, 1 QT, 2 JN, 3 K
  const bool big = nbf5 > 128 && nbf >= 256;
  if (arith == 0) {
    switch (stage_kind) {
      case 0: return big ? TileF64{5, 4} : TileF64{2, 2};
      case 1: return TileF64{2, 2};
      case 2: return big ? TileF64{2, 4} : TileF64{2, 2};
      default: return TileF64{2, 2};
    }
  }
  switch (stage_kind) {
    case 0: return big ? TileF64{4, 4} : TileF64{2, 2};
    case 1: return big ? TileF64{5, 2} : TileF64{2, 2};
    case 2: return big ? TileF64{4, 4} : TileF64{2, 2};
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
  std::mutex mu;
  bool prof = false;
  bool prof_recorded = false;
  hipEvent_t ev[ST_COUNT + 1] = {};
  int nbf = 0, naux = 0, naux_p = 0;
  int64_t ntri = 0;
  DevBuf btri;
  DevBuf bt16h, bt16l, bht16h, bht16l;  // split planes of b, built on the first split build
  DevBuf Cp, Hp, X, d, kpart, T, outs, stage, scal;
  DevBuf ct16h, ct16l, x16h, x16l, d16h, d16l;  // split workspace
  bool split_ready = false;
  // Tile override per arithmetic (0: FP64 kernels, 1: BF16x3 kernels) and stage;
  // {0, 0} selects the built-in choice.
  TileF64 tiles[2][ST_COUNT] = {};
  DevBuf* all[22];
  int nall = 0;
  pnjk_engine() {
    for (DevBuf* b : {&btri, &bt16h, &bt16l, &bht16h, &bht16l, &Cp, &Hp, &X, &d, &kpart, &T, &outs, &stage, &scal,
                      &ct16h, &ct16l, &x16h, &x16l, &d16h, &d16l})
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
  pnjk_engine* e = new pnjk_engine();
  e->device = device;
  hipError_t he = hipStreamCreateWithFlags(&e->own, hipStreamNonBlocking);
  if (he != hipSuccess) {
    delete e;
    return set_error(PNJK_EHIP, std::string("hipStreamCreate: ") + hipGetErrorString(he));
  }
  *out = e;
  return PNJK_OK;
}

int pnjk_destroy(pnjk_engine* e) {
  if (!e) return PNJK_OK;
  hipSetDevice(e->device);
  hipStreamSynchronize(e->own);
  for (int i = 0; i < e->nall; ++i) e->all[i]->release();
  for (hipEvent_t& ev : e->ev)
    if (ev) hipEventDestroy(ev);
  hipStreamDestroy(e->own);
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

uint64_t pnjk_workspace_bytes(int64_t nbf, int64_t naux, int64_t nbf5, int precision) {
  const int64_t naux_p = pnjk_packed_ld(nbf, naux);
  const int64_t nbf_p = roundup(nbf, 32), ldcp = roundup(nbf5, 32);
  uint64_t s = 0;
  s += (uint64_t)((nbf_p + 8) * ldcp + 1024) * 8;           // Cp
  s += (uint64_t)(nbf * nbf_p) * 8;                          // Hp
  s += (uint64_t)(nbf5 * nbf * naux_p) * 8;                  // X
  s += (uint64_t)(roundup(nbf5, 160) * naux_p) * 8;          // d
  s += (uint64_t)(nbf5 * nbf5 * (naux_p / 64)) * 8;          // kpart
  s += (uint64_t)(nbf * ldcp) * 8;                           // T
  if (precision != PNJK_FP64) {
    const int64_t ntri = nbf * (nbf + 1) / 2;
    s += (uint64_t)(ntri * naux_p) * 4;                      // btri16 hi/lo
    s += (uint64_t)(nbf * naux_p * nbf_p) * 4;               // bht16 hi/lo
    s += (uint64_t)(nbf5 * nbf_p) * 4;                       // Ct16 hi/lo
    s += (uint64_t)(nbf5 * nbf * naux_p) * 4;                // X16 hi/lo
    s += (uint64_t)(nbf5 * naux_p) * 4;                      // d16 hi/lo
  }
  return s;
}

// Builds the bf16 planes of b the split kernels read (once per uploaded b).
static int ensure_split_planes(pnjk_engine* e, hipStream_t s) {
  if (e->split_ready) return PNJK_OK;
  const int64_t ntri = e->ntri, naux_p = e->naux_p, nbf = e->nbf, ldk = roundup(nbf, 32);
  int rc;
  const size_t tri_bytes = (size_t)(ntri * naux_p) * 2, ht_bytes = (size_t)(nbf * naux_p * ldk) * 2;
  if ((rc = e->bt16h.ensure(tri_bytes)) || (rc = e->bt16l.ensure(tri_bytes))) return rc;
  if ((rc = e->bht16h.ensure(ht_bytes)) || (rc = e->bht16l.ensure(ht_bytes))) return rc;
  PNJK_HIP_CHECK(launch_split_btri(e->btri.as<double>(), ntri * naux_p, e->bt16h.as<uint16_t>(),
                                   e->bt16l.as<uint16_t>(), s));
  PNJK_HIP_CHECK(launch_build_bht(e->btri.as<double>(), (int)nbf, (int)naux_p, (int)ldk, e->bht16h.as<uint16_t>(),
                                  e->bht16l.as<uint16_t>(), s));
  e->split_ready = true;
  return PNJK_OK;
}

static int upload_common(pnjk_engine* e, int64_t nbf, int64_t naux) {
  e->split_ready = false;
  if (nbf < 1 || naux < 1) return set_error(PNJK_EINVAL, "upload_b: nbf and naux must be >= 1");
  if (nbf > 46000) return set_error(PNJK_EINVAL, "upload_b: nbf too large");
  e->nbf = (int)nbf;
  e->naux = (int)naux;
  e->naux_p = (int)pnjk_packed_ld(nbf, naux);
  e->ntri = nbf * (nbf + 1) / 2;
  if (e->ntri * (int64_t)e->naux_p > (int64_t)1 << 40) return set_error(PNJK_EINVAL, "upload_b: b too large");
  return e->btri.ensure((size_t)(e->ntri * e->naux_p) * sizeof(double));
}

int pnjk_upload_b(pnjk_engine* e, const double* b, int64_t nbf, int64_t naux, double sym_tol, uint64_t* footprint,
                  double* asym) {
  if (!e || !b) return set_error(PNJK_EINVAL, "pnjk_upload_b: null argument");
  std::lock_guard<std::mutex> lk(e->mu);
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

int pnjk_jk_build(pnjk_engine* e, const double* C, int64_t ldc, int64_t nbf5, const double* H, int64_t ldh,
                  int precision, uint32_t outputs, double* J, double* K, double* J_MO, double* K_MO, double* H_core,
                  void* stream) {
  if (!e) return set_error(PNJK_EINVAL, "pnjk_jk_build: engine is NULL");
  std::lock_guard<std::mutex> lk(e->mu);
  if (e->nbf == 0) return set_error(PNJK_ESTATE, "pnjk_jk_build: no b uploaded");
  if (!C) return set_error(PNJK_EINVAL, "pnjk_jk_build: C is NULL");
  if (nbf5 < 1 || nbf5 > e->nbf)
    return set_error(PNJK_EINVAL, "pnjk_jk_build: nbf5 must be in [1, nbf]; got " + std::to_string(nbf5));
  if (ldc < nbf5) return set_error(PNJK_EINVAL, "pnjk_jk_build: ldc < nbf5");
  if (outputs & ~PNJK_OUT_ALL) return set_error(PNJK_EINVAL, "pnjk_jk_build: unknown output bits");
  if ((outputs & PNJK_OUT_J) && !J) return set_error(PNJK_EINVAL, "pnjk_jk_build: J requested but NULL");
  if ((outputs & PNJK_OUT_K) && !K) return set_error(PNJK_EINVAL, "pnjk_jk_build: K requested but NULL");
  if ((outputs & PNJK_OUT_JMO) && !J_MO) return set_error(PNJK_EINVAL, "pnjk_jk_build: J_MO requested but NULL");
  if ((outputs & PNJK_OUT_KMO) && !K_MO) return set_error(PNJK_EINVAL, "pnjk_jk_build: K_MO requested but NULL");
  if ((outputs & PNJK_OUT_HCORE) && (!H_core || !H))
    return set_error(PNJK_EINVAL, "pnjk_jk_build: H_core requested but H or H_core is NULL");
  if ((outputs & PNJK_OUT_HCORE) && ldh < e->nbf) return set_error(PNJK_EINVAL, "pnjk_jk_build: ldh < nbf");
  if (precision != PNJK_FP64 && precision != PNJK_SPLIT && precision != PNJK_SPLIT_REFINE)
    return set_error(PNJK_EUNSUPPORTED, "pnjk_jk_build: unknown precision " + std::to_string(precision));
  PNJK_HIP_CHECK(hipSetDevice(e->device));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : e->own;
  auto mark = [&](int st) {
    if (e->prof) hipEventRecord(e->ev[st], s);
  };

  BuildGeom g;
  g.nbf = e->nbf;
  g.naux = e->naux;
  g.naux_p = e->naux_p;
  g.nbf5 = (int)nbf5;
  g.ntri = e->ntri;
  g.nbf_p = (int)roundup(e->nbf, 32);
  g.ldcp = (int)roundup(nbf5, 32);
  const int64_t nbf = g.nbf, n5 = nbf5;

  int rc;
  const int64_t cp_elems = (int64_t)(g.nbf_p + 8) * g.ldcp + 1024;
  if ((rc = e->Cp.ensure((size_t)cp_elems * 8))) return rc;
  if (precision != PNJK_SPLIT && (rc = e->X.ensure((size_t)(n5 * nbf * g.naux_p) * 8))) return rc;
  const int64_t d_rows = roundup(n5, 160);
  if ((rc = e->d.ensure((size_t)(d_rows * g.naux_p) * 8))) return rc;
  const int qt_bn = 64;
  if ((rc = e->kpart.ensure((size_t)(n5 * n5 * (g.naux_p / qt_bn)) * 8))) return rc;

  // Output staging for host destinations.
  const bool want_j = outputs & PNJK_OUT_J, want_k = outputs & PNJK_OUT_K, want_jmo = outputs & PNJK_OUT_JMO,
             want_kmo = outputs & PNJK_OUT_KMO, want_h = outputs & PNJK_OUT_HCORE;
  const int64_t jk_elems = n5 * nbf * nbf;
  double* dJ = J;
  double* dK = K;
  double* dJMO = J_MO;
  double* dKMO = K_MO;
  double* dH = H_core;
  const bool hJ = want_j && !is_device_ptr(J), hK = want_k && !is_device_ptr(K);
  const bool hJMO = want_jmo && !is_device_ptr(J_MO), hKMO = !want_kmo || !is_device_ptr(K_MO);
  const bool hH = want_h && !is_device_ptr(H_core);
  int64_t stage_elems = 0;
  const int64_t off_J = stage_elems;
  if (hJ) stage_elems += jk_elems;
  const int64_t off_K = stage_elems;
  if (hK) stage_elems += jk_elems;
  const int64_t off_JMO = stage_elems;
  stage_elems += n5 * n5;  // J_MO is always formed (cheap) for the energy path
  const int64_t off_KMO = stage_elems;
  if (hKMO) stage_elems += n5 * n5;
  const int64_t off_H = stage_elems;
  if (hH) stage_elems += n5;
  if ((rc = e->outs.ensure((size_t)std::max<int64_t>(stage_elems, 1) * 8))) return rc;
  double* ob = e->outs.as<double>();
  if (hJ) dJ = ob + off_J;
  if (hK) dK = ob + off_K;
  if (!want_jmo || hJMO) dJMO = ob + off_JMO;
  if (hKMO) dKMO = ob + off_KMO;
  if (hH) dH = ob + off_H;

  // C -> Cp[nbf_p + slack][ldcp], zero elsewhere.
  mark(ST_IN);
  double* Cp = e->Cp.as<double>();
  PNJK_HIP_CHECK(fill(Cp, cp_elems, 0.0, s));
  if (is_device_ptr(C)) {
    const int64_t n = nbf * n5;
    hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, C, ldc, Cp, (int64_t)g.ldcp,
                       (int)nbf, (int)n5);
    PNJK_HIP_CHECK(hipGetLastError());
  } else {
    PNJK_HIP_CHECK(hipMemcpy2DAsync(Cp, (size_t)g.ldcp * 8, C, (size_t)ldc * 8, (size_t)n5 * 8, (size_t)nbf,
                                    hipMemcpyHostToDevice, s));
  }

  // FP64 kernels run HT/QT in the fp64 and split_refine forms, J^n/K^n in fp64 only.
  const int a_mo = precision == PNJK_SPLIT ? 1 : 0, a_jk = precision == PNJK_FP64 ? 0 : 1;
  auto chosen = [&](int arith, int st, int kind) {
    return e->tiles[arith][st].fm ? e->tiles[arith][st] : default_tile(arith, kind, g.nbf, g.nbf5);
  };
  const TileF64 ht = chosen(a_mo, ST_HT, 0);
  const TileF64 qt = chosen(a_mo, ST_QT, 1);
  const TileF64 jt = chosen(a_jk, ST_JN, 2);
  const TileF64 kt = chosen(a_jk, ST_K, 3);
  int nt = 0;
  SplitGeom sg;
  sg.nbf = g.nbf;
  sg.naux_p = g.naux_p;
  sg.nbf5 = g.nbf5;
  sg.ldk = g.nbf_p;
  if (precision != PNJK_FP64) {
    if ((rc = ensure_split_planes(e, s))) return rc;
    const size_t ct_bytes = (size_t)(n5 * sg.ldk) * 2, x_bytes = (size_t)(n5 * nbf * g.naux_p) * 2,
                 d_bytes = (size_t)(n5 * g.naux_p) * 2;
    if ((rc = e->ct16h.ensure(ct_bytes)) || (rc = e->ct16l.ensure(ct_bytes))) return rc;
    if ((rc = e->x16h.ensure(x_bytes)) || (rc = e->x16l.ensure(x_bytes))) return rc;
    if ((rc = e->d16h.ensure(d_bytes)) || (rc = e->d16l.ensure(d_bytes))) return rc;
  }
  uint16_t *ct_h = e->ct16h.as<uint16_t>(), *ct_l = e->ct16l.as<uint16_t>();
  uint16_t *x_h = e->x16h.as<uint16_t>(), *x_l = e->x16l.as<uint16_t>();
  uint16_t *d_h = e->d16h.as<uint16_t>(), *d_l = e->d16l.as<uint16_t>();
  mark(ST_HT);
  if (precision == PNJK_SPLIT) {
    PNJK_HIP_CHECK(launch_split_ct(Cp, g.ldcp, g.nbf, g.nbf5, g.nbf5, sg.ldk, ct_h, ct_l, s));
    PNJK_HIP_CHECK(launch_ht_x3(sg, ht, ct_h, ct_l, e->bht16h.as<uint16_t>(), e->bht16l.as<uint16_t>(), x_h, x_l,
                                s));
  } else if (precision == PNJK_SPLIT_REFINE && want_k) {
    // FP64 X for the MO path, with its split planes for K^n written by the same epilogue.
    PNJK_HIP_CHECK(launch_ht_f64(g, ht, Cp, e->btri.as<double>(), e->X.as<double>(), s, x_h, x_l));
  } else {
    PNJK_HIP_CHECK(launch_ht_f64(g, ht, Cp, e->btri.as<double>(), e->X.as<double>(), s));
  }
  mark(ST_QT);
  if (precision == PNJK_SPLIT) {
    PNJK_HIP_CHECK(launch_qt_x3(sg, qt, ct_h, ct_l, x_h, x_l, e->kpart.as<double>(), e->d.as<double>(), d_h, d_l,
                                &nt, s));
  } else {
    PNJK_HIP_CHECK(
        launch_qt_f64(g, qt, Cp, e->X.as<double>(), e->kpart.as<double>(), e->d.as<double>(), &nt, s));
  }
  mark(ST_KMO);
  PNJK_HIP_CHECK(launch_kmo_reduce(g.nbf5, nt, e->kpart.as<double>(), dKMO, s));
  mark(ST_JMO);
  PNJK_HIP_CHECK(launch_syrk_f64(1, g.nbf5, g.naux_p, e->d.as<double>(), g.naux_p, 0, dJMO, g.nbf5, 0,
                                 TileF64{2, 2}, s));
  PNJK_HIP_CHECK(launch_symmetrize(1, g.nbf5, dJMO, g.nbf5, 0, s));
  if (precision == PNJK_SPLIT_REFINE && want_j) {
    // J^n in split form from the exact d.
    PNJK_HIP_CHECK(launch_split_btri(e->d.as<double>(), n5 * g.naux_p, d_h, d_l, s));
  }
  mark(ST_JN);
  if (want_j) {
    if (precision == PNJK_FP64)
      PNJK_HIP_CHECK(launch_jn_f64(g, jt, e->d.as<double>(), e->btri.as<double>(), dJ, s));
    else
      PNJK_HIP_CHECK(launch_jn_x3(sg, jt, d_h, d_l, e->bt16h.as<uint16_t>(), e->bt16l.as<uint16_t>(), dJ, s));
    PNJK_HIP_CHECK(launch_symmetrize(g.nbf5, g.nbf, dJ, nbf, nbf * nbf, s));
  }
  mark(ST_K);
  if (want_k) {
    if (precision == PNJK_FP64)
      PNJK_HIP_CHECK(launch_syrk_f64(g.nbf5, g.nbf, g.naux_p, e->X.as<double>(), g.naux_p, nbf * g.naux_p, dK, nbf,
                                     nbf * nbf, kt, s));
    else
      PNJK_HIP_CHECK(launch_kn_x3(sg, kt, x_h, x_l, dK, s));
    PNJK_HIP_CHECK(launch_symmetrize(g.nbf5, g.nbf, dK, nbf, nbf * nbf, s));
  }
  mark(ST_HC);
  if (want_h) {
    if ((rc = e->Hp.ensure((size_t)(nbf * g.nbf_p) * 8))) return rc;
    if ((rc = e->T.ensure((size_t)(nbf * g.ldcp) * 8))) return rc;
    double* Hp = e->Hp.as<double>();
    PNJK_HIP_CHECK(fill(Hp, nbf * g.nbf_p, 0.0, s));
    if (is_device_ptr(H)) {
      const int64_t n = nbf * nbf;
      hipLaunchKernelGGL(k_copy2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, H, ldh, Hp,
                         (int64_t)g.nbf_p, (int)nbf, (int)nbf);
      PNJK_HIP_CHECK(hipGetLastError());
    } else {
      PNJK_HIP_CHECK(hipMemcpy2DAsync(Hp, (size_t)g.nbf_p * 8, H, (size_t)ldh * 8, (size_t)nbf * 8, (size_t)nbf,
                                      hipMemcpyHostToDevice, s));
    }
    PNJK_HIP_CHECK(launch_hcore_f64(g, Hp, g.nbf_p, Cp, e->T.as<double>(), dH, s));
  }

  mark(ST_OUT);
  bool sync = false;
  if (hJ) {
    PNJK_HIP_CHECK(hipMemcpyAsync(J, dJ, (size_t)jk_elems * 8, hipMemcpyDeviceToHost, s));
    sync = true;
  }
  if (hK) {
    PNJK_HIP_CHECK(hipMemcpyAsync(K, dK, (size_t)jk_elems * 8, hipMemcpyDeviceToHost, s));
    sync = true;
  }
  if (want_jmo && hJMO) {
    PNJK_HIP_CHECK(hipMemcpyAsync(J_MO, dJMO, (size_t)(n5 * n5) * 8, hipMemcpyDeviceToHost, s));
    sync = true;
  }
  if (want_kmo && hKMO) {
    PNJK_HIP_CHECK(hipMemcpyAsync(K_MO, dKMO, (size_t)(n5 * n5) * 8, hipMemcpyDeviceToHost, s));
    sync = true;
  }
  if (hH) {
    PNJK_HIP_CHECK(hipMemcpyAsync(H_core, dH, (size_t)n5 * 8, hipMemcpyDeviceToHost, s));
    sync = true;
  }
  mark(ST_COUNT);
  if (e->prof) e->prof_recorded = true;
  if (sync) PNJK_HIP_CHECK(hipStreamSynchronize(s));
  return PNJK_OK;
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
  if (!reset && !known) return set_error(PNJK_EINVAL, "pnjk_set_tile: unsupported tile shape");
  if (!reset && stage == ST_QT && fn != 2)
    return set_error(PNJK_EINVAL, "pnjk_set_tile: the quarter transform needs fn = 2 (64 aux columns per tile)");
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
