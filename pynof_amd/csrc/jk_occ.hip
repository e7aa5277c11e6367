// The occupation step of a PNOF orbital optimisation on the device: occupations
// from the softmax or trigonometric parameters, the PNOF5 / PNOF7 / GNOF
// coefficient matrices C^J, C^K, the closed-shell energy and its gradient with
// respect to the occupation parameters (src/pynof/pnof.py:155-525, 692-810,
// 905-1025). Everything is O(nbf5^2) or O(nbf5^2 nv) scalar FP64 work on
// nbf5 x nbf5 matrices that stay in L2; the kernels are latency-bound, and the
// point is that J_MO, K_MO and H_core never leave the GPU between a JK build
// and the occupation update.
#include "jk_common.h"
#include "jk_kernels.h"

namespace pnjk {
namespace {

constexpr double kHCut = 0.028284271247461901;  // 0.02 sqrt(2), pnof.py:355
constexpr double kTiny = 1e-15;                 // derivative guards, pnof.py:229-231, :308, :456

// Subspace of orbital i: strong orbital no1 + g, or weak orbital
// no1 + ndns + (ndoc - g - 1) + ndoc t (pnof.py:187-190); -1 for the rest.
__device__ __forceinline__ int subspace(const PnofDev& p, int i, bool& strong) {
  strong = i >= p.no1 && i < p.no1 + p.ndoc;
  if (strong) return i - p.no1;
  const int r = i - p.no1 - p.ndns;
  if (r >= 0 && r < p.ndoc * p.ncwo) return p.ndoc - 1 - r % p.ndoc;
  return -1;
}

__device__ __forceinline__ int weak_index(const PnofDev& p, int g, int t) {
  return p.no1 + p.ndns + (p.ndoc - g - 1) + p.ndoc * t;
}

// Orbital ranges of the GNOF blocks: O = [no1, nbeta), S = [nbeta, nalpha), V = [nalpha, nbf5).
__device__ __forceinline__ bool in_o(const PnofDev& p, int i) { return i >= p.no1 && i < p.nbeta; }
__device__ __forceinline__ bool in_s(const PnofDev& p, int i) { return i >= p.nbeta && i < p.nalpha; }
__device__ __forceinline__ bool in_v(const PnofDev& p, int i) { return i >= p.nalpha; }

// C^J[i][j], C^K[i][j] (pnof.py:155-202, :248-289, :350-422), from n and the per-orbital
// Phi (fi), n_d and sqrt(n_d) (nd, nd12).
__device__ void coef(const PnofDev& p, const double* __restrict__ n, const double* __restrict__ fi,
                     const double* __restrict__ nd, const double* __restrict__ nd12, int i, int j, double& cj,
                     double& ck) {
  const double ni = n[i], nj = n[j];
  cj = 2.0 * ni * nj;
  ck = ni * nj;
  const bool ss = in_s(p, i) && in_s(p, j);
  if (p.ipnof == 7) ck = ni * nj + fi[i] * fi[j];
  if (p.ipnof != 8) {
    if (p.mspin == 0 && p.nsoc > 1 && ss) ck = 2.0 * ni * nj;
  } else {
    const bool oi = in_o(p, i), vi = in_v(p, i), oj = in_o(p, j), vj = in_v(p, j);
    if ((oi && vj) || (vi && oj) || (vi && vj)) ck += fi[i] * fi[j];
    if (p.mspin == 0 && p.nsoc > 0) {
      if (oi && in_s(p, j)) ck += 0.5 * (fi[i] * 0.5);
      if (in_s(p, i) && oj) ck += 0.5 * (0.5 * fi[j]);
      if (in_s(p, i) && vj) ck += 0.5 * fi[j];
      if (vi && in_s(p, j)) ck += fi[i] * 0.5;
    }
    if (p.mspin == 0 && p.nsoc > 1 && ss) ck = 0.5;
    if ((oi && vj) || (vi && oj)) ck += nd12[i] * nd12[j] - nd[i] * nd[j];
    if (vi && vj) ck += -nd12[i] * nd12[j] - nd[i] * nd[j];
  }
  bool si, sj;
  const int gi = subspace(p, i, si), gj = subspace(p, j, sj);
  if (gi >= 0 && gi == gj && !(si && sj)) {  // intra-pair blocks (pnof.py:184-200)
    cj = 0.0;
    if (si) ck = sqrt(ni * nj);
    else if (sj) ck = sqrt(nj * ni);
    else ck = -sqrt(ni * nj);
  }
}

// Derivative of C^J[i][j], C^K[i][j] with respect to parameter k through the row
// orbital i, as der_CJCKD5/7/8 form them (pnof.py:206-245, :291-347, :425-525).
// Orbital-major arrays: dn, dfi, dnd, dnd12 at [i * nv + k].
struct Der {
  const double* n;
  const double* fi;
  const double* nd;
  const double* nd12;
  const double* dn;
  const double* dfi;
  const double* dnd;
  const double* dnd12;
};

__device__ void dcoef(const PnofDev& p, const Der& d, int i, int j, int k, double& dcj, double& dck) {
  const int64_t ik = (int64_t)i * p.nv + k;
  const double dni = d.dn[ik];
  dcj = 2.0 * dni * d.n[j];
  dck = dni * d.n[j];
  if (p.ipnof == 7) dck = dni * d.n[j] + d.dfi[ik] * d.fi[j];
  if (p.ipnof == 8) {
    const bool oi = in_o(p, i), vi = in_v(p, i), oj = in_o(p, j), vj = in_v(p, j);
    const double dfi = d.dfi[ik];
    if (oi && vj) dck += dfi * d.fi[j];
    if (vi && (oj || vj)) dck += d.fi[j] * dfi;
    if (p.mspin == 0 && p.nsoc > 0 && in_s(p, j)) {
      if (oi) dck += 0.5 * dfi * 0.5;
      if (vi) dck += dfi * 0.5;
    }
    if (p.mspin == 0 && p.nsoc > 1 && in_s(p, i) && in_s(p, j)) dck = 0.0;
    if (oi && vj) dck += d.dnd12[ik] * d.nd12[j] - d.dnd[ik] * d.nd[j];
    if (vi && oj) dck += d.nd12[j] * d.dnd12[ik] - d.nd[j] * d.dnd[ik];
    if (vi && vj) dck += -d.nd12[j] * d.dnd12[ik] - d.nd[j] * d.dnd[ik];
  }
  bool si, sj;
  const int gi = subspace(p, i, si), gj = subspace(p, j, sj);
  if (gi >= 0 && gi == gj && !(si && sj)) {  // intra-pair blocks (pnof.py:220-243)
    dcj = 0.0;
    const double a = d.n[i] > kTiny ? d.n[i] : kTiny;
    if (si || sj) dck = 0.5 * 1.0 / sqrt(a) * dni * sqrt(d.n[j]);
    else dck = -0.5 * (1.0 / sqrt(a) * dni) * sqrt(d.n[j]);
  }
}

// Block-wide sum of one value per thread in a fixed order (blockDim.x = 256).
__device__ double block_sum256(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// ---------------------------------------------------------------- occupations
// One workgroup: n (nbf5) and dn (nbf5 x nv) from the parameters (pnof.py:704-810).
__global__ void __launch_bounds__(256) k_occ_n(PnofDev p, const double* __restrict__ x, double* __restrict__ n,
                                               double* __restrict__ dn) {
  const int nv = p.nv;
  for (int64_t i = threadIdx.x; i < (int64_t)p.nbf5 * nv; i += blockDim.x) dn[i] = 0.0;
  for (int i = threadIdx.x; i < p.nbf5; i += blockDim.x) {
    double v = 0.0;
    if (p.method == 1 && i < p.no1) v = 1.0;
    if (i >= p.no1 + p.ndoc && i < p.no1 + p.ndns) v = p.high_spin ? 1.0 : 0.5;
    if (p.method == 1 && i >= p.no1 && i < p.no1 + p.ndoc) {
      const double c = cos(x[i - p.no1]);
      v = 0.5 * (1.0 + c * c);
    }
    n[i] = v;
  }
  __syncthreads();
  for (int g = threadIdx.x; g < p.ndoc; g += blockDim.x) {
    const int m = p.ncwo;
    if (p.method == 0) {
      // softmax (pnof.py:783-808): the strong occupation and its derivative row sit at
      // index g, as in the reference
      double s = 0.0;
      for (int t = 0; t < m; ++t) s += exp(x[(p.ndoc - g - 1) + p.ndoc * t]);
      const double den = 1.0 + s;
      const double ns = 1.0 / den;
      n[g] = ns;
      for (int t = 0; t < m; ++t) {
        const int ct = (p.ndoc - g - 1) + p.ndoc * t;
        const double nw = exp(x[ct]) / den;
        const int w = weak_index(p, g, t);
        n[w] = nw;
        dn[(int64_t)g * nv + ct] = -ns * nw;
        for (int u = 0; u < m; ++u) {
          const int cu = (p.ndoc - g - 1) + p.ndoc * u;
          const double nu = exp(x[cu]) / den;
          dn[(int64_t)w * nv + cu] = -nw * nu + (u == t ? nw : 0.0);
        }
      }
    } else {
      // trigonometric (pnof.py:717-766): weak t takes hole * prod_{s<t} cos^2 g_s * sin^2 g_t
      const int s0 = p.no1 + g;
      const double hole = 1.0 - n[s0];
      const double dstrong = -0.5 * sin(2.0 * x[g]);
      dn[(int64_t)s0 * nv + g] = dstrong;
      const int gc0 = p.ndoc + (p.ndoc - g - 1);  // column of the first weak parameter
      double pre = 1.0;                           // prod_{s<t} cos^2
      for (int t = 0; t < m; ++t) {
        const int w = weak_index(p, g, t);
        const double gt = t < m - 1 ? x[gc0 + p.ndoc * t] : 0.0;
        const double st = sin(gt);
        const double last = t < m - 1 ? st * st : 1.0;
        n[w] = hole * pre * last;
        dn[(int64_t)w * nv + g] = -dstrong * pre * last;
        if (t < m - 1) dn[(int64_t)w * nv + gc0 + p.ndoc * t] = hole * pre * sin(2.0 * gt);
        const double ct = cos(gt);
        pre *= ct * ct;
      }
      // d/dg_j of the factor cos^2 g_j for rows t > j
      double prej = 1.0;  // prod_{s<j} cos^2
      for (int j = 0; j + 1 < m; ++j) {
        const double gj = x[gc0 + p.ndoc * j];
        double r = -hole * prej * sin(2.0 * gj);
        for (int t = j + 1; t < m; ++t) {
          const double gt = t < m - 1 ? x[gc0 + p.ndoc * t] : 0.0;
          const double st = sin(gt);
          const double last = t < m - 1 ? st * st : 1.0;
          dn[(int64_t)weak_index(p, g, t) * nv + gc0 + p.ndoc * j] = r * last;
          const double ct = cos(gt);
          r *= ct * ct;
        }
        const double cj = cos(gj);
        prej *= cj * cj;
      }
    }
  }
}

// Per-orbital Phi, n_d, sqrt(n_d) (pnof.py:252-257, :355-366).
__global__ void k_occ_orb(PnofDev p, const double* __restrict__ n, double* __restrict__ fi, double* __restrict__ nd,
                          double* __restrict__ nd12) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.nbf5) return;
  const double ni = n[i];
  double f;
  if (p.ipnof == 7 && p.ista != 0) {
    f = 2.0 * ni * (1.0 - ni);
  } else {
    const double q = ni * (1.0 - ni);
    f = q <= 0.0 ? 0.0 : sqrt(q);
  }
  fi[i] = f;
  bool strong;
  const int g = subspace(p, i, strong);
  double v = 0.0;
  if (g >= 0) {
    const double coc = (1.0 - n[p.no1 + g]) / kHCut;
    v = ni * exp(-(coc * coc));
  }
  nd[i] = v;
  nd12[i] = sqrt(v);
}

// Per-orbital derivatives dPhi, dn_d, dsqrt(n_d) at [i * nv + k] (pnof.py:301-309, :433-465).
__global__ void k_occ_orb_der(PnofDev p, const double* __restrict__ n, const double* __restrict__ dn,
                              const double* __restrict__ fi, const double* __restrict__ nd12,
                              double* __restrict__ dfi, double* __restrict__ dnd, double* __restrict__ dnd12) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)p.nbf5 * p.nv) return;
  const int i = (int)(idx / p.nv), k = (int)(idx % p.nv);
  const double ni = n[i], dni = dn[idx];
  double v = 0.0;
  if (i >= p.no1) {
    if (p.ipnof == 7 && p.ista != 0) {
      v = 2.0 * (1.0 - 2.0 * ni) * dni;
    } else {
      const double a = fi[i] > kTiny ? fi[i] : kTiny;
      v = 1.0 / (2.0 * a) * (1.0 - 2.0 * ni) * dni;
    }
  }
  dfi[idx] = v;
  bool strong;
  const int g = subspace(p, i, strong);
  double w = 0.0;
  if (p.ipnof == 8 && g >= 0) {
    const int s = p.no1 + g;
    const double ns = n[s];
    const double coc = (1.0 - ns) / kHCut;
    const double F = exp(-(coc * coc));
    const double dns = dn[(int64_t)s * p.nv + k];
    w = strong ? F * dns * (1.0 - ns * (-2.0 * coc / kHCut)) : F * (dni - (-2.0 * coc / kHCut) * (ni * dns));
  }
  dnd[idx] = w;
  dnd12[idx] = 0.5 * w / (nd12[i] > kTiny ? nd12[i] : kTiny);
}

__global__ void k_occ_coef(PnofDev p, const double* __restrict__ n, const double* __restrict__ fi,
                           const double* __restrict__ nd, const double* __restrict__ nd12, double* __restrict__ cj,
                           double* __restrict__ ck) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)p.nbf5 * p.nbf5) return;
  const int i = (int)(idx / p.nbf5), j = (int)(idx % p.nbf5);
  double a, b;
  coef(p, n, fi, nd, nd12, i, j, a, b);
  if (cj) cj[idx] = a;
  if (ck) ck[idx] = b;
}

// Transposed copy of an nbf5 x nbf5 matrix: t[i][j] = s[j][i].
__global__ void k_occ_transpose(const double* __restrict__ s, int n, double* __restrict__ t) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx % n);
  t[idx] = s[(int64_t)j * n + i];
}

// Energy (calce, pnof.py:905-988) in two passes. Pass 1: workgroup b takes rows
// i = 4b .. 4b + 3, one wave per row; lane sums j = lane + 64 t of C^J[i][j] J[j][i] -
// C^K[i][j] K[j][i] (j != i; with MSpin != 0 only i, j outside the singly occupied block
// S) plus, on lane 0, the row's one-body terms and, with MSpin != 0, its n-weighted sums
// over S; the workgroup's sum goes to part[b]. Pass 2 sums the parts in order. JT, KT are
// J_MO, K_MO transposed (JT[i][j] = J_MO[j][i]); with cjm / ckm null, C^J, C^K come from
// coef().
__global__ void __launch_bounds__(256) k_occ_energy_rows(PnofDev p, const double* __restrict__ n,
                                                         const double* __restrict__ fi, const double* __restrict__ nd,
                                                         const double* __restrict__ nd12,
                                                         const double* __restrict__ cjm, const double* __restrict__ ckm,
                                                         const double* __restrict__ JT, const double* __restrict__ KT,
                                                         const double* __restrict__ hc, double* __restrict__ part) {
  __shared__ double red[4];
  const int n5 = p.nbf5, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  double acc = 0.0;
  if (i < n5) {
    const bool si = in_s(p, i), open = p.mspin != 0;
    if (lane == 0) {
      acc = 2.0 * n[i] * hc[i];
      if (!si) acc += n[i] * JT[(int64_t)i * n5 + i];
      if (open) {  // pnof.py:966-988
        double sj = 0.0, sk = 0.0;
        for (int t = p.nbeta; t < p.nalpha; ++t) {
          sj += JT[(int64_t)i * n5 + t];
          sk += KT[(int64_t)i * n5 + t];
        }
        if (!si)
          acc += 2.0 * n[i] * sj - n[i] * sk;
        else
          acc += 0.5 * (n[i] * sj - n[i] * JT[(int64_t)i * n5 + i]) - 0.5 * (n[i] * sk + n[i] * KT[(int64_t)i * n5 + i]);
      }
    }
    for (int j = lane; j < n5; j += 64) {
      if (j == i || (open && (si || in_s(p, j)))) continue;
      const int64_t idx = (int64_t)i * n5 + j;
      double a, b;
      if (cjm) {
        a = cjm[idx];
        b = ckm[idx];
      } else {
        coef(p, n, fi, nd, nd12, i, j, a, b);
      }
      acc += a * JT[idx] - b * KT[idx];
    }
  }
  const double tot = block_sum256(acc, red);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256) k_occ_energy_sum(const double* __restrict__ part, int nparts,
                                                        double* __restrict__ E) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) acc += part[b];
  const double tot = block_sum256(acc, red);
  if (threadIdx.x == 0) E[0] = tot;
}

// Occupation gradient (calcoccg, pnof.py:1004-1098): one workgroup per parameter k.
// Row i contributes only where one of its derivative factors at k is nonzero; every
// term of dC[i][j] carries one of them, so skipped rows add exact zeros.
__global__ void __launch_bounds__(256) k_occ_grad(PnofDev p, Der d, const double* __restrict__ JT,
                                                  const double* __restrict__ KT, const double* __restrict__ hc,
                                                  double* __restrict__ grad) {
  __shared__ double red[4];
  const int n5 = p.nbf5, k = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool open = p.mspin != 0;
  double acc = 0.0;
  for (int i = threadIdx.x; i < n5; i += blockDim.x) {
    // MSpin != 0 keeps rows [no1, nbeta) and [nalpha, nbf5) only (pnof.py:1029-1098)
    if (open && (i < p.no1 || in_s(p, i))) continue;
    const double dni = d.dn[(int64_t)i * p.nv + k];
    acc += dni * (2.0 * hc[i] + JT[(int64_t)i * n5 + i]);
    if (open && dni != 0.0) {
      double sj = 0.0, sk = 0.0;
      for (int t = p.nbeta; t < p.nalpha; ++t) {
        sj += JT[(int64_t)i * n5 + t];
        sk += KT[(int64_t)i * n5 + t];
      }
      acc += 2.0 * dni * sj - dni * sk;
    }
  }
  for (int i = wave; i < n5; i += 4) {
    if (open && (i < p.no1 || in_s(p, i))) continue;
    const int64_t ik = (int64_t)i * p.nv + k;
    bool live = d.dn[ik] != 0.0;
    if (p.ipnof >= 7) live = live || d.dfi[ik] != 0.0;
    if (p.ipnof == 8) live = live || d.dnd[ik] != 0.0 || d.dnd12[ik] != 0.0;
    if (!live) continue;
    for (int j = lane; j < n5; j += 64) {
      if (j == i || (open && in_s(p, j))) continue;
      double a, b;
      dcoef(p, d, i, j, k, a, b);
      acc += 2.0 * a * JT[(int64_t)i * n5 + j] - 2.0 * b * KT[(int64_t)i * n5 + j];
    }
  }
  const double tot = block_sum256(acc, red);
  if (threadIdx.x == 0) grad[k] = tot;
}

inline unsigned blocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

hipError_t launch_occupation(const PnofDev& p, const double* x, double* n, double* dn, hipStream_t s) {
  hipLaunchKernelGGL(k_occ_n, dim3(1), dim3(256), 0, s, p, x, n, dn);
  return hipGetLastError();
}

hipError_t launch_pnof_orbitals(const PnofDev& p, const double* n, const double* dn, double* fi, double* nd,
                                double* nd12, double* dfi, double* dnd, double* dnd12, hipStream_t s) {
  hipLaunchKernelGGL(k_occ_orb, dim3(blocks(p.nbf5, 256)), dim3(256), 0, s, p, n, fi, nd, nd12);
  if (dn && p.nv > 0)
    hipLaunchKernelGGL(k_occ_orb_der, dim3(blocks((int64_t)p.nbf5 * p.nv, 256)), dim3(256), 0, s, p, n, dn, fi, nd12,
                       dfi, dnd, dnd12);
  return hipGetLastError();
}

hipError_t launch_pnof_coef(const PnofDev& p, const double* n, const double* fi, const double* nd, const double* nd12,
                            double* cj, double* ck, hipStream_t s) {
  hipLaunchKernelGGL(k_occ_coef, dim3(blocks((int64_t)p.nbf5 * p.nbf5, 256)), dim3(256), 0, s, p, n, fi, nd, nd12, cj,
                     ck);
  return hipGetLastError();
}

hipError_t launch_occ_transpose(const double* src, int n, double* dst, hipStream_t s) {
  hipLaunchKernelGGL(k_occ_transpose, dim3(blocks((int64_t)n * n, 256)), dim3(256), 0, s, src, n, dst);
  return hipGetLastError();
}

hipError_t launch_occ_energy(const PnofDev& p, const double* n, const double* fi, const double* nd, const double* nd12,
                             const double* cjm, const double* ckm, const double* JT, const double* KT,
                             const double* hc, double* part, double* E, hipStream_t s) {
  const int nparts = (p.nbf5 + 3) / 4;
  hipLaunchKernelGGL(k_occ_energy_rows, dim3(nparts), dim3(256), 0, s, p, n, fi, nd, nd12, cjm, ckm, JT, KT, hc, part);
  hipLaunchKernelGGL(k_occ_energy_sum, dim3(1), dim3(256), 0, s, part, nparts, E);
  return hipGetLastError();
}

hipError_t launch_occ_grad(const PnofDev& p, const double* n, const double* fi, const double* nd, const double* nd12,
                           const double* dn, const double* dfi, const double* dnd, const double* dnd12,
                           const double* JT, const double* KT, const double* hc, double* grad, hipStream_t s) {
  if (p.nv <= 0) return hipSuccess;
  Der d{n, fi, nd, nd12, dn, dfi, dnd, dnd12};
  hipLaunchKernelGGL(k_occ_grad, dim3(p.nv), dim3(256), 0, s, p, d, JT, KT, hc, grad);
  return hipGetLastError();
}

}  // namespace pnjk
