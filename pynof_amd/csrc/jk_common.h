// Shared definitions for the MI355X (gfx950) RI JK build kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

namespace pnjk {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));

// Tile shape of one GEMM launch (FP64 or BF16x3 core): BM = 32*fm rows,
// BN = 32*fn columns per workgroup.
struct TileF64 {
  int fm;
  int fn;
};

// One output tile of a GEMM-shaped problem.
struct Tile {
  int batch, mt, nt, m0, n0;
};

// Records a message as the calling thread's last error; returns `code`.
int set_error(int code, const std::string& msg);

#define PNJK_HIP_CHECK(expr)                                                        \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess)                                                           \
      return ::pnjk::set_error(2, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Pointers that a kernel reads out of a problem table in device memory (the
// grouped launches) carry no address space, so the compiler would emit flat_*
// accesses; flat loads also count against lgkmcnt, which makes every LDS wait a
// wait for the HBM loads as well. gmem() restores the global address space.
#define PNJK_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const PNJK_GLOBAL T* gmem(const T* p) {
  return (const PNJK_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ PNJK_GLOBAL T* gmem(T* p) {
  return (PNJK_GLOBAL T*)p;
}

// Workgroup -> XCD placement on MI355X: consecutive workgroup ids are dealt
// round-robin over the 8 XCDs. This maps a workgroup id to a logical tile id so
// that each XCD receives one contiguous run of logical tiles, keeping tiles that
// share an operand in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  constexpr int kXcd = 8;
  const int per = nblocks / kXcd, rem = nblocks % kXcd;
  const int x = bid % kXcd, i = bid / kXcd;
  return x * per + (x < rem ? x : rem) + i;
}

// bf16 bit pattern of f, rounded to nearest-even (inputs are finite).
__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// x ~ hi + lo with hi = bf16(x), lo = bf16(x - hi).
__device__ __forceinline__ void split2(double x, uint16_t& hi, uint16_t& lo) {
  hi = bf16_rne((float)x);
  lo = bf16_rne((float)(x - (double)bf16_to_f(hi)));
}

// Cyclic-window schedule of the quarter transform Y_p[q][P] = sum_nu C[nu][q] X[p][nu][P].
// Y_p[q] = Y_q[p] (b is symmetric in its two orbital indices), so each unordered pair
// {p, q} of the n = nbf5 orbitals is needed once. Batch p computes the window rows
// r < qt_window(n), q = (p + r) mod n: every batch has the same n / 2 + 1 rows (the lower
// triangle would give batch p its p + 1 rows q <= p). Distances (q - p) mod n below the
// window are owned by batch p and the others by batch q; for even n the antipodal pair
// (r = n / 2) is reached from both ends and only the batch p < n / 2 keeps it.
__host__ __device__ inline int qt_window(int n) { return n / 2 + 1; }
__host__ __device__ inline int qt_row(int p, int r, int n) {
  const int q = p + r;
  return q >= n ? q - n : q;
}
__host__ __device__ inline bool qt_owns(int p, int r, int n) { return !((n & 1) == 0 && 2 * r == n && 2 * p >= n); }
// Columns of the FP64 quarter transform's window copy of C (jk_f64.hip): batch offsets below
// n plus whole tiles of at most 160 rows over the window.
__host__ __device__ inline int qt_window_cols(int n) { return (n + qt_window(n) + 160 + 31) / 32 * 32; }

// Lower-triangle tile pair number rem -> (mt, nt), nt <= mt, rem = mt*(mt+1)/2 + nt.
__host__ __device__ inline void tri_pair(int rem, int& mt, int& nt) {
  int m = (int)((sqrt(8.0 * rem + 1.0) - 1.0) * 0.5);
  while ((m + 1) * (m + 2) / 2 <= rem) ++m;
  while (m * (m + 1) / 2 > rem) --m;
  mt = m;
  nt = rem - m * (m + 1) / 2;
}

// Tile number rem of a symmetric product's lower triangle of tiles whose first npair
// numbers are diagonal pair tiles: rem < npair is pair s = rem, covering tiles (2s, 2s),
// (2s + 1, 2s) and (2s + 1, 2s + 1) (returned as mt = 2s, nt = -1); the other numbers are
// the tiles (mt, nt), nt <= mt, that no pair covers, in row-major order (rows 2s and
// 2s + 1 keep their 2s tiles left of the pair, rows from 2 npair on are whole). With
// npair = 0 this is tri_pair. A triangle of T tile rows has T (T + 1) / 2 - 2 npair numbers.
__host__ __device__ inline void syrk_tile(int rem, int npair, int& mt, int& nt) {
  if (rem < npair) {
    mt = 2 * rem;
    nt = -1;
    return;
  }
  rem -= npair;
  const int paired = 2 * npair * (npair - 1);  // tiles of rows 0 .. 2 npair - 1
  if (rem < paired) {
    int s = (int)((1.0 + sqrt(1.0 + 2.0 * rem)) * 0.5);
    while (2 * s * (s - 1) > rem) --s;
    while (2 * (s + 1) * s <= rem) ++s;
    const int r = rem - 2 * s * (s - 1);
    mt = r < 2 * s ? 2 * s : 2 * s + 1;
    nt = r < 2 * s ? r : r - 2 * s;
  } else {
    tri_pair(rem - paired + npair * (2 * npair + 1), mt, nt);
  }
}

// Row index of the packed lower triangle: t = mu*(mu+1)/2 + nu for mu >= nu.
__device__ __forceinline__ int64_t tri_index(int a, int b) {
  const int64_t hi = a > b ? a : b, lo = a > b ? b : a;
  return hi * (hi + 1) / 2 + lo;
}

// Inverse of tri_index for t < 2^40: returns mu (row), nu = t - mu*(mu+1)/2.
__device__ __forceinline__ int tri_row(int64_t t) {
  int mu = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((int64_t)(mu + 1) * (mu + 2) / 2 <= t) ++mu;
  while ((int64_t)mu * (mu + 1) / 2 > t) --mu;
  return mu;
}

}  // namespace pnjk
