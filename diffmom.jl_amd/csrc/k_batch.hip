// k_batch.hip — Arnoldi kernels for lockstep batches of right-hand sides.
//
// A batch block is n x R complex128 with column r at offset r * n; a Krylov
// basis is a sequence of such blocks (block i at offset i * n * R).  Every
// kernel takes act[r]: a column with act[r] == 0 (converged, or not part of
// this solve) is skipped.  h arrays are R x kcap complex, column r's
// coefficient i at r * kcap + i.
#include <hip/hip_runtime.h>

#include "dm_common.h"
#include "k_batch.h"

namespace {

constexpr int BT = 256;          // threads per workgroup
constexpr int BROWS = 4;         // rows per thread in the dot kernel
constexpr int BCHUNK = BT * BROWS;
constexpr int BKMAX = 512;       // basis vectors per column

__device__ inline double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// part[(r * kcap + i) * nch + c] = sum over the rows of chunk c of conj(V_i[r]) . W[r]
// (each workgroup keeps its chunk of W[r] in registers and walks the k basis vectors)
__global__ __launch_bounds__(BT) void k_bdots(i64 n, int R, int k, int kcap, const double2* __restrict__ V,
                                              const double2* __restrict__ W, const int* __restrict__ act,
                                              double2* __restrict__ part, int nch) {
  const int c = blockIdx.x, r = blockIdx.y;
  if (!act[r]) return;
  __shared__ double2 red[BT / 64][BKMAX];
  const i64 base = (i64)c * BCHUNK + threadIdx.x;
  const i64 NR = n * R;
  double2 w[BROWS];
#pragma unroll
  for (int e = 0; e < BROWS; ++e) {
    i64 m = base + e * BT;
    w[e] = m < n ? W[(i64)r * n + m] : make_double2(0.0, 0.0);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = 0; i < k; ++i) {
    const double2* v = V + (i64)i * NR + (i64)r * n;
    double ar = 0, ai = 0;
#pragma unroll
    for (int e = 0; e < BROWS; ++e) {
      i64 m = base + e * BT;
      if (m < n) {
        double2 a = v[m];
        ar += a.x * w[e].x + a.y * w[e].y;
        ai += a.x * w[e].y - a.y * w[e].x;
      }
    }
    ar = wsum(ar);
    ai = wsum(ai);
    if (lane == 0) red[wave][i] = make_double2(ar, ai);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < k; i += BT) {
    double sr = 0, si = 0;
#pragma unroll
    for (int q = 0; q < BT / 64; ++q) {
      sr += red[q][i].x;
      si += red[q][i].y;
    }
    part[((i64)r * kcap + i) * nch + c] = make_double2(sr, si);
  }
}

__global__ void k_bdots_reduce(int R, int k, int kcap, int nch, const double2* __restrict__ part,
                               const int* __restrict__ act, double2* __restrict__ h) {
  const int r = blockIdx.x;
  if (!act[r]) return;
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    const double2* p = part + ((i64)r * kcap + i) * nch;
    double sr = 0, si = 0;
    for (int c = 0; c < nch; ++c) {
      sr += p[c].x;
      si += p[c].y;
    }
    h[(i64)r * kcap + i] = make_double2(sr, si);
  }
}

// W[r] -= sum_{i<k} V_i[r] h[r, i]
__global__ __launch_bounds__(BT) void k_bupdate(i64 n, int R, int k, int kcap, const double2* __restrict__ V,
                                                const double2* __restrict__ h, const int* __restrict__ act,
                                                double2* __restrict__ W) {
  const int r = blockIdx.y;
  if (!act[r]) return;
  __shared__ double2 hs[BKMAX];
  for (int i = threadIdx.x; i < k; i += BT) hs[i] = h[(i64)r * kcap + i];
  __syncthreads();
  const i64 m = blockIdx.x * (i64)BT + threadIdx.x;
  if (m >= n) return;
  const i64 NR = n * R, off = (i64)r * n + m;
  double ar = 0, ai = 0;
  for (int i = 0; i < k; ++i) {
    double2 a = V[(i64)i * NR + off], b = hs[i];
    ar += a.x * b.x - a.y * b.y;
    ai += a.x * b.y + a.y * b.x;
  }
  double2 o = W[off];
  W[off] = make_double2(o.x - ar, o.y - ai);
}

// X[r] += sum_{i<cnt[r]} V_i[r] y[r, i]
__global__ __launch_bounds__(BT) void k_bcombine(i64 n, int R, int kcap, const double2* __restrict__ V,
                                                 const double2* __restrict__ y, const int* __restrict__ cnt,
                                                 double2* __restrict__ X) {
  const int r = blockIdx.y;
  const int k = cnt[r];
  if (k <= 0) return;
  __shared__ double2 ys[BKMAX];
  for (int i = threadIdx.x; i < k; i += BT) ys[i] = y[(i64)r * kcap + i];
  __syncthreads();
  const i64 m = blockIdx.x * (i64)BT + threadIdx.x;
  if (m >= n) return;
  const i64 NR = n * R, off = (i64)r * n + m;
  double ar = 0, ai = 0;
  for (int i = 0; i < k; ++i) {
    double2 a = V[(i64)i * NR + off], b = ys[i];
    ar += a.x * b.x - a.y * b.y;
    ai += a.x * b.y + a.y * b.x;
  }
  double2 o = X[off];
  X[off] = make_double2(o.x + ar, o.y + ai);
}

// out[r] = ||W[r]||^2 (one workgroup of 1024 threads per column)
__global__ __launch_bounds__(1024) void k_bnorm2(i64 n, const double2* __restrict__ W, const int* __restrict__ act,
                                                 double* __restrict__ out) {
  const int r = blockIdx.x;
  if (act && !act[r]) return;
  const double2* w = W + (i64)r * n;
  double s = 0;
  for (i64 m = threadIdx.x; m < n; m += 1024) {
    double2 a = w[m];
    s += a.x * a.x + a.y * a.y;
  }
  s = wsum(s);
  __shared__ double red[16];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int q = 0; q < 16; ++q) t += red[q];
    out[r] = t;
  }
}

// dst[r] = src[r] * sc[r] (real per-column factor; sc = 0 clears the column;
// src == dst allowed)
__global__ void k_bscale(i64 n, const double2* src, const double* __restrict__ sc, double2* dst) {
  const int r = blockIdx.y;
  const i64 m = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double f = sc[r];
  const i64 off = (i64)r * n + m;
  double2 v = src[off];
  dst[off] = make_double2(v.x * f, v.y * f);
}

unsigned row_blocks(i64 n) { return (unsigned)((n + BT - 1) / BT); }

}  // namespace

namespace dmb {

int max_basis() { return BKMAX; }

int dot_chunks(i64 n) { return (int)((n + BCHUNK - 1) / BCHUNK); }

void dots(i64 n, int R, int k, int kcap, const double* V, const double* W, const int* act, double* part, double* h,
          hipStream_t s) {
  if (k <= 0) return;
  DM_REQUIRE(k <= BKMAX && k <= kcap, DM_ERR_ARGUMENT, "batched Krylov basis larger than 512 vectors");
  DM_REQUIRE(R >= 1 && R <= 65535, DM_ERR_ARGUMENT, "batch width out of range");
  const int nch = dot_chunks(n);
  k_bdots<<<dim3((unsigned)nch, (unsigned)R), BT, 0, s>>>(n, R, k, kcap, reinterpret_cast<const double2*>(V),
                                                          reinterpret_cast<const double2*>(W), act,
                                                          reinterpret_cast<double2*>(part), nch);
  k_bdots_reduce<<<R, 256, 0, s>>>(R, k, kcap, nch, reinterpret_cast<const double2*>(part), act,
                                   reinterpret_cast<double2*>(h));
  HIP_OK(hipGetLastError());
}

void update(i64 n, int R, int k, int kcap, const double* V, const double* h, const int* act, double* W,
            hipStream_t s) {
  if (k <= 0) return;
  DM_REQUIRE(k <= BKMAX && k <= kcap, DM_ERR_ARGUMENT, "batched Krylov basis larger than 512 vectors");
  k_bupdate<<<dim3(row_blocks(n), (unsigned)R), BT, 0, s>>>(n, R, k, kcap, reinterpret_cast<const double2*>(V),
                                                            reinterpret_cast<const double2*>(h), act,
                                                            reinterpret_cast<double2*>(W));
  HIP_OK(hipGetLastError());
}

void combine(i64 n, int R, int kcap, const double* V, const double* y, const int* cnt, double* X, hipStream_t s) {
  DM_REQUIRE(kcap <= BKMAX, DM_ERR_ARGUMENT, "batched Krylov basis larger than 512 vectors");
  k_bcombine<<<dim3(row_blocks(n), (unsigned)R), BT, 0, s>>>(n, R, kcap, reinterpret_cast<const double2*>(V),
                                                             reinterpret_cast<const double2*>(y), cnt,
                                                             reinterpret_cast<double2*>(X));
  HIP_OK(hipGetLastError());
}

void norm2(i64 n, int R, const double* W, const int* act, double* out, hipStream_t s) {
  k_bnorm2<<<R, 1024, 0, s>>>(n, reinterpret_cast<const double2*>(W), act, out);
  HIP_OK(hipGetLastError());
}

void scale(i64 n, int R, const double* src, const double* sc, double* dst, hipStream_t s) {
  k_bscale<<<dim3(row_blocks(n), (unsigned)R), BT, 0, s>>>(n, reinterpret_cast<const double2*>(src), sc,
                                                           reinterpret_cast<double2*>(dst));
  HIP_OK(hipGetLastError());
}

}  // namespace dmb
