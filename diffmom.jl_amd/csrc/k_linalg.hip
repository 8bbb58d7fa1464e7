// k_linalg.hip — dense EFIE products, Krylov building blocks, preconditioner
// and direct-solve kernels (all in the internal spatial basis order).
//
// Z is symmetric (Z = Z^T, src/assembly/EFIE.jl:737-742), so y = Z x is formed
// as column dot products y_n = sum_m Z[m, n] x_m: every wave streams whole
// columns of the column-major matrix with 16-byte loads, which is the HBM-bound
// access pattern the single-right-hand-side product needs.
#include <hip/hip_runtime.h>

#include "dm_common.h"
#include "k_linalg.h"

#include <algorithm>

namespace {

typedef double nd2 __attribute__((ext_vector_type(2)));
typedef float nf4 __attribute__((ext_vector_type(4)));
// Streaming loads: the matrix is read once per product, so bypass cache retention.
__device__ inline double2 ld_stream(const double2* p) {
  nd2 v = __builtin_nontemporal_load(reinterpret_cast<const nd2*>(p));
  return make_double2(v.x, v.y);
}
__device__ inline float4 ld_stream(const float4* p) {
  nf4 v = __builtin_nontemporal_load(reinterpret_cast<const nf4*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// The same sum through DPP moves (no LDS round trip per step): pairs, quads,
// half rows, rows, then row broadcasts 15 and 31; lane 63 holds the total,
// which readlane hands to every lane.  Each step moves the two 32-bit halves
// of the double; lanes outside a step's row mask add zero.
template <int CTRL, int ROWS>
__device__ inline double dpp_add_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffLL), CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xF, false);
  return v + __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline double wave_sum_dpp(double x) {
  x = dpp_add_f64<0xB1, 0xF>(x);   // quad_perm 1,0,3,2
  x = dpp_add_f64<0x4E, 0xF>(x);   // quad_perm 2,3,0,1
  x = dpp_add_f64<0x141, 0xF>(x);  // row_half_mirror
  x = dpp_add_f64<0x140, 0xF>(x);  // row_mirror
  x = dpp_add_f64<0x142, 0xA>(x);  // row_bcast15
  x = dpp_add_f64<0x143, 0xC>(x);  // row_bcast31
  const long long b = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffLL), 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (long long)lo);
}

// y = Z x, complex128 Z.  Each wave owns NC consecutive columns; lane l reads
// rows l, l+64, ...; x is shared by the NC columns from registers.
template <int NC>
__global__ __launch_bounds__(256) void k_gemv_z64(i64 n, i64 ld, const double2* __restrict__ z,
                                                  const double2* __restrict__ x, double2* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const i64 c0 = ((i64)blockIdx.x * 4 + (threadIdx.x >> 6)) * NC;
  if (c0 >= n) return;
  double ar[NC], ai[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) ar[c] = ai[c] = 0.0;
  const double2* zc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) zc[c] = z + (c0 + c < n ? c0 + c : c0) * ld;
  i64 m = lane;
  for (; m + 64 < n; m += 128) {
    double2 x0 = x[m], x1 = x[m + 64];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double2 a0 = ld_stream(zc[c] + m);
      double2 a1 = ld_stream(zc[c] + m + 64);
      ar[c] += a0.x * x0.x - a0.y * x0.y + a1.x * x1.x - a1.y * x1.y;
      ai[c] += a0.x * x0.y + a0.y * x0.x + a1.x * x1.y + a1.y * x1.x;
    }
  }
  for (; m < n; m += 64) {
    double2 x0 = x[m];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double2 a0 = zc[c][m];
      ar[c] += a0.x * x0.x - a0.y * x0.y;
      ai[c] += a0.x * x0.y + a0.y * x0.x;
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double sr = wave_sum(ar[c]), si = wave_sum(ai[c]);
    if (lane == 0 && c0 + c < n) y[c0 + c] = make_double2(sr, si);
  }
}

// y = Z x with complex64 storage and FP64 arithmetic (the product is HBM
// bound, so FP64 accumulation costs nothing).  Lane reads two consecutive
// entries (16 bytes) per column.
template <int NC>
__global__ __launch_bounds__(256) void k_gemv_z32(i64 n, i64 ld, const float4* __restrict__ z4,
                                                  const double2* __restrict__ x, double2* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const i64 c0 = ((i64)blockIdx.x * 4 + (threadIdx.x >> 6)) * NC;
  if (c0 >= n) return;
  double ar[NC], ai[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) ar[c] = ai[c] = 0.0;
  const i64 ld2 = ld >> 1;  // ld is even
  const i64 n2 = n >> 1;
  for (i64 m2 = lane; m2 < n2; m2 += 64) {
    double2 x0 = x[2 * m2], x1 = x[2 * m2 + 1];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      i64 col = c0 + c < n ? c0 + c : c0;
      float4 a = ld_stream(z4 + col * ld2 + m2);
      ar[c] += (double)a.x * x0.x - (double)a.y * x0.y + (double)a.z * x1.x - (double)a.w * x1.y;
      ai[c] += (double)a.x * x0.y + (double)a.y * x0.x + (double)a.z * x1.y + (double)a.w * x1.x;
    }
  }
  if ((n & 1) && lane == 0) {
    i64 m = n - 1;
    const float2* z2 = reinterpret_cast<const float2*>(z4);
    double2 x0 = x[m];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      i64 col = c0 + c < n ? c0 + c : c0;
      float2 a = z2[col * ld + m];
      ar[c] += (double)a.x * x0.x - (double)a.y * x0.y;
      ai[c] += (double)a.x * x0.y + (double)a.y * x0.x;
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    double sr = wave_sum(ar[c]), si = wave_sum(ai[c]);
    if (lane == 0 && c0 + c < n) y[c0 + c] = make_double2(sr, si);
  }
}

__global__ void k_permute(i64 n, const int* __restrict__ idx, const double2* __restrict__ src,
                          double2* __restrict__ dst, int gather) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (gather) dst[i] = src[idx[i]];
  else dst[idx[i]] = src[i];
}

// ---------------------------------------------------------------- Krylov ops
// Row-chunked dot products for the Arnoldi step: workgroup (c, g) forms the
// partial sums of vectors g*kDotVG .. over 1024 rows (w read once per group,
// every V entry once); k_update adds them over the chunks.
constexpr int kDotVG = 4;
constexpr int kDotRows = 4;  // rows per thread

// jd (captured steps): k = *jd + 1 from device memory, the grid sized for
// the largest k of the graph (workgroups past k exit).
__global__ __launch_bounds__(256) void k_dots_part(i64 n, i64 ldv, int k, const int* __restrict__ jd,
                                                   const double2* __restrict__ V, const double2* __restrict__ w,
                                                   double2* __restrict__ part, int nch) {
  __shared__ double2 red[4][kDotVG];
  if (jd) k = *jd + 1;
  const int c = blockIdx.x, i0 = blockIdx.y * kDotVG;
  if (i0 >= k) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const i64 m0 = (i64)c * 256 * kDotRows + threadIdx.x;
  double2 wm[kDotRows];
#pragma unroll
  for (int u = 0; u < kDotRows; ++u) {
    const i64 m = m0 + 256 * u;
    wm[u] = m < n ? w[m] : make_double2(0.0, 0.0);
  }
  const int cnt = min(kDotVG, k - i0);
  // all kDotVG x kDotRows loads are issued before any sum: vectors past cnt
  // re-read the last one (clamped, unconditional) and are not stored
  double2 a[kDotVG][kDotRows];
#pragma unroll
  for (int q = 0; q < kDotVG; ++q) {
    const double2* v = V + (i64)(i0 + min(q, cnt - 1)) * ldv;
#pragma unroll
    for (int u = 0; u < kDotRows; ++u) {
      const i64 m = m0 + 256 * u;
      a[q][u] = v[min(m, n - 1)];
    }
  }
#pragma unroll
  for (int q = 0; q < kDotVG; ++q) {
    double ar = 0, ai = 0;
#pragma unroll
    for (int u = 0; u < kDotRows; ++u) {
      // rows past n carry w = 0
      ar += a[q][u].x * wm[u].x + a[q][u].y * wm[u].y;
      ai += a[q][u].x * wm[u].y - a[q][u].y * wm[u].x;
    }
    ar = wave_sum_dpp(ar);
    ai = wave_sum_dpp(ai);
    if (lane == 0 && q < cnt) red[wave][q] = make_double2(ar, ai);
  }
  __syncthreads();
  if (threadIdx.x < cnt) {
    const int q = threadIdx.x;
    part[(i64)c * k + i0 + q] = make_double2(red[0][q].x + red[1][q].x + red[2][q].x + red[3][q].x,
                                             red[0][q].y + red[1][q].y + red[2][q].y + red[3][q].y);
  }
}

// w[m] -= sum_{i<k} V[m, i] h[i], where h[i] = sum_c part[c * k + i] are
// the chunk partials of k_dots_part: every workgroup adds them itself in the
// same order (for k <= 256 four waves over interleaved chunks, then their
// four sums; else thread i over the chunks), so all workgroups hold the same
// h, and workgroup 0 stores h (at h + hsel k).  A workgroup takes
// kUpdRows rows; wave q sums the vectors i = q mod 4 (a quarter of the
// dependent loads per wave), and wave 0 adds the four sums in order.  nrm,
// when given, receives each workgroup's sum of |w[m]|^2 over its rows after
// the update.  jd (captured steps): k = *jd + 1 from device memory.
constexpr int kUpdRows = 64;
__global__ __launch_bounds__(256) void k_update(i64 n, i64 ldv, int k, const int* __restrict__ jd, int hsel,
                                                const double2* __restrict__ V, const double2* __restrict__ part,
                                                int nch, double2* __restrict__ h, double2* __restrict__ w,
                                                double* __restrict__ nrm) {
  __shared__ double2 hs[1024];
  __shared__ double2 acc[3][kUpdRows];
  __shared__ double2 ps[4][256];
  if (jd) k = *jd + 1;
  h += (i64)hsel * k;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (k <= 256) {
    // wave q adds the chunks c = q mod 4 of every vector (a quarter of the
    // dependent loads per thread), then the four sums in order
    for (int i = lane; i < k; i += 64) {
      double sr = 0, si = 0;
#pragma unroll 4
      for (int c = wave; c < nch; c += 4) {
        const double2 v = part[(i64)c * k + i];
        sr += v.x;
        si += v.y;
      }
      ps[wave][i] = make_double2(sr, si);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < k; i += 256) {
      const double sr = ((ps[0][i].x + ps[1][i].x) + ps[2][i].x) + ps[3][i].x;
      const double si = ((ps[0][i].y + ps[1][i].y) + ps[2][i].y) + ps[3][i].y;
      hs[i] = make_double2(sr, si);
      if (blockIdx.x == 0) h[i] = make_double2(sr, si);
    }
  } else {
    for (int i = threadIdx.x; i < k; i += 256) {
      double sr = 0, si = 0;
#pragma unroll 8
      for (int c = 0; c < nch; ++c) {
        const double2 v = part[(i64)c * k + i];
        sr += v.x;
        si += v.y;
      }
      hs[i] = make_double2(sr, si);
      if (blockIdx.x == 0) h[i] = make_double2(sr, si);
    }
  }
  __syncthreads();
  const i64 m = blockIdx.x * (i64)kUpdRows + lane;
  double ar = 0, ai = 0;
  if (m < n) {
    for (int i = wave; i < k; i += 4) {
      const double2 a = V[(i64)i * ldv + m], b = hs[i];
      ar += a.x * b.x - a.y * b.y;
      ai += a.x * b.y + a.y * b.x;
    }
  }
  if (wave > 0) acc[wave - 1][lane] = make_double2(ar, ai);
  __syncthreads();
  if (wave != 0) return;
  double s2 = 0;
  if (m < n) {
    ar += acc[0][lane].x;
    ai += acc[0][lane].y;
    ar += acc[1][lane].x;
    ai += acc[1][lane].y;
    ar += acc[2][lane].x;
    ai += acc[2][lane].y;
    const double2 o = w[m];
    const double2 u = make_double2(o.x - ar, o.y - ai);
    w[m] = u;
    s2 = u.x * u.x + u.y * u.y;
  }
  if (nrm) {
    s2 = wave_sum(s2);
    if (lane == 0) nrm[blockIdx.x] = s2;
  }
}

// dst column *jd of base (leading dimension ldv) = src
__global__ void k_copy_col(i64 n, const double2* __restrict__ src, double2* __restrict__ base, i64 ldv,
                           const int* __restrict__ jd) {
  const i64 m = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (m < n) base[(i64)(*jd) * ldv + m] = src[m];
}

__global__ void k_step_inc(int* jd) {
  if (threadIdx.x == 0) *jd += 1;
}

// out[0] = ||w||^2 (single workgroup of 1024 threads)
__global__ __launch_bounds__(1024) void k_norm2(i64 n, const double2* __restrict__ w, double* out) {
  double s = 0;
  for (i64 m = threadIdx.x; m < n; m += 1024) {
    double2 a = w[m];
    s += a.x * a.x + a.y * a.y;
  }
  s = wave_sum(s);
  __shared__ double part[16];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < 16; ++i) t += part[i];
    out[0] = t;
  }
}

// y = a*x + b*y (complex a, b).  With b == 0 the old y is not read: it may be
// uninitialised memory (a fresh Krylov column), whose NaN bits 0 * y would keep.
__global__ void k_axpby(i64 n, double2 a, const double2* x, double2 b, double2* y) {
  i64 m = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (m >= n) return;
  double2 xv = x[m];
  double2 r = make_double2(a.x * xv.x - a.y * xv.y, a.x * xv.y + a.y * xv.x);
  if (b.x != 0.0 || b.y != 0.0) {
    double2 yv = y[m];
    r.x += b.x * yv.x - b.y * yv.y;
    r.y += b.x * yv.y + b.y * yv.x;
  }
  y[m] = r;
}

// x += V y (k columns, y on device)
__global__ __launch_bounds__(256) void k_combine(i64 n, i64 ldv, int k, const double2* __restrict__ V,
                                                 const double2* __restrict__ yv, double2* __restrict__ x) {
  __shared__ double2 hs[1024];
  for (int i = threadIdx.x; i < k; i += 256) hs[i] = yv[i];
  __syncthreads();
  i64 m = blockIdx.x * (i64)256 + threadIdx.x;
  if (m >= n) return;
  double ar = 0, ai = 0;
  for (int i = 0; i < k; ++i) {
    double2 a = V[(i64)i * ldv + m], b = hs[i];
    ar += a.x * b.x - a.y * b.y;
    ai += a.x * b.y + a.y * b.x;
  }
  double2 o = x[m];
  x[m] = make_double2(o.x + ar, o.y + ai);
}

__global__ void k_cast_matrix(i64 n, i64 ld, const double2* __restrict__ src, float2* __restrict__ dst) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= ld * n) return;
  double2 v = src[idx];
  dst[idx] = make_float2((float)v.x, (float)v.y);
}

// ------------------------------------------------------- block preconditioner
// Extract diagonal blocks of Z (internal order): blocks[b] = Z[b*bs .. , b*bs ..],
// identity-padded past n.
__global__ void k_extract_blocks(i64 n, ZRef z, int bs, double2* __restrict__ blocks) {
  int b = blockIdx.y;
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= (i64)bs * bs) return;
  int r = idx % bs, c = idx / bs;
  i64 gr = (i64)b * bs + r, gc = (i64)b * bs + c;
  blocks[(i64)b * bs * bs + idx] = (gr < n && gc < n) ? zat(z, gr, gc) : make_double2(r == c ? 1.0 : 0.0, 0.0);
}

// In-place Gauss-Jordan inverse with partial pivoting of each bs x bs block
// (column-major), one workgroup per block.  The singular-pivot flag is set in
// status[b] (the host turns it into an error).
__global__ __launch_bounds__(256) void k_block_inverse(int bs, double2* __restrict__ blocks, int* status) {
  double2* A = blocks + (i64)blockIdx.x * bs * bs;
  extern __shared__ double2 sh[];
  double2* rowk = sh;         // bs
  double2* colk = sh + bs;    // bs
  __shared__ int piv;
  __shared__ double pmax;
  // perm tracks row swaps to un-permute the columns at the end
  int* perm = reinterpret_cast<int*>(sh + 2 * bs);
  for (int i = threadIdx.x; i < bs; i += blockDim.x) perm[i] = i;
  __syncthreads();
  for (int k = 0; k < bs; ++k) {
    if (threadIdx.x == 0) {
      int p = k;
      double m = -1;
      for (int i = k; i < bs; ++i) {
        double2 v = A[i + (i64)k * bs];
        double a = v.x * v.x + v.y * v.y;
        if (a > m) { m = a; p = i; }
      }
      piv = p;
      pmax = m;
    }
    __syncthreads();
    if (pmax == 0.0) {
      if (threadIdx.x == 0) status[blockIdx.x] = 1;
      return;
    }
    int p = piv;
    if (p != k) {
      for (int c = threadIdx.x; c < bs; c += blockDim.x) {
        double2 t = A[k + (i64)c * bs];
        A[k + (i64)c * bs] = A[p + (i64)c * bs];
        A[p + (i64)c * bs] = t;
      }
      if (threadIdx.x == 0) { int t = perm[k]; perm[k] = perm[p]; perm[p] = t; }
    }
    __syncthreads();
    double2 d = A[k + (i64)k * bs];
    double dn = d.x * d.x + d.y * d.y;
    double2 inv = make_double2(d.x / dn, -d.y / dn);
    // row k scaled by 1/d (with the pivot column replaced by 1/d)
    for (int c = threadIdx.x; c < bs; c += blockDim.x) {
      double2 v = c == k ? make_double2(1.0, 0.0) : A[k + (i64)c * bs];
      rowk[c] = make_double2(v.x * inv.x - v.y * inv.y, v.x * inv.y + v.y * inv.x);
    }
    for (int r = threadIdx.x; r < bs; r += blockDim.x) colk[r] = r == k ? make_double2(0.0, 0.0) : A[r + (i64)k * bs];
    __syncthreads();
    for (i64 idx = threadIdx.x; idx < (i64)bs * bs; idx += blockDim.x) {
      int r = idx % bs, c = idx / bs;
      if (r == k) {
        A[idx] = rowk[c];
      } else {
        double2 base = c == k ? make_double2(0.0, 0.0) : A[idx];
        double2 f = colk[r], g = rowk[c];
        A[idx] = make_double2(base.x - (f.x * g.x - f.y * g.y), base.y - (f.x * g.y + f.y * g.x));
      }
    }
    __syncthreads();
  }
  // undo the row permutation of the input: inverse columns are permuted
  // (A P)^-1 handling: Gauss-Jordan on P A gives (P A)^-1 = A^-1 P^T, so
  // A^-1 = (P A)^-1 P: column j of A^-1 is column perm^-1(j) of the result.
  for (int r = 0; r < bs; ++r) {
    for (int c = threadIdx.x; c < bs; c += blockDim.x) rowk[perm[c]] = A[r + (i64)c * bs];
    __syncthreads();
    for (int c = threadIdx.x; c < bs; c += blockDim.x) A[r + (i64)c * bs] = rowk[c];
    __syncthreads();
  }
}

// y_b = inv_b x_b for every block (block-diagonal product).
__global__ __launch_bounds__(256) void k_block_apply(i64 n, int bs, const double2* __restrict__ inv,
                                                     const double2* __restrict__ x, double2* __restrict__ y) {
  int b = blockIdx.y;
  int r = blockIdx.x * 256 + threadIdx.x;
  extern __shared__ double2 xs[];
  for (int c = threadIdx.x; c < bs; c += 256) {
    i64 g = (i64)b * bs + c;
    xs[c] = g < n ? x[g] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  if (r >= bs) return;
  i64 gr = (i64)b * bs + r;
  if (gr >= n) return;
  const double2* A = inv + (i64)b * bs * bs;
  double ar = 0, ai = 0;
  for (int c = 0; c < bs; ++c) {
    double2 a = A[r + (i64)c * bs], v = xs[c];
    ar += a.x * v.x - a.y * v.y;
    ai += a.x * v.y + a.y * v.x;
  }
  y[gr] = make_double2(ar, ai);
}

__global__ void k_diag_inv(i64 n, ZRef z, double2* __restrict__ d) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 v = zat(z, i, i);
  double m = v.x * v.x + v.y * v.y;
  d[i] = make_double2(v.x / m, -v.y / m);
}

__global__ void k_diag_apply(i64 n, const double2* __restrict__ d, const double2* __restrict__ x, double2* __restrict__ y) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 a = d[i], v = x[i];
  y[i] = make_double2(a.x * v.x - a.y * v.y, a.x * v.y + a.y * v.x);
}

// ------------------------------------------------------------ dense LU
// Blocked right-looking LU with partial pivoting (the dense direct method,
// solve_forward :direct), panels of kLuW columns: one workgroup factors the
// panel (pivot search, row swap inside the panel, scaling and the panel's own
// rank-1 updates, column by column, then the inverse of its unit lower
// triangle L11), k_lu_laswp applies the panel's swaps to the other columns,
// U12 = L11^-1 A12 is a small complex128 GEMM with that inverse, and the trailing
// matrix takes A22 -= L21 U12 as one complex128 GEMM on the FP64 matrix cores
// (gemm_c128, subtract mode).  Pivots: the largest |a|^2 in the column, the
// smallest row index among equal ones.
constexpr int kLuW = 32;

__global__ __launch_bounds__(1024) void k_lu_panel(i64 n, i64 ld, int k0, int w, double2* __restrict__ A,
                                                   int* __restrict__ ipiv, int* __restrict__ status) {
  __shared__ double wbest[16];
  __shared__ int widx[16];
  __shared__ double2 prow[kLuW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (int j = 0; j < w; ++j) {
    const i64 c = (i64)k0 + j;
    double m = -1;
    int p = (int)c;
    for (i64 i = c + tid; i < n; i += 1024) {
      const double2 v = A[i + c * ld];
      const double a = v.x * v.x + v.y * v.y;
      if (a > m) {
        m = a;
        p = (int)i;
      }
    }
    // largest |a|^2, smallest row among equal ones: within the wave, then
    // over the 16 waves (every thread reduces the same 16 values)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double om = __shfl_xor(m, o, 64);
      const int op = __shfl_xor(p, o, 64);
      if (om > m || (om == m && op < p)) {
        m = om;
        p = op;
      }
    }
    if (lane == 0) {
      wbest[wave] = m;
      widx[wave] = p;
    }
    __syncthreads();
    double bm = wbest[0];
    int bp = widx[0];
    for (int q = 1; q < 16; ++q) {
      if (wbest[q] > bm || (wbest[q] == bm && widx[q] < bp)) {
        bm = wbest[q];
        bp = widx[q];
      }
    }
    const bool sing = !(bm > 0.0);
    const int pv = sing ? (int)c : bp;
    if (tid == 0) {
      ipiv[c] = pv;
      if (sing && status[0] == 0) status[0] = (int)c + 1;
    }
    // swap rows c and pv inside the panel; the pivot row to LDS
    if (tid < w) {
      const i64 col = (i64)k0 + tid;
      const double2 rc = A[c + col * ld];
      double2 pr = rc;
      if (pv != c) {
        pr = A[pv + col * ld];
        A[c + col * ld] = pr;
        A[pv + col * ld] = rc;
      }
      prow[tid] = pr;
    }
    __syncthreads();
    if (!sing) {
      const double2 d = prow[j];
      const double dn = d.x * d.x + d.y * d.y;
      const double2 inv = make_double2(d.x / dn, -d.y / dn);
      for (i64 i = c + 1 + tid; i < n; i += 1024) {
        const double2 ej = A[i + c * ld];
        const double2 l = make_double2(ej.x * inv.x - ej.y * inv.y, ej.x * inv.y + ej.y * inv.x);
        A[i + c * ld] = l;
        double2* row = A + i + (i64)k0 * ld;
        // the row's entries past column j, eight loads issued together per round
#pragma unroll 1
        for (int q0 = 0; q0 < kLuW; q0 += 8) {
          if (q0 + 7 <= j || q0 >= w) continue;
          double2 e[8];
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (q0 + u > j && q0 + u < w) e[u] = row[(i64)(q0 + u) * ld];
#pragma unroll
          for (int u = 0; u < 8; ++u)
            if (q0 + u > j && q0 + u < w) {
              const double2 pu = prow[q0 + u];
              row[(i64)(q0 + u) * ld] =
                  make_double2(e[u].x - (l.x * pu.x - l.y * pu.y), e[u].y - (l.x * pu.y + l.y * pu.x));
            }
        }
      }
    }
    __syncthreads();
  }

}

// inverse of the panel's unit lower triangle L11 (w x w, column-major with
// leading dimension kLuW): thread t forward-substitutes e_t
__global__ __launch_bounds__(64) void k_lu_linv(i64 ld, int k0, int w, const double2* __restrict__ A,
                                                double2* __restrict__ linv) {
  const int tid = threadIdx.x;
  if (tid >= w) return;
  double2 x[kLuW];
#pragma unroll
  for (int i = 0; i < kLuW; ++i) x[i] = make_double2(i == tid ? 1.0 : 0.0, 0.0);
#pragma unroll
  for (int i = 1; i < kLuW; ++i) {
    if (i < w) {
      double2 a = x[i];
#pragma unroll
      for (int q = 0; q < i; ++q) {
        const double2 l = A[((i64)k0 + i) + ((i64)k0 + q) * ld];
        a.x -= l.x * x[q].x - l.y * x[q].y;
        a.y -= l.x * x[q].y + l.y * x[q].x;
      }
      x[i] = a;
    }
  }
#pragma unroll
  for (int i = 0; i < kLuW; ++i)
    if (i < w) linv[i + (i64)tid * kLuW] = x[i];
}

// the panel's row swaps (in order) on every column outside the panel
__global__ __launch_bounds__(256) void k_lu_laswp(i64 n, i64 ld, int k0, int w, double2* __restrict__ A,
                                                  const int* __restrict__ ipiv) {
  const i64 col = blockIdx.x * (i64)256 + threadIdx.x;
  if (col >= n || (col >= k0 && col < (i64)k0 + w)) return;
  for (int j = 0; j < w; ++j) {
    const i64 r = (i64)k0 + j;
    const i64 p = ipiv[r];
    if (p != r) {
      const double2 t = A[r + col * ld];
      A[r + col * ld] = A[p + col * ld];
      A[p + col * ld] = t;
    }
  }
}

// Solve with the LU factors for one right-hand side per workgroup (in place):
// the row swaps in order (one thread), then L y = P b and U x = y in blocks
// of 32 rows: one wave solves the block's triangle with its rows in
// registers, then every thread updates its rows below (above) the block with
// the 32 solved values.  b sits in LDS (dynamic) when n <= kLuSolveLds.
constexpr i64 kLuSolveLds = 8192;

__device__ inline double2 cmul2(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }

template <bool LDSB>
__global__ __launch_bounds__(1024) void k_lu_solve(i64 n, i64 ld, const double2* __restrict__ A,
                                                   const int* __restrict__ ipiv, double2* __restrict__ B, i64 ldb) {
  extern __shared__ double2 sb[];
  __shared__ double2 xs[32];
  double2* bg = B + (i64)blockIdx.x * ldb;
  double2* b = LDSB ? sb : bg;
  const int tid = threadIdx.x, lane = tid & 63;
  if (LDSB) {
    for (i64 i = tid; i < n; i += 1024) sb[i] = bg[i];
    __syncthreads();
  }
  if (tid == 0)
    for (i64 k = 0; k < n; ++k) {
      const int p = ipiv[k];
      if (p != k) {
        const double2 t = b[k];
        b[k] = b[p];
        b[p] = t;
      }
    }
  __syncthreads();
  // L y = P b (unit lower triangle)
  for (i64 k0 = 0; k0 < n; k0 += 32) {
    const int w = (int)min((i64)32, n - k0);
    if (tid < 64) {
      double2 v = lane < w ? b[k0 + lane] : make_double2(0.0, 0.0);
      for (int j = 0; j < w; ++j) {
        const double2 xj = make_double2(__shfl(v.x, j, 64), __shfl(v.y, j, 64));
        if (lane > j && lane < w) {
          const double2 t = cmul2(A[(k0 + lane) + (k0 + j) * ld], xj);
          v.x -= t.x;
          v.y -= t.y;
        }
      }
      if (lane < w) {
        b[k0 + lane] = v;
        xs[lane] = v;
      }
    }
    __syncthreads();
    for (i64 i = k0 + w + tid; i < n; i += 1024) {
      double2 v = b[i];
      for (int j = 0; j < w; ++j) {
        const double2 t = cmul2(A[i + (k0 + j) * ld], xs[j]);
        v.x -= t.x;
        v.y -= t.y;
      }
      b[i] = v;
    }
    __syncthreads();
  }
  // U x = y
  for (i64 k1 = n; k1 > 0; k1 -= 32) {
    const i64 k0 = k1 > 32 ? k1 - 32 : 0;
    const int w = (int)(k1 - k0);
    if (tid < 64) {
      double2 v = lane < w ? b[k0 + lane] : make_double2(0.0, 0.0);
      for (int j = w - 1; j >= 0; --j) {
        if (lane == j) {
          const double2 d = A[(k0 + j) + (k0 + j) * ld];
          const double dn = d.x * d.x + d.y * d.y;
          v = make_double2((v.x * d.x + v.y * d.y) / dn, (v.y * d.x - v.x * d.y) / dn);
        }
        const double2 xj = make_double2(__shfl(v.x, j, 64), __shfl(v.y, j, 64));
        if (lane < j) {
          const double2 t = cmul2(A[(k0 + lane) + (k0 + j) * ld], xj);
          v.x -= t.x;
          v.y -= t.y;
        }
      }
      if (lane < w) {
        b[k0 + lane] = v;
        xs[lane] = v;
      }
    }
    __syncthreads();
    for (i64 i = tid; i < k0; i += 1024) {
      double2 v = b[i];
      for (int j = 0; j < w; ++j) {
        const double2 t = cmul2(A[i + (k0 + j) * ld], xs[j]);
        v.x -= t.x;
        v.y -= t.y;
      }
      b[i] = v;
    }
    __syncthreads();
  }
  if (LDSB)
    for (i64 i = tid; i < n; i += 1024) bg[i] = sb[i];
}

// ------------------------------------------------ GMRES Givens on the device
// Column j of the Hessenberg matrix from the two CGS2 passes and the norm of
// the orthogonalised vector (hbuf = [h1 (k complex), h2 (k complex)], k =
// j + 1; ||w||^2 = the sum of the nb per-workgroup partials nrm of the last
// update, added by one wave in a fixed order), the previous rotations
// applied, a new rotation chosen so that
// H[j+1][j] = 0, and g (the rotated right-hand side beta e_1) updated.
// st[2j], st[2j+1] = |g[j+1]| (the residual estimate) and the subdiagonal
// (0 = breakdown); sc[0] = 1 / subdiagonal (0 at breakdown).  The launch
// covers the n rows: every workgroup forms the same norm and writes its rows
// of v_{j+1} = sc w; workgroup 0 also does the rotations.
// jd (captured steps): j = *jd from device memory, v_{j+1} written to column
// j + 1 of Vb (leading dimension ldv) and to vcur (the next step's input).
__global__ __launch_bounds__(256) void k_givens(int j, const int* __restrict__ jd, int ldh,
                                                const double* __restrict__ hbuf, const double* __restrict__ nrm,
                                                int nb, i64 n, const double2* __restrict__ w,
                                                double2* __restrict__ vnext, double2* __restrict__ Vb, i64 ldv,
                                                double2* __restrict__ vcur, double2* __restrict__ H,
                                                double* __restrict__ cs, double2* __restrict__ sn,
                                                double2* __restrict__ g, double* __restrict__ st,
                                                double* __restrict__ sc) {
  if (jd) {
    j = *jd;
    vnext = Vb + (i64)(j + 1) * ldv;
  }
  const int k = j + 1;
  double2* Hj = H + (i64)j * ldh;
  // the column and the previous rotations are staged in LDS by all threads;
  // the rotation chain then runs in one thread without global-memory latency
  extern __shared__ double2 sh[];
  double2* col = sh;          // k + 1
  double2* rs = sh + k + 1;   // j rotations' s
  double* rc = reinterpret_cast<double*>(rs + j);  // j rotations' c
  if (blockIdx.x == 0) {
    for (int i = threadIdx.x; i < k; i += blockDim.x)
      col[i] = make_double2(hbuf[2 * i] + hbuf[2 * k + 2 * i], hbuf[2 * i + 1] + hbuf[2 * k + 2 * i + 1]);
    for (int i = threadIdx.x; i < j; i += blockDim.x) {
      rs[i] = sn[i];
      rc[i] = cs[i];
    }
  }
  // ||w||^2 from the nb partials: every thread of the (256-thread) block
  // adds a few, then the four wave sums in order
  __shared__ double wn2, wpart[4];
  {
    double t = 0;
    for (int i = threadIdx.x; i < nb; i += 256) t += nrm[i];
    t = wave_sum(t);
    if ((threadIdx.x & 63) == 0) wpart[threadIdx.x >> 6] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) wn2 = (wpart[0] + wpart[1]) + (wpart[2] + wpart[3]);
  __syncthreads();
  {
    const double hn = sqrt(wn2), scl = hn > 0 ? 1.0 / hn : 0.0;
    const i64 m = blockIdx.x * (i64)blockDim.x + threadIdx.x;
    if (m < n) {
      const double2 v = w[m];
      const double2 u = make_double2(scl * v.x, scl * v.y);
      vnext[m] = u;
      if (vcur) vcur[m] = u;
    }
  }
  if (blockIdx.x != 0) return;
  if (threadIdx.x == 0) {
    const double hn = sqrt(wn2);
    // previous rotations: [c s; -conj(s) c]
    double2 a = col[0];
    for (int i = 0; i < j; ++i) {
      const double c = rc[i];
      const double2 s = rs[i], b = col[i + 1];
      // t = c a + s b ; b' = -conj(s) a + c b
      col[i] = make_double2(c * a.x + s.x * b.x - s.y * b.y, c * a.y + s.x * b.y + s.y * b.x);
      a = make_double2(-(s.x * a.x + s.y * a.y) + c * b.x, -(s.x * a.y - s.y * a.x) + c * b.y);
    }
    const double bb = hn;  // H[j+1][j] before the rotation (real, >= 0)
    const double aa = hypot(a.x, a.y);
    const double rho = hypot(aa, bb);
    double c;
    double2 s;
    if (aa == 0) {
      c = 0;
      s = make_double2(1.0, 0.0);
    } else {
      c = aa / rho;
      // s = (a / |a|) conj(b) / rho with b real
      s = make_double2(a.x / aa * bb / rho, a.y / aa * bb / rho);
    }
    cs[j] = c;
    sn[j] = s;
    // H[j][j] = c a + s b
    col[j] = make_double2(c * a.x + s.x * bb, c * a.y + s.y * bb);
    col[k] = make_double2(0.0, 0.0);
    const double2 gj = g[j];
    // g[j+1] = -conj(s) g[j] ; g[j] = c g[j]
    const double2 gk = make_double2(-(s.x * gj.x + s.y * gj.y), -(s.x * gj.y - s.y * gj.x));
    g[k] = gk;
    g[j] = make_double2(c * gj.x, c * gj.y);
    st[2 * j] = hypot(gk.x, gk.y);
    st[2 * j + 1] = hn;
    sc[0] = hn > 0 ? 1.0 / hn : 0.0;
  }
  __syncthreads();
  for (int i = threadIdx.x; i <= k; i += blockDim.x) Hj[i] = col[i];
}

// y = R^{-1} g for the leading j x j upper triangle of the rotated H
// (column-major, leading dimension ldh): column-oriented back substitution,
// the right-hand side in LDS, one column update per step across the block.
__global__ __launch_bounds__(256) void k_backsolve(int j, int ldh, const double2* __restrict__ H,
                                                   const double2* __restrict__ g, double2* __restrict__ y) {
  extern __shared__ double2 rhs[];
  __shared__ double2 yi_s;
  for (int i = threadIdx.x; i < j; i += blockDim.x) rhs[i] = g[i];
  __syncthreads();
  for (int i = j - 1; i >= 0; --i) {
    if (threadIdx.x == 0) {
      const double2 d = H[(i64)i * ldh + i], acc = rhs[i];
      const double dd = d.x * d.x + d.y * d.y;
      const double2 yi = make_double2((acc.x * d.x + acc.y * d.y) / dd, (acc.y * d.x - acc.x * d.y) / dd);
      yi_s = yi;
      y[i] = yi;
    }
    __syncthreads();
    const double2 yi = yi_s;
    const double2* hc = H + (i64)i * ldh;
    for (int l = threadIdx.x; l < i; l += blockDim.x) {
      const double2 h = hc[l];
      rhs[l].x -= h.x * yi.x - h.y * yi.y;
      rhs[l].y -= h.x * yi.y + h.y * yi.x;
    }
    __syncthreads();
  }
}

inline unsigned nblocks(i64 n, int t) { return (unsigned)((n + t - 1) / t); }
// Caller-order copy of a block of columns: out[r + (c - c0)*n] = Z[int(r) + int(c)*ld]
__global__ void k_unpermute(i64 n, ZRef z, i64 c0, i64 nc, const int* __restrict__ iofe, double2* __restrict__ out) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= n * nc) return;
  i64 r = idx % n, c = c0 + idx / n;
  out[idx] = zat(z, iofe[r], iofe[c]);
}

__global__ void k_entries(ZRef z, i64 count, const int* __restrict__ iofe, const i64* __restrict__ m,
                          const i64* __restrict__ nn, double2* __restrict__ out) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= count) return;
  out[i] = zat(z, iofe[m[i]], iofe[nn[i]]);
}

__global__ void k_conj(i64 n, double2* v) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i < n) v[i].y = -v[i].y;
}


}  // namespace

void launch_gemv_z64(i64 n, i64 ld, const double* z, const double* x, double* y, hipStream_t s) {
  constexpr int NC = 4;
  i64 cols_per_block = 4 * NC;
  k_gemv_z64<NC><<<nblocks(n, (int)cols_per_block), 256, 0, s>>>(
      n, ld, reinterpret_cast<const double2*>(z), reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}

void launch_gemv_z32(i64 n, i64 ld, const float* z, const double* x, double* y, hipStream_t s) {
  constexpr int NC = 4;
  DM_REQUIRE((ld & 1) == 0, DM_ERR_ARGUMENT, "complex64 matrix leading dimension must be even");
  k_gemv_z32<NC><<<nblocks(n, 4 * NC), 256, 0, s>>>(
      n, ld, reinterpret_cast<const float4*>(z), reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}

void launch_permute(i64 n, const int* idx, const double* src, double* dst, int gather, hipStream_t s) {
  k_permute<<<nblocks(n, 256), 256, 0, s>>>(n, idx, reinterpret_cast<const double2*>(src),
                                             reinterpret_cast<double2*>(dst), gather);
  HIP_OK(hipGetLastError());
}

void launch_cast_c128_to_c64_matrix(i64 n, i64 ld, const double* src, float* dst, hipStream_t s) {
  k_cast_matrix<<<nblocks(ld * n, 256), 256, 0, s>>>(n, ld, reinterpret_cast<const double2*>(src),
                                                      reinterpret_cast<float2*>(dst));
  HIP_OK(hipGetLastError());
}

namespace dml {

size_t dots_part_count(i64 n, int k) { return (size_t)k * (size_t)nblocks(n, 256 * kDotRows); }

int norm_part_count(i64 n) { return (int)nblocks(n, kUpdRows); }
void dots(i64 n, i64 ldv, int k, const double* V, const double* w, double* part, hipStream_t s) {
  if (k <= 0) return;
  const int nch = (int)nblocks(n, 256 * kDotRows);
  dim3 g((unsigned)nch, (unsigned)((k + kDotVG - 1) / kDotVG));
  k_dots_part<<<g, 256, 0, s>>>(n, ldv, k, nullptr, reinterpret_cast<const double2*>(V),
                                reinterpret_cast<const double2*>(w), reinterpret_cast<double2*>(part), nch);
  HIP_OK(hipGetLastError());
}
void update(i64 n, i64 ldv, int k, const double* V, const double* part, double* h, double* w, double* nrm,
            hipStream_t s) {
  if (k <= 0) return;
  DM_REQUIRE(k <= 1024, DM_ERR_ARGUMENT, "Krylov basis larger than 1024 vectors");
  const int nch = (int)nblocks(n, 256 * kDotRows);
  k_update<<<nblocks(n, kUpdRows), 256, 0, s>>>(n, ldv, k, nullptr, 0, reinterpret_cast<const double2*>(V),
                                            reinterpret_cast<const double2*>(part), nch, reinterpret_cast<double2*>(h),
                                            reinterpret_cast<double2*>(w), nrm);
  HIP_OK(hipGetLastError());
}
void combine(i64 n, i64 ldv, int k, const double* V, const double* y, double* x, hipStream_t s) {
  if (k <= 0) return;
  DM_REQUIRE(k <= 1024, DM_ERR_ARGUMENT, "Krylov basis larger than 1024 vectors");
  k_combine<<<nblocks(n, 256), 256, 0, s>>>(n, ldv, k, reinterpret_cast<const double2*>(V),
                                             reinterpret_cast<const double2*>(y), reinterpret_cast<double2*>(x));
  HIP_OK(hipGetLastError());
}
void givens(int j, int ldh, const double* hbuf, const double* nrm, int nb, i64 n, const double* w, double* vnext,
            double* H, double* cs, double* sn, double* g, double* st, double* sc, hipStream_t s) {
  const size_t smem = sizeof(double2) * (2 * (size_t)j + 2) + sizeof(double) * j;
  DM_REQUIRE(smem <= 60 * 1024, DM_ERR_ARGUMENT, "Krylov basis too large for the device Givens step");
  k_givens<<<nblocks(n, 256), 256, smem, s>>>(j, nullptr, ldh, hbuf, nrm, nb, n,
                                              reinterpret_cast<const double2*>(w), reinterpret_cast<double2*>(vnext),
                                              nullptr, 0, nullptr, reinterpret_cast<double2*>(H), cs,
                                              reinterpret_cast<double2*>(sn), reinterpret_cast<double2*>(g), st, sc);
  HIP_OK(hipGetLastError());
}

void dots_dev(i64 n, i64 ldv, int kmax, const int* jd, const double* V, const double* w, double* part,
              hipStream_t s) {
  const int nch = (int)nblocks(n, 256 * kDotRows);
  dim3 g((unsigned)nch, (unsigned)((kmax + kDotVG - 1) / kDotVG));
  k_dots_part<<<g, 256, 0, s>>>(n, ldv, kmax, jd, reinterpret_cast<const double2*>(V),
                                reinterpret_cast<const double2*>(w), reinterpret_cast<double2*>(part), nch);
  HIP_OK(hipGetLastError());
}
void update_dev(i64 n, i64 ldv, const int* jd, int hsel, const double* V, const double* part, double* h, double* w,
                double* nrm, hipStream_t s) {
  const int nch = (int)nblocks(n, 256 * kDotRows);
  k_update<<<nblocks(n, kUpdRows), 256, 0, s>>>(n, ldv, 0, jd, hsel, reinterpret_cast<const double2*>(V),
                                            reinterpret_cast<const double2*>(part), nch, reinterpret_cast<double2*>(h),
                                            reinterpret_cast<double2*>(w), nrm);
  HIP_OK(hipGetLastError());
}
void givens_dev(int jmax, const int* jd, int ldh, const double* hbuf, const double* nrm, int nb, i64 n,
                const double* w, double* V, i64 ldv, double* vcur, double* H, double* cs, double* sn, double* g,
                double* st, double* sc, hipStream_t s) {
  const size_t smem = sizeof(double2) * (2 * (size_t)jmax + 2) + sizeof(double) * jmax;
  DM_REQUIRE(smem <= 60 * 1024, DM_ERR_ARGUMENT, "Krylov basis too large for the device Givens step");
  k_givens<<<nblocks(n, 256), 256, smem, s>>>(0, jd, ldh, hbuf, nrm, nb, n, reinterpret_cast<const double2*>(w),
                                              nullptr, reinterpret_cast<double2*>(V), ldv,
                                              reinterpret_cast<double2*>(vcur), reinterpret_cast<double2*>(H), cs,
                                              reinterpret_cast<double2*>(sn), reinterpret_cast<double2*>(g), st, sc);
  HIP_OK(hipGetLastError());
}
void copy_col_dev(i64 n, const double* src, double* base, i64 ldv, const int* jd, hipStream_t s) {
  k_copy_col<<<nblocks(n, 256), 256, 0, s>>>(n, reinterpret_cast<const double2*>(src),
                                              reinterpret_cast<double2*>(base), ldv, jd);
  HIP_OK(hipGetLastError());
}
void step_inc(int* jd, hipStream_t s) {
  k_step_inc<<<1, 64, 0, s>>>(jd);
  HIP_OK(hipGetLastError());
}

void backsolve(int j, int ldh, const double* H, const double* g, double* y, hipStream_t s) {
  const size_t smem = sizeof(double2) * (size_t)std::max(j, 1);
  DM_REQUIRE(smem <= 64 * 1024, DM_ERR_ARGUMENT, "Krylov basis too large for the device back substitution");
  k_backsolve<<<1, 256, smem, s>>>(j, ldh, reinterpret_cast<const double2*>(H),
                                   reinterpret_cast<const double2*>(g), reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}


void norm2(i64 n, const double* w, double* out, hipStream_t s) {
  k_norm2<<<1, 1024, 0, s>>>(n, reinterpret_cast<const double2*>(w), out);
  HIP_OK(hipGetLastError());
}
void axpby(i64 n, double ar, double ai, const double* x, double br, double bi, double* y, hipStream_t s) {
  k_axpby<<<nblocks(n, 256), 256, 0, s>>>(n, make_double2(ar, ai), reinterpret_cast<const double2*>(x),
                                           make_double2(br, bi), reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}
void extract_blocks(i64 n, const ZRef& z, int bs, int nb, double* blocks, hipStream_t s) {
  dim3 g(nblocks((i64)bs * bs, 256), nb);
  k_extract_blocks<<<g, 256, 0, s>>>(n, z, bs, reinterpret_cast<double2*>(blocks));
  HIP_OK(hipGetLastError());
}
void block_inverse(int bs, int nb, double* blocks, int* status, hipStream_t s) {
  size_t sm = sizeof(double2) * 2 * bs + sizeof(int) * bs;
  k_block_inverse<<<nb, 256, sm, s>>>(bs, reinterpret_cast<double2*>(blocks), status);
  HIP_OK(hipGetLastError());
}
void block_apply(i64 n, int bs, int nb, const double* inv, const double* x, double* y, hipStream_t s) {
  dim3 g(nblocks(bs, 256), nb);
  k_block_apply<<<g, 256, sizeof(double2) * bs, s>>>(n, bs, reinterpret_cast<const double2*>(inv),
                                                      reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}
void diag_inv(i64 n, const ZRef& z, double* d, hipStream_t s) {
  k_diag_inv<<<nblocks(n, 256), 256, 0, s>>>(n, z, reinterpret_cast<double2*>(d));
  HIP_OK(hipGetLastError());
}
void diag_apply(i64 n, const double* d, const double* x, double* y, hipStream_t s) {
  k_diag_apply<<<nblocks(n, 256), 256, 0, s>>>(n, reinterpret_cast<const double2*>(d), reinterpret_cast<const double2*>(x),
                                                reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}
void lu_factor(i64 n, i64 ld, double* A, int* ipiv, int* status, hipStream_t s) {
  DM_REQUIRE(n >= 1 && n < (1LL << 31) && ld >= n, DM_ERR_ARGUMENT, "dense LU: bad dimensions");
  double2* a = reinterpret_cast<double2*>(A);
  DBuf<double> ws, linv;
  ws.alloc(gemm_c128_workspace_bytes(kLuW, (int)std::max<i64>(n - kLuW, 1)) / sizeof(double));
  linv.alloc((size_t)2 * kLuW * kLuW);
  for (i64 k0 = 0; k0 < n; k0 += kLuW) {
    const int w = (int)std::min<i64>(kLuW, n - k0);
    const i64 rest = n - k0 - w;
    k_lu_panel<<<1, 1024, 0, s>>>(n, ld, (int)k0, w, a, ipiv, status);
    k_lu_laswp<<<nblocks(n, 256), 256, 0, s>>>(n, ld, (int)k0, w, a, ipiv);
    if (rest > 0) {
      k_lu_linv<<<1, 64, 0, s>>>(ld, (int)k0, w, a, reinterpret_cast<double2*>(linv.p));
      double* a12 = A + 2 * (k0 + (k0 + w) * ld);
      // U12 = L11^-1 A12 (in place: A12 is copied into the product's planes first)
      gemm_c128(w, w, linv.p, kLuW, a12, ld, (int)rest, a12, ld, ws.p, s);
      // A22 -= L21 U12
      gemm_c128(rest, w, A + 2 * ((k0 + w) + k0 * ld), ld, a12, ld, (int)rest, A + 2 * ((k0 + w) + (k0 + w) * ld),
                ld, ws.p, s, true);
    }
  }
  HIP_OK(hipGetLastError());
}
void lu_solve(i64 n, i64 ld, const double* A, const int* ipiv, double* B, i64 ldb, int nrhs, hipStream_t s) {
  if (n <= kLuSolveLds) {
    const size_t smem = (size_t)n * sizeof(double2);
    HIP_OK(hipFuncSetAttribute((const void*)k_lu_solve<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    k_lu_solve<true><<<nrhs, 1024, smem, s>>>(n, ld, reinterpret_cast<const double2*>(A), ipiv,
                                               reinterpret_cast<double2*>(B), ldb);
  } else {
    k_lu_solve<false><<<nrhs, 1024, 0, s>>>(n, ld, reinterpret_cast<const double2*>(A), ipiv,
                                             reinterpret_cast<double2*>(B), ldb);
  }
  HIP_OK(hipGetLastError());
}

void unpermute(i64 n, const ZRef& z, i64 c0, i64 nc, const int* iofe, double* out, hipStream_t s) {
  k_unpermute<<<nblocks(n * nc, 256), 256, 0, s>>>(n, z, c0, nc, iofe, reinterpret_cast<double2*>(out));
  HIP_OK(hipGetLastError());
}
void entries(const ZRef& z, i64 count, const int* iofe, const i64* m, const i64* n, double* out, hipStream_t s) {
  k_entries<<<nblocks(count, 256), 256, 0, s>>>(z, count, iofe, m, n, reinterpret_cast<double2*>(out));
  HIP_OK(hipGetLastError());
}
void conj(i64 n, double* v, hipStream_t s) {
  k_conj<<<nblocks(n, 256), 256, 0, s>>>(n, reinterpret_cast<double2*>(v));
  HIP_OK(hipGetLastError());
}
}  // namespace dml
