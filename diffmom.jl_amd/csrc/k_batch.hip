// k_batch.hip — Arnoldi kernels for lockstep batches of right-hand sides.
//
// A batch block is n x R complex128 with column r at offset r * n; a Krylov
// basis is a sequence of such blocks (block i at offset i * n * R).  Every
// kernel takes act[r]: a column with act[r] == 0 (converged, or not part of
// this solve) is skipped.  h arrays are R x kcap complex, column r's
// coefficient i at r * kcap + i.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dm_common.h"
#include "k_batch.h"

namespace {

constexpr int BT = 256;          // threads per workgroup
constexpr int BROWS = 16;        // rows per thread in the dot kernel (one wave reduction per 16 KB read)
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

// W[r] -= sum_{i<k} V_i[r] h[r, i]; each thread takes BUROWS rows (stride
// BT), so a workgroup's staging of h serves BT * BUROWS rows
constexpr int BUROWS = 4;
__global__ __launch_bounds__(BT) void k_bupdate(i64 n, int R, int k, int kcap, const double2* __restrict__ V,
                                                const double2* __restrict__ h, const int* __restrict__ act,
                                                double2* __restrict__ W) {
  const int r = blockIdx.y;
  if (!act[r]) return;
  __shared__ double2 hs[BKMAX];
  for (int i = threadIdx.x; i < k; i += BT) hs[i] = h[(i64)r * kcap + i];
  __syncthreads();
  const i64 m0 = blockIdx.x * (i64)(BT * BUROWS) + threadIdx.x;
  const i64 NR = n * R, off = (i64)r * n + m0;
  double ar[BUROWS], ai[BUROWS];
#pragma unroll
  for (int u = 0; u < BUROWS; ++u) ar[u] = ai[u] = 0.0;
  for (int i = 0; i < k; ++i) {
    const double2 b = hs[i];
#pragma unroll
    for (int u = 0; u < BUROWS; ++u) {
      if (m0 + BT * u < n) {
        const double2 a = V[(i64)i * NR + off + BT * u];
        ar[u] += a.x * b.x - a.y * b.y;
        ai[u] += a.x * b.y + a.y * b.x;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < BUROWS; ++u) {
    if (m0 + BT * u < n) {
      const double2 o = W[off + BT * u];
      W[off + BT * u] = make_double2(o.x - ar[u], o.y - ai[u]);
    }
  }
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

namespace {

// Per-column state of the device Givens step: column r's Hessenberg matrix at
// H + r ldh m (ldh = m + 1 rows, m columns), rotations at cs / sn + r m,
// rotated right-hand side at g + r ldh.
__global__ void k_bg_init(int R, int ldh, const double* __restrict__ beta, const int* __restrict__ act,
                          double2* __restrict__ g, double* __restrict__ sc, int* __restrict__ jcnt,
                          int* __restrict__ its) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= R) return;
  g[(i64)r * ldh] = make_double2(act[r] ? beta[r] : 0.0, 0.0);
  sc[r] = act[r] && beta[r] > 0 ? 1.0 / beta[r] : 0.0;
  jcnt[r] = 0;
  its[r] = 0;
}

// Step j of every active column (one block each): the Hessenberg column from
// the two CGS2 passes (h1, h2: pitch kcap) and ||w||^2 (nr), previous
// rotations applied from LDS by one thread, the new rotation, g updated; a
// column whose estimate |g[j+1]| reaches eps[r] (or that breaks down) is
// deactivated.  sc[r] = 1 / h_{j+1,j} for the next basis vector (0 when
// inactive).
__global__ __launch_bounds__(64) void k_bgivens(int j, int m, int kcap, const double2* __restrict__ h1,
                                                const double2* __restrict__ h2, const double* __restrict__ nr,
                                                double2* __restrict__ H, double* __restrict__ cs,
                                                double2* __restrict__ sn, double2* __restrict__ g,
                                                double* __restrict__ sc, int* __restrict__ act,
                                                const double* __restrict__ eps, int* __restrict__ jcnt,
                                                int* __restrict__ its, int* __restrict__ solved,
                                                double* __restrict__ est_out) {
  const int r = blockIdx.x;
  if (!act[r]) return;
  const int k = j + 1, ldh = m + 1;
  extern __shared__ double2 sh[];
  double2* col = sh;         // k + 1
  double2* rs = sh + k + 1;  // j
  double* rc = reinterpret_cast<double*>(rs + j);
  const double2* a1 = h1 + (i64)r * kcap;
  const double2* a2 = h2 + (i64)r * kcap;
  for (int i = threadIdx.x; i < k; i += blockDim.x) col[i] = make_double2(a1[i].x + a2[i].x, a1[i].y + a2[i].y);
  for (int i = threadIdx.x; i < j; i += blockDim.x) {
    rs[i] = sn[(i64)r * m + i];
    rc[i] = cs[(i64)r * m + i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double hn = sqrt(nr[r]);
    double2 a = col[0];
    for (int i = 0; i < j; ++i) {
      const double c = rc[i];
      const double2 s = rs[i], b = col[i + 1];
      col[i] = make_double2(c * a.x + s.x * b.x - s.y * b.y, c * a.y + s.x * b.y + s.y * b.x);
      a = make_double2(-(s.x * a.x + s.y * a.y) + c * b.x, -(s.x * a.y - s.y * a.x) + c * b.y);
    }
    const double bb = hn, aa = hypot(a.x, a.y), rho = hypot(aa, bb);
    double c;
    double2 s;
    if (aa == 0) {
      c = 0;
      s = make_double2(1.0, 0.0);
    } else {
      c = aa / rho;
      s = make_double2(a.x / aa * bb / rho, a.y / aa * bb / rho);
    }
    cs[(i64)r * m + j] = c;
    sn[(i64)r * m + j] = s;
    col[j] = make_double2(c * a.x + s.x * bb, c * a.y + s.y * bb);
    col[k] = make_double2(0.0, 0.0);
    double2* gr = g + (i64)r * ldh;
    const double2 gj = gr[j];
    const double2 gk = make_double2(-(s.x * gj.x + s.y * gj.y), -(s.x * gj.y - s.y * gj.x));
    gr[k] = gk;
    gr[j] = make_double2(c * gj.x, c * gj.y);
    jcnt[r] = k;
    its[r] += 1;
    const double est = hypot(gk.x, gk.y);
    if (est_out) est_out[r] = est;
    if (est <= eps[r] || hn == 0) {
      if (est <= eps[r]) solved[r] = 1;
      act[r] = 0;
      sc[r] = 0.0;
    } else {
      sc[r] = 1.0 / hn;
    }
  }
  __syncthreads();
  double2* Hj = H + (i64)r * ldh * m + (i64)j * ldh;
  for (int i = threadIdx.x; i <= k; i += blockDim.x) Hj[i] = col[i];
}

// y_r = R_r^{-1} g_r over the column's jcnt[r] steps (one block per column).
__global__ __launch_bounds__(64) void k_bbacksolve(int m, int kcap, const double2* __restrict__ H,
                                                   const double2* __restrict__ g, const int* __restrict__ jcnt,
                                                   double2* __restrict__ y) {
  const int r = blockIdx.x, jn = jcnt[r], ldh = m + 1;
  extern __shared__ double2 rhs[];
  __shared__ double2 yi_s;
  const double2* Hr = H + (i64)r * ldh * m;
  double2* yr = y + (i64)r * kcap;
  for (int i = threadIdx.x; i < jn; i += blockDim.x) rhs[i] = g[(i64)r * ldh + i];
  __syncthreads();
  for (int i = jn - 1; i >= 0; --i) {
    if (threadIdx.x == 0) {
      const double2 d = Hr[(i64)i * ldh + i], acc = rhs[i];
      const double dd = d.x * d.x + d.y * d.y;
      const double2 yi = make_double2((acc.x * d.x + acc.y * d.y) / dd, (acc.y * d.x - acc.x * d.y) / dd);
      yi_s = yi;
      yr[i] = yi;
    }
    __syncthreads();
    const double2 yi = yi_s;
    const double2* hc = Hr + (i64)i * ldh;
    for (int l = threadIdx.x; l < i; l += blockDim.x) {
      const double2 h = hc[l];
      rhs[l].x -= h.x * yi.x - h.y * yi.y;
      rhs[l].y -= h.x * yi.y + h.y * yi.x;
    }
    __syncthreads();
  }
}

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
  const dim3 g((unsigned)((n + BT * BUROWS - 1) / (BT * BUROWS)), (unsigned)R);
  k_bupdate<<<g, BT, 0, s>>>(n, R, k, kcap, reinterpret_cast<const double2*>(V), reinterpret_cast<const double2*>(h),
                             act, reinterpret_cast<double2*>(W));
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

void givens_init(int R, int m, const double* beta, const int* act, double* g, double* sc, int* jcnt, int* its,
                 hipStream_t s) {
  k_bg_init<<<(R + 63) / 64, 64, 0, s>>>(R, m + 1, beta, act, reinterpret_cast<double2*>(g), sc, jcnt, its);
  HIP_OK(hipGetLastError());
}

void givens(int R, int j, int m, int kcap, const double* h1, const double* h2, const double* nr, double* H,
            double* cs, double* sn, double* g, double* sc, int* act, const double* eps, int* jcnt, int* its,
            int* solved, hipStream_t s, double* est_out) {
  DM_REQUIRE(j < m && m < BKMAX, DM_ERR_ARGUMENT, "Krylov basis too large for the device Givens step");
  const size_t smem = sizeof(double2) * (2 * (size_t)j + 2) + sizeof(double) * j;
  k_bgivens<<<R, 64, smem, s>>>(j, m, kcap, reinterpret_cast<const double2*>(h1), reinterpret_cast<const double2*>(h2),
                                nr, reinterpret_cast<double2*>(H), cs, reinterpret_cast<double2*>(sn),
                                reinterpret_cast<double2*>(g), sc, act, eps, jcnt, its, solved, est_out);
  HIP_OK(hipGetLastError());
}

void backsolve(int R, int m, int kcap, const double* H, const double* g, const int* jcnt, double* y, hipStream_t s) {
  const size_t smem = sizeof(double2) * (size_t)std::max(m, 1);
  k_bbacksolve<<<R, 64, smem, s>>>(m, kcap, reinterpret_cast<const double2*>(H), reinterpret_cast<const double2*>(g),
                                   jcnt, reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}

void scale(i64 n, int R, const double* src, const double* sc, double* dst, hipStream_t s) {
  k_bscale<<<dim3(row_blocks(n), (unsigned)R), BT, 0, s>>>(n, reinterpret_cast<const double2*>(src), sc,
                                                           reinterpret_cast<double2*>(dst));
  HIP_OK(hipGetLastError());
}

}  // namespace dmb
