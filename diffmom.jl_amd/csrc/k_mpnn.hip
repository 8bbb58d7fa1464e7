// k_mpnn.hip — forward, backward and Adam kernels of the learned
// preconditioner's message-passing network (layout and equations in k_mpnn.h).
//
// Training products with the samples' matrices A = Z S are small (n up to a
// few thousand): a row per thread for A y, a wave per row for A^H q.
//
// Layers run 64 nodes per workgroup: the layer's two 64 x 64 weight matrices
// (bfloat16 in memory, widened to FP32 in LDS, rows padded to 65 floats), the
// nodes' hidden states and their aggregated neighbour messages are staged in
// LDS; each of the 256 threads produces 16 outputs of one node.  Weight
// gradients are sums of per-node outer products, reduced per workgroup into
// FP32 partials and across workgroups in FP64 (fixed order).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "k_mpnn.h"

namespace {

using dmn::DEG;
using dmn::FIN;
using dmn::H;

constexpr int NT = 64;       // nodes per workgroup
constexpr int HP = H + 1;    // padded LDS row
constexpr int OUTER_G = 256; // workgroups of the outer-product kernels

__device__ inline float bf(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline uint16_t to_bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ inline double bsum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
  return t;
}

// one workgroup per sample: c from r, then tau = rms |c z| and den
__global__ __launch_bounds__(256) void k_stats(const int64_t* __restrict__ ptr, int64_t node0,
                                               const double2* __restrict__ z, const double2* __restrict__ r,
                                               double* __restrict__ cden) {
  __shared__ double red[4];
  const int s = blockIdx.x;
  const int64_t a = ptr[s] - node0, b = ptr[s + 1] - node0;
  double r2 = 0, qr = 0, qi = 0, z2 = 0;
  for (int64_t i = a + threadIdx.x; i < b; i += 256) {
    const double2 v = r[i];
    const double m2 = v.x * v.x + v.y * v.y, m = sqrt(m2);
    r2 += m2;
    qr += v.x * m;
    qi += v.y * m;
    const double2 w = z[i];
    z2 += w.x * w.x + w.y * w.y;
  }
  r2 = bsum(r2, red);
  qr = bsum(qr, red);
  qi = bsum(qi, red);
  z2 = bsum(z2, red);
  if (threadIdx.x == 0) {
    double cr = 0, ci = 0;
    const double rho = b > a ? sqrt(r2 / (double)(b - a)) : 0.0;
    if (rho > 0) {
      const double qa = sqrt(qr * qr + qi * qi);
      if (qa > 0) {
        cr = qr / (qa * rho);
        ci = -qi / (qa * rho);
      } else {
        cr = 1.0 / rho;
      }
    }
    const double c2 = cr * cr + ci * ci;
    cden[4 * s] = cr;
    cden[4 * s + 1] = ci;
    cden[4 * s + 2] = c2 * r2;
    cden[4 * s + 3] = b > a ? sqrt(c2 * z2 / (double)(b - a)) : 0.0;
  }
}

__global__ void k_features(i64 n, const int* __restrict__ smp, int s0, const double* __restrict__ cden,
                           const double2* __restrict__ z, const double2* __restrict__ r, const float2* __restrict__ geom,
                           float* __restrict__ u) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = smp[i] - s0;
  const double cr = cden[4 * s], ci = cden[4 * s + 1], tau = cden[4 * s + 3];
  const double it = tau > 0 ? 1.0 / tau : 0.0;
  const double2 zv = z[i], rv = r[i];
  float* ui = u + i * FIN;
  ui[0] = (float)((cr * zv.x - ci * zv.y) * it);
  ui[1] = (float)((cr * zv.y + ci * zv.x) * it);
  ui[2] = (float)(cr * rv.x - ci * rv.y);
  ui[3] = (float)(cr * rv.y + ci * rv.x);
  const float2 g = geom[i];
  ui[4] = g.x;
  ui[5] = g.y;
}

__global__ void k_encode(i64 n, const float* __restrict__ u, const uint16_t* __restrict__ we, float* __restrict__ h) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= n * H) return;
  i64 i = idx / H;
  int o = (int)(idx % H);
  float acc = 0.f;
#pragma unroll
  for (int k = 0; k < FIN; ++k) acc += bf(we[o * FIN + k]) * u[i * FIN + k];
  h[idx] = acc;
}

// h_out = h + tanh(W_s h + W_n m + b), m_i = dinv_i sum_q sgn_iq h_{nbr_iq}
__global__ __launch_bounds__(256) void k_layer(i64 n, i64 gnode0, const int* __restrict__ nbr,
                                               const float* __restrict__ sgn, const float* __restrict__ dinv,
                                               const float* __restrict__ hin, const uint16_t* __restrict__ prm,
                                               float* __restrict__ hout, float* __restrict__ tcache,
                                               float* __restrict__ mcache) {
  __shared__ float Ws[H * HP], Wn[H * HP];
  __shared__ float bs[H];
  __shared__ float hs[NT * HP], ms[NT * HP];
  const int tid = threadIdx.x;
  const i64 i0 = (i64)blockIdx.x * NT;
  for (int e = tid; e < H * H; e += 256) {
    int o = e / H, k = e % H;
    Ws[o * HP + k] = bf(prm[e]);
    Wn[o * HP + k] = bf(prm[H * H + e]);
  }
  if (tid < H) bs[tid] = bf(prm[2 * H * H + tid]);
  for (int e = tid; e < NT * H; e += 256) {
    int nd = e / H, k = e % H;
    i64 gi = i0 + nd;
    float hv = 0.f, mv = 0.f;
    if (gi < n) {
      hv = hin[gi * H + k];
#pragma unroll
      for (int q = 0; q < DEG; ++q) {
        int j = nbr[gi * DEG + q];
        if (j >= 0) mv += sgn[gi * DEG + q] * hin[(i64)(j - gnode0) * H + k];
      }
      mv *= dinv[gi];
    }
    hs[nd * HP + k] = hv;
    ms[nd * HP + k] = mv;
  }
  __syncthreads();
  const int nd = tid >> 2, cg = tid & 3;
  float acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = bs[cg * 16 + j];
  for (int k = 0; k < H; ++k) {
    const float hk = hs[nd * HP + k], mk = ms[nd * HP + k];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int o = cg * 16 + j;
      acc[j] += Ws[o * HP + k] * hk + Wn[o * HP + k] * mk;
    }
  }
  const i64 gi = i0 + nd;
  if (gi >= n) return;
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int o = cg * 16 + j;
    const float tt = tanhf(acc[j]);
    hout[gi * H + o] = hs[nd * HP + o] + tt;
    if (tcache) tcache[gi * H + o] = tt;
    if (mcache) mcache[gi * H + o] = ms[nd * HP + o];
  }
}

__global__ void k_decode(i64 n, const float* __restrict__ h, const uint16_t* __restrict__ wd, float* __restrict__ y) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a = 0.f, b = 0.f;
  for (int k = 0; k < H; ++k) {
    float hv = h[i * H + k];
    a += bf(wd[k]) * hv;
    b += bf(wd[H + k]) * hv;
  }
  y[2 * i] = a;
  y[2 * i + 1] = b;
}

__global__ __launch_bounds__(1024) void k_sum(i64 n, const double* __restrict__ v, double* __restrict__ out) {
  __shared__ double red[16];
  double s = 0;
  for (i64 i = threadIdx.x; i < n; i += 1024) s += v[i];
  s = bsum(s, red);
  if (threadIdx.x == 0) out[0] = s;
}

// gh = W_d^T gy
__global__ void k_decode_back(i64 n, const float2* __restrict__ gy, const uint16_t* __restrict__ wd,
                              float* __restrict__ gh) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= n * H) return;
  i64 i = idx / H;
  int k = (int)(idx % H);
  float2 g = gy[i];
  gh[idx] = bf(wd[k]) * g.x + bf(wd[H + k]) * g.y;
}

// ga = gh (1 - t^2); gh_out = gh + W_s^T ga; gm = dinv W_n^T ga
__global__ __launch_bounds__(256) void k_layer_back(i64 n, const float* __restrict__ gh, const float* __restrict__ t,
                                                    const float* __restrict__ dinv, const uint16_t* __restrict__ prm,
                                                    float* __restrict__ gh_out, float* __restrict__ gm) {
  __shared__ float Ws[H * HP], Wn[H * HP];
  __shared__ float gas[NT * HP];
  const int tid = threadIdx.x;
  const i64 i0 = (i64)blockIdx.x * NT;
  for (int e = tid; e < H * H; e += 256) {
    int o = e / H, k = e % H;
    Ws[o * HP + k] = bf(prm[e]);
    Wn[o * HP + k] = bf(prm[H * H + e]);
  }
  for (int e = tid; e < NT * H; e += 256) {
    int nd = e / H, o = e % H;
    i64 gi = i0 + nd;
    float v = 0.f;
    if (gi < n) {
      float tt = t[gi * H + o];
      v = gh[gi * H + o] * (1.f - tt * tt);
    }
    gas[nd * HP + o] = v;
  }
  __syncthreads();
  const int nd = tid >> 2, cg = tid & 3;
  float as[16], an[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    as[j] = 0.f;
    an[j] = 0.f;
  }
  for (int o = 0; o < H; ++o) {
    const float g = gas[nd * HP + o];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int k = cg * 16 + j;
      as[j] += Ws[o * HP + k] * g;
      an[j] += Wn[o * HP + k] * g;
    }
  }
  const i64 gi = i0 + nd;
  if (gi >= n) return;
  const float di = dinv[gi];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int k = cg * 16 + j;
    gh_out[gi * H + k] = gh[gi * H + k] + as[j];
    gm[gi * H + k] = an[j] * di;
  }
}

// gh_out_j += sum_q sgn_jq gm_{nbr_jq}  (adjoint of the aggregation; the graph is symmetric)
__global__ void k_scatter(i64 n, i64 gnode0, const int* __restrict__ nbr, const float* __restrict__ sgn,
                          const float* __restrict__ gm, float* __restrict__ gh_out) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= n * H) return;
  i64 i = idx / H;
  int k = (int)(idx % H);
  float acc = 0.f;
#pragma unroll
  for (int q = 0; q < DEG; ++q) {
    int j = nbr[i * DEG + q];
    if (j >= 0) acc += sgn[i * DEG + q] * gm[(i64)(j - gnode0) * H + k];
  }
  gh_out[idx] += acc;
}

// part[b][a * KB1 + c] = sum over this workgroup's nodes of A'[i][a] B[i][c],
// A' = A (1 - T^2) when T is given, B column KB (bias) = 1 when bias.
template <int LA, int KB>
__global__ __launch_bounds__(256) void k_outer(i64 n, const float* __restrict__ A, const float* __restrict__ T,
                                               const float* __restrict__ B, int bias, float* __restrict__ part) {
  constexpr int KB1 = KB + 1;
  constexpr int O = LA * KB1;
  constexpr int PER = (O + 255) / 256;
  __shared__ float As[NT * LA];
  __shared__ float Bs[NT * KB1];
  const int tid = threadIdx.x;
  float acc[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) acc[j] = 0.f;
  const i64 ntile = (n + NT - 1) / NT;
  for (i64 tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
    const i64 i0 = tile * NT;
    __syncthreads();
    for (int e = tid; e < NT * LA; e += 256) {
      int nd = e / LA, a = e % LA;
      i64 gi = i0 + nd;
      float v = 0.f;
      if (gi < n) {
        v = A[gi * LA + a];
        if (T) {
          float tt = T[gi * LA + a];
          v *= 1.f - tt * tt;
        }
      }
      As[nd * LA + a] = v;
    }
    for (int e = tid; e < NT * KB1; e += 256) {
      int nd = e / KB1, c = e % KB1;
      i64 gi = i0 + nd;
      Bs[nd * KB1 + c] = gi < n ? (c < KB ? B[gi * KB + c] : 1.f) : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int o = tid + 256 * j;
      if (o < O) {
        const int a = o / KB1, c = o % KB1;
        if (c < KB || bias) {
          float s = 0.f;
          for (int nd = 0; nd < NT; ++nd) s += As[nd * LA + a] * Bs[nd * KB1 + c];
          acc[j] += s;
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int o = tid + 256 * j;
    if (o < O) part[(i64)blockIdx.x * O + o] = acc[j];
  }
}

// grad[woff + a * KB + c] (c < KB), grad[boff + a] (c == KB, bias) from G partials
__global__ void k_outer_reduce(int G, int LA, int KB, int bias, const float* __restrict__ part, float* __restrict__ grad,
                               int woff, int boff) {
  const int KB1 = KB + 1, O = LA * KB1;
  int o = blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= O) return;
  const int a = o / KB1, c = o % KB1;
  if (c == KB && !bias) return;
  double s = 0;
  for (int b = 0; b < G; ++b) s += part[(i64)b * O + o];
  if (c < KB)
    grad[woff + a * KB + c] = (float)s;
  else
    grad[boff + a] = (float)s;
}

__global__ void k_adam(int np, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float b1t, float b2t, float lr, uint16_t* __restrict__ bf16) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float gi = g[i];
  float mi = b1 * m[i] + (1.f - b1) * gi;
  float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
  float pi = p[i] - lr * mh / (sqrtf(vh) + eps);
  p[i] = pi;
  bf16[i] = to_bf(pi);
}

__global__ void k_to_bf16(int np, const float* __restrict__ p, uint16_t* __restrict__ bf16) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) bf16[i] = to_bf(p[i]);
}

// d = (y0 + i y1) / c = (y0 + i y1) conj(c) / |c|^2
__global__ void k_correction(i64 n, const float2* __restrict__ y, const double* __restrict__ cden,
                             double2* __restrict__ d) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double cr = cden[0], ci = cden[1];
  const double c2 = cr * cr + ci * ci;
  if (!(c2 > 0)) {
    d[i] = make_double2(0.0, 0.0);
    return;
  }
  const float2 yv = y[i];
  d[i] = make_double2(((double)yv.x * cr + (double)yv.y * ci) / c2, ((double)yv.y * cr - (double)yv.x * ci) / c2);
}

__global__ void k_identity(i64 n, double2* __restrict__ out) {
  const i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx < n * n) out[idx] = make_double2(idx % n == idx / n ? 1.0 : 0.0, 0.0);
}

__global__ void k_add(i64 n, const double2* __restrict__ a, const double2* __restrict__ b, double2* __restrict__ out) {
  i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i < n) out[i] = make_double2(a[i].x + b[i].x, a[i].y + b[i].y);
}

// out[:, c] = A x[:, c] for c < k: one row per thread, the row's k sums in
// FP64; A is read down its columns, which a wave reads as consecutive rows.
template <int KMAX, class TX>
__global__ __launch_bounds__(256) void k_acols(i64 n, int k, const float2* __restrict__ A, const TX* __restrict__ x,
                                               double2* __restrict__ out) {
  const i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double ar[KMAX], ai[KMAX];
#pragma unroll
  for (int c = 0; c < KMAX; ++c) ar[c] = ai[c] = 0.0;
  for (i64 j = 0; j < n; ++j) {
    const float2 af = A[i + j * n];
    const double zr = af.x, zi = af.y;
#pragma unroll
    for (int c = 0; c < KMAX; ++c) {
      if (c < k) {
        const TX xv = x[j + c * n];
        const double xr = xv.x, xi = xv.y;
        ar[c] += zr * xr - zi * xi;
        ai[c] += zr * xi + zi * xr;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < KMAX; ++c)
    if (c < k) out[i + c * n] = make_double2(ar[c], ai[c]);
}

// out[:, c] = A^H x[:, c] for c < k: one wave per output row i (column i of
// A, read contiguously by the lanes), FP64 lane sums reduced across the wave.
template <int KMAX>
__global__ __launch_bounds__(256) void k_acols_adj(i64 n, int k, const float2* __restrict__ A,
                                                   const float2* __restrict__ x, double2* __restrict__ out) {
  const i64 i = blockIdx.x * (i64)(blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  double ar[KMAX], ai[KMAX];
#pragma unroll
  for (int c = 0; c < KMAX; ++c) ar[c] = ai[c] = 0.0;
  const float2* col = A + i * n;
  for (i64 j = lane; j < n; j += 64) {
    const float2 af = col[j];
    const double zr = af.x, zi = -(double)af.y;
#pragma unroll
    for (int c = 0; c < KMAX; ++c) {
      if (c < k) {
        const float2 xv = x[j + c * n];
        ar[c] += zr * xv.x - zi * xv.y;
        ai[c] += zr * xv.y + zi * xv.x;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < KMAX; ++c) {
    for (int o = 32; o > 0; o >>= 1) {
      ar[c] += __shfl_xor(ar[c], o, 64);
      ai[c] += __shfl_xor(ai[c], o, 64);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int c = 0; c < KMAX; ++c)
      if (c < k) out[i + c * n] = make_double2(ar[c], ai[c]);
  }
}

__global__ void k_gather_c64(i64 n, const double2* __restrict__ a, const int* __restrict__ iofe,
                             float2* __restrict__ out) {
  const i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= n * n) return;
  const i64 r = idx % n, c = idx / n;
  const double2 v = a[(i64)iofe[r] + (i64)iofe[c] * n];
  out[idx] = make_float2((float)v.x, (float)v.y);
}

// p = r - Ar
__global__ void k_res_rhs(i64 nk, const double2* __restrict__ r, const double2* __restrict__ ar,
                          double2* __restrict__ p) {
  const i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (i >= nk) return;
  p[i] = make_double2(r[i].x - ar[i].x, r[i].y - ar[i].y);
}

// q = Ay - c p, parts = |q|^2 / (den ns)
__global__ void k_res_q(i64 n, int k, const double2* __restrict__ zy, const double2* __restrict__ p,
                        const double* __restrict__ cden, int ns, float2* __restrict__ q, double* __restrict__ parts) {
  const i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= n * k) return;
  const int c = (int)(idx / n);
  const double cr = cden[4 * c], ci = cden[4 * c + 1], den = cden[4 * c + 2];
  const double2 pv = p[idx], v = zy[idx];
  const double qr = v.x - (cr * pv.x - ci * pv.y), qi = v.y - (cr * pv.y + ci * pv.x);
  q[idx] = make_float2((float)qr, (float)qi);
  parts[idx] = den > 0 ? (qr * qr + qi * qi) / den / ns : 0.0;
}

// gy = 2 A^H q / (den ns)
__global__ void k_res_gy(i64 n, int k, const double2* __restrict__ g, const double* __restrict__ cden, int ns,
                         float2* __restrict__ gy) {
  const i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= n * k) return;
  const double den = cden[4 * (idx / n) + 2];
  const double f = den > 0 ? 2.0 / den / ns : 0.0;
  gy[idx] = make_float2((float)(f * g[idx].x), (float)(f * g[idx].y));
}

unsigned nb256(i64 n) { return (unsigned)((n + 255) / 256); }

constexpr int ZC_K = 8;  // columns per pass of the small products

template <class TX>
void acols(i64 n, int k, const float* A, const TX* x, double* out, hipStream_t s) {
  for (int c0 = 0; c0 < k; c0 += ZC_K) {
    const int kk = std::min(ZC_K, k - c0);
    k_acols<ZC_K, TX><<<nb256(n), 256, 0, s>>>(n, kk, reinterpret_cast<const float2*>(A), x + (i64)c0 * n,
                                               reinterpret_cast<double2*>(out) + (i64)c0 * n);
  }
}

void acols_adj(i64 n, int k, const float* A, const float2* x, double* out, hipStream_t s) {
  for (int c0 = 0; c0 < k; c0 += ZC_K) {
    const int kk = std::min(ZC_K, k - c0);
    k_acols_adj<ZC_K><<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, kk, reinterpret_cast<const float2*>(A),
                                                              x + (i64)c0 * n,
                                                              reinterpret_cast<double2*>(out) + (i64)c0 * n);
  }
}

template <int LA, int KB>
void outer(i64 n, const float* A, const float* T, const float* B, int bias, float* part, float* grad, int woff,
           int boff, hipStream_t s) {
  const int G = (int)std::min<i64>(OUTER_G, std::max<i64>(1, (n + NT - 1) / NT));
  k_outer<LA, KB><<<G, 256, 0, s>>>(n, A, T, B, bias, part);
  const int O = LA * (KB + 1);
  k_outer_reduce<<<(O + 255) / 256, 256, 0, s>>>(G, LA, KB, bias, part, grad, woff, boff);
}

}  // namespace

namespace dmn {

void stats(int ns, const int64_t* ptr, int64_t node0, const double* z, const double* r, double* cden, hipStream_t s) {
  if (ns <= 0) return;
  k_stats<<<ns, 256, 0, s>>>(ptr, node0, reinterpret_cast<const double2*>(z), reinterpret_cast<const double2*>(r),
                             cden);
  HIP_OK(hipGetLastError());
}

void features(i64 n, const int* smp, int s0, const double* cden, const double* z, const double* r, const float* geom,
              float* u, hipStream_t s) {
  k_features<<<nb256(n), 256, 0, s>>>(n, smp, s0, cden, reinterpret_cast<const double2*>(z),
                                      reinterpret_cast<const double2*>(r), reinterpret_cast<const float2*>(geom), u);
  HIP_OK(hipGetLastError());
}

void forward(const Graph& g, i64 gnode0, const uint16_t* prm, const Cache& c, float* hbuf, hipStream_t s) {
  const i64 n = g.n;
  // training keeps every layer's input in c.h; inference ping-pongs between
  // c.h (layers 0, 2, 4) and hbuf (layers 1, 3)
  auto hl = [&](int l) -> float* {
    if (c.t) return c.h + (i64)l * n * H;
    return (l & 1) ? hbuf : c.h;
  };
  k_encode<<<nb256(n * H), 256, 0, s>>>(n, c.u, prm + OFF_WE, hl(0));
  const unsigned gl = (unsigned)((n + NT - 1) / NT);
  for (int l = 0; l < L; ++l) {
    k_layer<<<gl, 256, 0, s>>>(n, gnode0, g.nbr, g.sgn, g.dinv, hl(l), prm + off_ws(l), hl(l + 1),
                               c.t ? c.t + (i64)l * n * H : nullptr, c.m ? c.m + (i64)l * n * H : nullptr);
  }
  k_decode<<<nb256(n), 256, 0, s>>>(n, hl(L), prm + OFF_WD, c.y);
  HIP_OK(hipGetLastError());
}

void sum(i64 n, const double* parts, double* out, hipStream_t s) {
  k_sum<<<1, 1024, 0, s>>>(n, parts, out);
  HIP_OK(hipGetLastError());
}

size_t backward_floats(i64 n) { return (size_t)3 * n * H; }
size_t outer_floats() { return (size_t)OUTER_G * H * (H + 1); }

void backward(const Graph& g, i64 gnode0, const uint16_t* prm, const Cache& c, const float* gy, float* grad,
              float* work, float* part, hipStream_t s) {
  const i64 n = g.n;
  float* gh = work;
  float* gh2 = work + n * H;
  float* gm = work + 2 * n * H;
  outer<2, H>(n, gy, nullptr, c.h + (i64)L * n * H, 0, part, grad, OFF_WD, 0, s);
  k_decode_back<<<nb256(n * H), 256, 0, s>>>(n, reinterpret_cast<const float2*>(gy), prm + OFF_WD, gh);
  const unsigned gl = (unsigned)((n + NT - 1) / NT);
  for (int l = L - 1; l >= 0; --l) {
    const float* hin = c.h + (i64)l * n * H;
    const float* tl = c.t + (i64)l * n * H;
    const float* ml = c.m + (i64)l * n * H;
    outer<H, H>(n, gh, tl, hin, 1, part, grad, off_ws(l), off_b(l), s);
    outer<H, H>(n, gh, tl, ml, 0, part, grad, off_wn(l), 0, s);
    k_layer_back<<<gl, 256, 0, s>>>(n, gh, tl, g.dinv, prm + off_ws(l), gh2, gm);
    k_scatter<<<nb256(n * H), 256, 0, s>>>(n, gnode0, g.nbr, g.sgn, gm, gh2);
    std::swap(gh, gh2);
  }
  outer<H, FIN>(n, gh, nullptr, c.u, 0, part, grad, OFF_WE, 0, s);
  HIP_OK(hipGetLastError());
}

void adam(float* p, const float* grad, float* m, float* v, int step, float lr, uint16_t* bf16, hipStream_t s) {
  const float b1t = powf(0.9f, (float)step), b2t = powf(0.999f, (float)step);
  k_adam<<<(NPARAM + 255) / 256, 256, 0, s>>>(NPARAM, p, grad, m, v, b1t, b2t, lr, bf16);
  HIP_OK(hipGetLastError());
}

void to_bf16(const float* p, uint16_t* bf16, hipStream_t s) {
  k_to_bf16<<<(NPARAM + 255) / 256, 256, 0, s>>>(NPARAM, p, bf16);
  HIP_OK(hipGetLastError());
}

void gather_c64(i64 n, const double* a_int, const int* iofe, float* out, hipStream_t s) {
  k_gather_c64<<<nb256(n * n), 256, 0, s>>>(n, reinterpret_cast<const double2*>(a_int), iofe,
                                            reinterpret_cast<float2*>(out));
  HIP_OK(hipGetLastError());
}

void res_rhs(i64 n, int k, const float* A, const double* r, double* p, hipStream_t s) {
  // A r into p, then p = r - p
  acols(n, k, A, reinterpret_cast<const double2*>(r), p, s);
  k_res_rhs<<<nb256(n * k), 256, 0, s>>>(n * k, reinterpret_cast<const double2*>(r),
                                         reinterpret_cast<const double2*>(p), reinterpret_cast<double2*>(p));
  HIP_OK(hipGetLastError());
}

void res_loss(i64 n, int k, const float* A, const float* y, const double* p, const double* cden, int ns, float* q,
              double* parts, float* gy, double* work, hipStream_t s) {
  acols(n, k, A, reinterpret_cast<const float2*>(y), work, s);
  k_res_q<<<nb256(n * k), 256, 0, s>>>(n, k, reinterpret_cast<const double2*>(work),
                                       reinterpret_cast<const double2*>(p), cden, ns, reinterpret_cast<float2*>(q),
                                       parts);
  if (gy) {
    acols_adj(n, k, A, reinterpret_cast<const float2*>(q), work, s);
    k_res_gy<<<nb256(n * k), 256, 0, s>>>(n, k, reinterpret_cast<const double2*>(work), cden, ns,
                                          reinterpret_cast<float2*>(gy));
  }
  HIP_OK(hipGetLastError());
}

void correction(i64 n, const float* y, const double* cden, double* d, hipStream_t s) {
  k_correction<<<nb256(n), 256, 0, s>>>(n, reinterpret_cast<const float2*>(y), cden, reinterpret_cast<double2*>(d));
  HIP_OK(hipGetLastError());
}

void identity(i64 n, double* out, hipStream_t s) {
  k_identity<<<nb256(n * n), 256, 0, s>>>(n, reinterpret_cast<double2*>(out));
  HIP_OK(hipGetLastError());
}

void add(i64 n, const double* a, const double* b, double* out, hipStream_t s) {
  k_add<<<nb256(n), 256, 0, s>>>(n, reinterpret_cast<const double2*>(a), reinterpret_cast<const double2*>(b),
                                 reinterpret_cast<double2*>(out));
  HIP_OK(hipGetLastError());
}

}  // namespace dmn
