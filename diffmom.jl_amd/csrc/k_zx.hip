// k_zx.hip — multi-right-hand-side product Y = Z X on MFMA (complex64 Z).
//
// The complex64 matrix is column-major with interleaved (re, im) parts, i.e.
// it already is the real 2N x N matrix A' with rows (Re z_m, Im z_m).  With
// P = A' Re(X) and Q = A' Im(X) (two real products sharing every A' tile):
//   Re Y_m = P_{2m} - Q_{2m+1},   Im Y_m = Q_{2m} + P_{2m+1}
// so the complex product costs exactly 8 N^2 R flops with no repacking of Z.
// Tiles: a workgroup (4 waves) owns 64 complex rows x 64 right-hand sides for
// one K chunk; each wave owns 32 real rows x 64 columns of P and Q as 2x4x2
// v_mfma_f32_16x16x4f32 accumulators.  A' and X tiles are staged through LDS.
// K chunks of KC columns accumulate in FP32 and are summed in FP64 by a second
// kernel, which bounds the FP32 accumulation length.
#include <hip/hip_runtime.h>

#include "dm_common.h"
#include "k_linalg.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int ZX_BM = 64;   // complex rows per workgroup
constexpr int ZX_BN = 64;   // right-hand sides per workgroup
constexpr int ZX_BK = 16;   // complex columns of Z per LDS stage
constexpr int ZX_KC = 4096; // columns per FP32 accumulation chunk

constexpr int ZX_BK2 = 32;  // complex columns of Z per stage (FP32 kernel)

// Xr, Xi: np x R real planes of the right-hand sides (FP32; rows n .. np - 1
// zero, np a multiple of 4).  part: [kchunk][R][n] complex128 partials.
// Stages of ZX_BK2 columns: 16-byte global loads into registers for stage
// s + 1 while stage s is multiplied out of LDS (two LDS buffers, one barrier
// per stage).
__global__ __launch_bounds__(256) void k_zx_mfma(i64 n, i64 ld, const float* __restrict__ Z, const float* __restrict__ Xr,
                                                 const float* __restrict__ Xi, i64 np, int R, double2* __restrict__ part) {
  __shared__ float As[2][ZX_BK2][2 * ZX_BM + 4];
  __shared__ float Br[2][ZX_BK2][ZX_BN + 4];
  __shared__ float Bi[2][ZX_BK2][ZX_BN + 4];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const i64 m0 = (i64)blockIdx.x * ZX_BM;   // complex row base
  const int c0 = blockIdx.y * ZX_BN;        // rhs base
  const i64 k_begin = (i64)blockIdx.z * ZX_KC;
  const i64 k_end = k_begin + ZX_KC < n ? k_begin + ZX_KC : n;
  const int nst = (int)((k_end - k_begin + ZX_BK2 - 1) / ZX_BK2);
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 ra[4], rb[2], rc[2];
  // stage loads: A' 32 columns x 128 real rows (thread: 4 x float4); X planes
  // 64 rhs x 32 rows (thread: 2 x float4 each)
  auto gload = [&](i64 k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q, kk = idx >> 5, v = idx & 31;
      const i64 col = k0 + kk;
      ra[q] = col < k_end ? *reinterpret_cast<const f4*>(Z + col * 2 * ld + 2 * m0 + 4 * v) : f4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = tid + 256 * q, jj = idx >> 3, v = idx & 7;
      const i64 row = k0 + 4 * v;
      const int col = c0 + jj;
      const bool ok = col < R && row < np;
      rb[q] = ok ? *reinterpret_cast<const f4*>(Xr + row + (i64)col * np) : f4{0.f, 0.f, 0.f, 0.f};
      rc[q] = ok ? *reinterpret_cast<const f4*>(Xi + row + (i64)col * np) : f4{0.f, 0.f, 0.f, 0.f};
    }
  };
  auto lstore = [&](int b) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q, kk = idx >> 5, v = idx & 31;
      *reinterpret_cast<f4*>(&As[b][kk][4 * v]) = ra[q];
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = tid + 256 * q, jj = idx >> 3, v = idx & 7;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        Br[b][4 * v + e][jj] = rb[q][e];
        Bi[b][4 * v + e][jj] = rc[q][e];
      }
    }
  };
  v4f accP[2][4], accQ[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      accP[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};
      accQ[a][b] = (v4f){0.f, 0.f, 0.f, 0.f};
    }
  gload(k_begin);
  lstore(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int b = st & 1;
    if (st + 1 < nst) gload(k_begin + (i64)(st + 1) * ZX_BK2);
#pragma unroll
    for (int kq = 0; kq < ZX_BK2 / 4; ++kq) {
      const int kk = 4 * kq + (lane >> 4);
      float a[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) a[r] = As[b][kk][32 * wave + 16 * r + (lane & 15)];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const float br = Br[b][kk][16 * cb + (lane & 15)];
        const float bi = Bi[b][kk][16 * cb + (lane & 15)];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          accP[r][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], br, accP[r][cb], 0, 0, 0);
          accQ[r][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], bi, accQ[r][cb], 0, 0, 0);
        }
      }
    }
    if (st + 1 < nst) lstore(b ^ 1);
    __syncthreads();
  }
  // lane holds real rows 4*(lane/16) + r (r = 0..3) of each 16-row block and
  // column lane%16: rows (0,1) and (2,3) are the (re, im) rows of two complex rows
  double2* out = part + (i64)blockIdx.z * (i64)R * n;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int col = c0 + 16 * cb + (lane & 15);
      if (col >= R) continue;
      v4f p = accP[r][cb], q = accQ[r][cb];
      const i64 real_row = 32 * wave + 16 * r + 4 * (lane >> 4);  // even
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        i64 m = m0 + (real_row >> 1) + h;
        if (m >= n) continue;
        float pr = h ? p[2] : p[0], pi = h ? p[3] : p[1];
        float qr = h ? q[2] : q[0], qi = h ? q[3] : q[1];
        out[m + (i64)col * n] = make_double2((double)pr - (double)qi, (double)qr + (double)pi);
      }
    }
}

__global__ void k_zx_reduce(i64 n, int R, int nk, const double2* __restrict__ part, double2* __restrict__ Y, i64 ldy) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= n * R) return;
  i64 m = idx % n, j = idx / n;
  double sr = 0, si = 0;
  for (int c = 0; c < nk; ++c) {
    double2 v = part[(i64)c * R * n + idx];
    sr += v.x;
    si += v.y;
  }
  Y[m + j * ldy] = make_double2(sr, si);
}

// complex128 n x R (ldx) -> FP32 real / imaginary planes (n x R)
// rows n .. np - 1 of each plane column are zero
__global__ void k_split_planes(i64 n, i64 np, int R, const double2* __restrict__ X, i64 ldx, float* __restrict__ Xr,
                               float* __restrict__ Xi) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= np * R) return;
  i64 m = idx % np, j = idx / np;
  double2 v = m < n ? X[m + j * ldx] : make_double2(0.0, 0.0);
  Xr[idx] = (float)v.x;
  Xi[idx] = (float)v.y;
}


typedef double v4d __attribute__((ext_vector_type(4)));

// FP64 variant (complex128 Z, v_mfma_f64_16x16x4f64): the refinement residual
// of many right-hand sides.  Same tiling; one K chunk (FP64 accumulation).
// Rows mr (complex) and inner dimension n may differ: Y (mr x R) = Z (mr x n,
// leading dimension ld) X (n x R).
__global__ __launch_bounds__(256, 2) void k_zx_mfma64(i64 mr, i64 n, i64 ld, const double* __restrict__ Z,
                                                   const double* __restrict__ Xr, const double* __restrict__ Xi, i64 np,
                                                   int R, double2* __restrict__ Y, i64 ldy) {
  // stages of ZX_BK columns: 16-byte loads (one complex of Z, two rows of a
  // plane) into registers for stage s + 1 while stage s is multiplied out of
  // LDS (two buffers, one barrier per stage).  Xr, Xi: np x R, rows past n zero.
  __shared__ double As[2][ZX_BK][2 * ZX_BM + 2];
  __shared__ double Br[2][ZX_BK][ZX_BN + 2];
  __shared__ double Bi[2][ZX_BK][ZX_BN + 2];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const i64 m0 = (i64)blockIdx.x * ZX_BM;
  const int c0 = blockIdx.y * ZX_BN;
  const int nst = (int)((n + ZX_BK - 1) / ZX_BK);
  double2 ra[4], rb[2], rc[2];
  auto gload = [&](i64 k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q, kk = idx >> 6, v = idx & 63;
      const i64 col = k0 + kk, m = m0 + v;
      ra[q] = (col < n && m < mr) ? *reinterpret_cast<const double2*>(Z + col * 2 * ld + 2 * m) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = tid + 256 * q, jj = idx >> 3, v = idx & 7;
      const i64 row = k0 + 2 * v;
      const int col = c0 + jj;
      const bool ok = col < R && row < np;
      rb[q] = ok ? *reinterpret_cast<const double2*>(Xr + row + (i64)col * np) : make_double2(0.0, 0.0);
      rc[q] = ok ? *reinterpret_cast<const double2*>(Xi + row + (i64)col * np) : make_double2(0.0, 0.0);
    }
  };
  auto lstore = [&](int b) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int idx = tid + 256 * q, kk = idx >> 6, v = idx & 63;
      As[b][kk][2 * v] = ra[q].x;
      As[b][kk][2 * v + 1] = ra[q].y;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int idx = tid + 256 * q, jj = idx >> 3, v = idx & 7;
      Br[b][2 * v][jj] = rb[q].x;
      Br[b][2 * v + 1][jj] = rb[q].y;
      Bi[b][2 * v][jj] = rc[q].x;
      Bi[b][2 * v + 1][jj] = rc[q].y;
    }
  };
  v4d accP[2][4], accQ[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      accP[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
      accQ[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
    }
  gload(0);
  lstore(0);
  __syncthreads();
  for (int st = 0; st < nst; ++st) {
    const int b = st & 1;
    if (st + 1 < nst) gload((i64)(st + 1) * ZX_BK);
#pragma unroll
    for (int kq = 0; kq < ZX_BK / 4; ++kq) {
      const int kk = 4 * kq + (lane >> 4);
      double a[2];
#pragma unroll
      for (int r = 0; r < 2; ++r) a[r] = As[b][kk][32 * wave + 16 * r + (lane & 15)];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const double br = Br[b][kk][16 * cb + (lane & 15)];
        const double bi = Bi[b][kk][16 * cb + (lane & 15)];
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          accP[r][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], br, accP[r][cb], 0, 0, 0);
          accQ[r][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], bi, accQ[r][cb], 0, 0, 0);
        }
      }
    }
    if (st + 1 < nst) lstore(b ^ 1);
    __syncthreads();
  }
  // FP64 MFMA output layout: lane l holds real rows (l / 16) + 4 r (r = 0..3)
  // of each 16-row block, column l % 16.  The (re, im) rows 2m, 2m + 1 of one
  // complex row sit in lanes l and l ^ 16: swap, then the even lane writes.
  const int g = lane >> 4;
#pragma unroll
  for (int r2 = 0; r2 < 2; ++r2)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int col = c0 + 16 * cb + (lane & 15);
      v4d p = accP[r2][cb], q = accQ[r2][cb];
      double po[4], qo[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        po[r] = __shfl_xor(p[r], 16, 64);
        qo[r] = __shfl_xor(q[r], 16, 64);
      }
      if (col >= R || (g & 1)) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const i64 m = m0 + ((32 * wave + 16 * r2 + g + 4 * r) >> 1);
        if (m >= mr) continue;
        Y[m + (i64)col * ldy] = make_double2(p[r] - qo[r], q[r] + po[r]);
      }
    }
}

__global__ void k_split_planes64(i64 n, i64 np, int R, const double2* __restrict__ X, i64 ldx,
                                 double* __restrict__ Xr, double* __restrict__ Xi) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= np * R) return;
  i64 m = idx % np, j = idx / np;
  double2 v = m < n ? X[m + j * ldx] : make_double2(0.0, 0.0);
  Xr[idx] = v.x;
  Xi[idx] = v.y;
}

}  // namespace

namespace dml {

size_t zx_workspace_bytes(i64 n, int R) {
  i64 nk = (n + ZX_KC - 1) / ZX_KC;
  const i64 np = (n + 31) / 32 * 32;
  size_t a = (size_t)nk * R * n * sizeof(double2) + 2 * (size_t)np * R * sizeof(float) + 64;
  size_t b = 2 * (size_t)((n + 15) / 16 * 16) * R * sizeof(double);
  return a > b ? a : b;
}

size_t gemm_c128_workspace_bytes(i64 n, int R) { return 2 * (size_t)((n + 15) / 16 * 16) * R * sizeof(double); }

// Y = Z64 X with FP64 MFMA (complex128 everything).
void zx_mfma64(i64 n, i64 ld, const double* z64, const double* X, i64 ldx, int R, double* Y, i64 ldy, void* ws,
               hipStream_t s) {
  const i64 np = (n + 15) / 16 * 16;
  double* Xr = reinterpret_cast<double*>(ws);
  double* Xi = Xr + np * R;
  k_split_planes64<<<(unsigned)((np * R + 255) / 256), 256, 0, s>>>(n, np, R, reinterpret_cast<const double2*>(X), ldx,
                                                                     Xr, Xi);
  dim3 grid((unsigned)((n + ZX_BM - 1) / ZX_BM), (unsigned)((R + ZX_BN - 1) / ZX_BN), 1);
  DM_REQUIRE(grid.y <= 65535, DM_ERR_ARGUMENT, "too many right-hand sides for one product");
  k_zx_mfma64<<<grid, 256, 0, s>>>(n, n, ld, z64, Xr, Xi, np, R, reinterpret_cast<double2*>(Y), ldy);
  HIP_OK(hipGetLastError());
}

void gemm_c128(i64 mr, i64 n, const double* A, i64 lda, const double* X, i64 ldx, int R, double* Y, i64 ldy,
               void* ws, hipStream_t s) {
  DM_REQUIRE(lda >= mr && ldx >= n && ldy >= mr && R >= 1, DM_ERR_ARGUMENT, "complex GEMM: bad dimensions");
  const i64 np = (n + 15) / 16 * 16;
  double* Xr = reinterpret_cast<double*>(ws);
  double* Xi = Xr + np * R;
  k_split_planes64<<<(unsigned)((np * R + 255) / 256), 256, 0, s>>>(n, np, R, reinterpret_cast<const double2*>(X), ldx,
                                                                     Xr, Xi);
  dim3 grid((unsigned)((mr + ZX_BM - 1) / ZX_BM), (unsigned)((R + ZX_BN - 1) / ZX_BN), 1);
  DM_REQUIRE(grid.x < (1u << 31) && grid.y <= 65535, DM_ERR_ARGUMENT, "complex GEMM too large for one launch");
  k_zx_mfma64<<<grid, 256, 0, s>>>(mr, n, lda, A, Xr, Xi, np, R, reinterpret_cast<double2*>(Y), ldy);
  HIP_OK(hipGetLastError());
}

// Y (n x R, ldy, complex128) = Z32 X (X: n x R, ldx, complex128).  ws: at
// least zx_workspace_bytes(n, R).
void zx_mfma(i64 n, i64 ld, const float* z32, const double* X, i64 ldx, int R, double* Y, i64 ldy, void* ws,
             hipStream_t s) {
  DM_REQUIRE((ld & 1) == 0, DM_ERR_ARGUMENT, "complex64 matrix leading dimension must be even");
  i64 nk = (n + ZX_KC - 1) / ZX_KC;
  double2* part = reinterpret_cast<double2*>(ws);
  DM_REQUIRE(ld % ZX_BM == 0 && ld >= n, DM_ERR_ARGUMENT, "complex64 matrix leading dimension must be a multiple of 64");
  const i64 np = (n + 31) / 32 * 32;  // plane rows, zero past n (16-byte loads never cross a column)
  float* Xr = reinterpret_cast<float*>(reinterpret_cast<uintptr_t>(part + nk * R * n + 1) & ~uintptr_t(15));
  float* Xi = Xr + np * R;
  i64 tot = n * R;
  k_split_planes<<<(unsigned)((np * R + 255) / 256), 256, 0, s>>>(n, np, R, reinterpret_cast<const double2*>(X), ldx,
                                                                   Xr, Xi);
  dim3 grid((unsigned)((n + ZX_BM - 1) / ZX_BM), (unsigned)((R + ZX_BN - 1) / ZX_BN), (unsigned)nk);
  DM_REQUIRE(grid.y <= 65535 && grid.z <= 65535, DM_ERR_ARGUMENT, "too many right-hand sides for one product");
  k_zx_mfma<<<grid, 256, 0, s>>>(n, ld, z32, Xr, Xi, np, R, part);
  k_zx_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, R, (int)nk, part, reinterpret_cast<double2*>(Y), ldy);
  HIP_OK(hipGetLastError());
}

}  // namespace dml
