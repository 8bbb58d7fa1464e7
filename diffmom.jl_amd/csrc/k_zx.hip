// k_zx.hip — multi-right-hand-side products Y = Z X on the matrix cores.
//
// The complex64 matrix is column-major with interleaved (re, im) parts, i.e.
// it already is the real 2N x N matrix A' with rows (Re z_m, Im z_m).  With
// P = A' Re(X) and Q = A' Im(X) (two real products sharing every A' tile):
//   Re Y_m = P_{2m} - Q_{2m+1},   Im Y_m = Q_{2m} + P_{2m+1}
// so the complex product costs 8 N^2 R useful flops.  k_zx_bf16 computes it
// on the bf16 matrix cores at FP32 accuracy (three-part split of every FP32
// value, six part products); K chunks of KC columns accumulate in FP32 and
// are summed in FP64 by k_zx_reduce, which bounds the FP32 accumulation
// length.  k_zx_mfma64 is the complex128 product on the FP64 matrix cores
// (refinement residuals of many right-hand sides, far-field GEMMs).
#include <hip/hip_runtime.h>

#include "dm_common.h"
#include "k_linalg.h"

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int ZX_BM = 64;   // complex rows per workgroup
constexpr int ZX_BN = 64;   // right-hand sides per workgroup
constexpr int ZX_BK = 16;   // complex columns of Z per LDS stage
constexpr int ZX_KC = 4096; // columns per FP32 accumulation chunk

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

typedef double v4d __attribute__((ext_vector_type(4)));

// FP64 variant (complex128 Z, v_mfma_f64_16x16x4f64): the refinement residual
// of many right-hand sides.  Same tiling; one K chunk (FP64 accumulation).
// Rows mr (complex) and inner dimension n may differ: Y (mr x R) = Z (mr x n,
// leading dimension ld) X (n x R).
__global__ __launch_bounds__(256, 2) void k_zx_mfma64(i64 mr, i64 n, i64 ld, const double* __restrict__ Z,
                                                   const double* __restrict__ Xr, const double* __restrict__ Xi, i64 np,
                                                   int R, double2* __restrict__ Y, i64 ldy, int sub) {
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
        const double2 v = make_double2(p[r] - qo[r], q[r] + po[r]);
        if (sub) {  // Y -= Z X
          const double2 o = Y[m + (i64)col * ldy];
          Y[m + (i64)col * ldy] = make_double2(o.x - v.x, o.y - v.y);
        } else {
          Y[m + (i64)col * ldy] = v;
        }
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


// ------------------------------------------------------------------ bf16 x 3
// The complex64 product on the bf16 matrix cores at FP32 accuracy: every FP32
// value x is split exactly into three bf16 parts, x = h + m + l (h = bf16(x),
// m = bf16(x - h), l = x - h - m; each difference is exact in FP32 and l
// fits bf16), and A X is formed from the six part products h h', h m', m h',
// h l', l h', m m' (the dropped ones are below 2^-24 relative), accumulated
// in FP32 by v_mfma_f32_16x16x32_bf16 (lane l supplies row l % 16 and eight
// consecutive k from 8 (l / 16); the output layout is that of the FP32
// 16x16 product).  Planes of A' are stored in groups of 8 columns, (k, r) at
// ((k / 8) R2 + r) 8 + k % 8, so an operand is one 16-byte read; the X planes
// are column-major with k contiguous.  A workgroup covers 64 complex rows x
// 64 right-hand sides for one K chunk (four waves of 32 real rows each); the
// four column blocks of a row block are dispatched back to back on one XCD,
// which share its A' stages in L2.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
constexpr int ZB_LDA = 40;  // bf16 per LDS row (32 + 8 padding)

__device__ inline void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = (__bf16)x;
  const float r1 = x - (float)h;
  m = (__bf16)r1;
  l = (__bf16)(r1 - (float)m);
}

// A' (FP32, 2 ld x n column-major) -> three bf16 planes in 8-column groups,
// ngroups groups (columns past n zero).  One thread per (row, group).
__global__ void k_split_a_bf16(i64 R2, i64 n, i64 ngroups, const float* __restrict__ z, __bf16* __restrict__ out,
                               i64 plane) {
  const i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= R2 * ngroups) return;
  const i64 r = idx % R2, g = idx / R2;
  v8bf h, m, l;
#pragma unroll
  for (int kk = 0; kk < 8; ++kk) {
    const i64 k = 8 * g + kk;
    const float x = k < n ? z[k * R2 + r] : 0.f;
    __bf16 a, b, c;
    split3(x, a, b, c);
    h[kk] = a;
    m[kk] = b;
    l[kk] = c;
  }
  const i64 at = (g * R2 + r) * 8;
  *reinterpret_cast<v8bf*>(out + at) = h;
  *reinterpret_cast<v8bf*>(out + plane + at) = m;
  *reinterpret_cast<v8bf*>(out + 2 * plane + at) = l;
}

// complex128 X (n x R, ldx) -> six bf16 planes np x Rp (Re h, m, l, Im h, m,
// l), rows past n and columns past R zero.
__global__ void k_split_x_bf16(i64 n, i64 np, int R, int Rp, const double2* __restrict__ X, i64 ldx,
                               __bf16* __restrict__ out) {
  const i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  const i64 plane = np * Rp;
  if (idx >= plane) return;
  const i64 k = idx % np, c = idx / np;
  const double2 v = (k < n && c < R) ? X[k + c * ldx] : make_double2(0.0, 0.0);
  __bf16 a, b, d;
  split3((float)v.x, a, b, d);
  out[idx] = a;
  out[plane + idx] = b;
  out[2 * plane + idx] = d;
  split3((float)v.y, a, b, d);
  out[3 * plane + idx] = a;
  out[4 * plane + idx] = b;
  out[5 * plane + idx] = d;
}

// NT = 6: the six part products (FP32 accuracy); NT = 3: h h', h m', m h'
// only (the dropped terms are each below 2^-18 relative, ~1e-5 in all), for
// relaxed products late in a Krylov solve.
template <int NT>
__global__ __launch_bounds__(256, 1) void k_zx_bf16(i64 n, i64 R2, int nrb, int ncb, const __bf16* __restrict__ Ap,
                                                    i64 aplane,
                                                    const __bf16* __restrict__ Xp, i64 np, i64 xplane, int R,
                                                    i64 kc, double2* __restrict__ part) {
  __shared__ __align__(16) __bf16 Bs[2][6][ZX_BN][ZB_LDA];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // XCD-aware tile order: workgroup w runs on XCD w % 8; its index j = w / 8
  // on that XCD gives row block 8 (j / ncb) + w % 8 and column block j % ncb
  const int w = blockIdx.x, j = w / 8;
  const int rb_idx = 8 * (j / ncb) + w % 8, cb_idx = j % ncb;
  if (rb_idx >= nrb) return;
  const i64 m0 = (i64)rb_idx * ZX_BM;  // complex row base
  const int c0 = cb_idx * ZX_BN;       // rhs base
  const i64 k_begin = (i64)blockIdx.z * kc;
  const i64 k_end = k_begin + kc < n ? k_begin + kc : n;
  const int nst = (int)((k_end - k_begin + 31) / 32);
  typedef unsigned int u4v __attribute__((ext_vector_type(4)));
  // A' rows belong to one wave each (32 real rows per wave): they go from
  // global memory straight into the MFMA operand layout (lane l: row l % 16,
  // columns 8 (l / 16) .. + 7 of the stage, one 16-byte load from the
  // 8-column-group planes).  X is shared by the four waves and staged in LDS
  // (two buffers, one barrier per stage): 6 planes x 64 columns x 4 groups of
  // 8 rows = 1536 16-byte pieces, 6 per thread.
  const int g = lane >> 4;
  const i64 arow = 2 * m0 + 32 * wave + (lane & 15);
  v8bf ra[3][2];
  u4v rb[6];
  auto gload = [&](i64 k0) {
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int r = 0; r < 2; ++r)
        ra[q][r] = *reinterpret_cast<const v8bf*>(Ap + q * aplane + ((k0 / 8 + g) * R2 + arow + 16 * r) * 8);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int c = tid + 256 * i, q = c >> 8, rem = c & 255, col = rem >> 2, kq = rem & 3;
      rb[i] = *reinterpret_cast<const u4v*>(Xp + q * xplane + (i64)(c0 + col) * np + k0 + 8 * kq);
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const int c = tid + 256 * i, q = c >> 8, rem = c & 255, col = rem >> 2, kq = rem & 3;
      *reinterpret_cast<u4v*>(&Bs[buf][q][col][8 * kq]) = rb[i];
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
    const int buf = st & 1;
    // all operands of the stage, then the six part products, each over the
    // wave's 16 independent accumulator tiles (no back-to-back dependence)
    v8bf A[3][2], Br[3][4], Bi[3][4];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
      for (int r = 0; r < 2; ++r) A[q][r] = ra[q][r];
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        Br[q][cb] = *reinterpret_cast<const v8bf*>(&Bs[buf][q][16 * cb + (lane & 15)][8 * g]);
        Bi[q][cb] = *reinterpret_cast<const v8bf*>(&Bs[buf][3 + q][16 * cb + (lane & 15)][8 * g]);
      }
    }
    if (st + 1 < nst) gload(k_begin + (i64)(st + 1) * 32);
    constexpr int kPa[6] = {1, 0, 2, 0, 1, 0}, kPb[6] = {1, 2, 0, 1, 0, 0};  // smallest terms first
#pragma unroll
    for (int t = 6 - NT; t < 6; ++t)
#pragma unroll
      for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          accP[r][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[kPa[t]][r], Br[kPb[t]][cb], accP[r][cb], 0, 0, 0);
          accQ[r][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[kPa[t]][r], Bi[kPb[t]][cb], accQ[r][cb], 0, 0, 0);
        }
    // the other buffer was last read in stage st - 1, before the barrier that
    // ended it
    if (st + 1 < nst) {
      lstore(buf ^ 1);
      __syncthreads();
    }
  }
  // output layout of the 16x16 FP32-accumulator MFMA: lane holds real rows
  // 4 (lane / 16) + 0..3 of each 16-row block, column lane % 16
  double2* out = part + (i64)blockIdx.z * (i64)R * n;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
      const int col = c0 + 16 * cb + (lane & 15);
      if (col >= R) continue;
      v4f p = accP[r][cb], q = accQ[r][cb];
      const i64 real_row = 32 * wave + 16 * r + 4 * (lane >> 4);
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

}  // namespace

namespace dml {

size_t zx_workspace_bytes(i64 n, int R) {
  const size_t a = zx_bf16_workspace_bytes(n, R);
  const size_t b = 2 * (size_t)((n + 15) / 16 * 16) * R * sizeof(double);
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
  k_zx_mfma64<<<grid, 256, 0, s>>>(n, n, ld, z64, Xr, Xi, np, R, reinterpret_cast<double2*>(Y), ldy, 0);
  HIP_OK(hipGetLastError());
}

void gemm_c128(i64 mr, i64 n, const double* A, i64 lda, const double* X, i64 ldx, int R, double* Y, i64 ldy,
               void* ws, hipStream_t s, bool subtract) {
  DM_REQUIRE(lda >= mr && ldx >= n && ldy >= mr && R >= 1, DM_ERR_ARGUMENT, "complex GEMM: bad dimensions");
  const i64 np = (n + 15) / 16 * 16;
  double* Xr = reinterpret_cast<double*>(ws);
  double* Xi = Xr + np * R;
  k_split_planes64<<<(unsigned)((np * R + 255) / 256), 256, 0, s>>>(n, np, R, reinterpret_cast<const double2*>(X), ldx,
                                                                     Xr, Xi);
  dim3 grid((unsigned)((mr + ZX_BM - 1) / ZX_BM), (unsigned)((R + ZX_BN - 1) / ZX_BN), 1);
  DM_REQUIRE(grid.x < (1u << 31) && grid.y <= 65535, DM_ERR_ARGUMENT, "complex GEMM too large for one launch");
  k_zx_mfma64<<<grid, 256, 0, s>>>(mr, n, lda, A, Xr, Xi, np, R, reinterpret_cast<double2*>(Y), ldy,
                                   subtract ? 1 : 0);
  HIP_OK(hipGetLastError());
}

// bf16 x 3 form of Z32 (see k_zx_bf16): planes of A' built once per matrix.
i64 zx_bf16_groups(i64 n) { return (n + 31) / 32 * 4; }
size_t zx_bf16_planes_elems(i64 ld, i64 n) { return (size_t)3 * zx_bf16_groups(n) * 2 * ld * 8; }
void zx_bf16_split(i64 n, i64 ld, const float* z32, void* planes, hipStream_t s) {
  const i64 R2 = 2 * ld, ng = zx_bf16_groups(n);
  const i64 tot = R2 * ng;
  k_split_a_bf16<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(R2, n, ng, z32, reinterpret_cast<__bf16*>(planes),
                                                               ng * R2 * 8);
  HIP_OK(hipGetLastError());
}
size_t zx_bf16_workspace_bytes(i64 n, int R) {
  const i64 nk = (n + ZX_KC - 1) / ZX_KC, np = (n + 31) / 32 * 32, Rp = ((i64)R + ZX_BN - 1) / ZX_BN * ZX_BN;
  return (size_t)nk * R * n * sizeof(double2) + 64 + (size_t)6 * np * Rp * 2;
}
std::vector<int> zx_bf16_chunks() { return {ZX_KC, 2 * ZX_KC}; }
void zx_bf16(i64 n, i64 ld, const void* planes, const double* X, i64 ldx, int R, double* Y, i64 ldy, void* ws,
             hipStream_t s, int kc, int nterms) {
  DM_REQUIRE(nterms == 6 || nterms == 3, DM_ERR_ARGUMENT, "bf16 part products must be 6 or 3");
  DM_REQUIRE(ld % ZX_BM == 0 && ld >= n, DM_ERR_ARGUMENT, "complex64 matrix leading dimension must be a multiple of 64");
  if (kc <= 0) kc = ZX_KC;
  // the workspace holds the partials of ZX_KC-column chunks; longer chunks need fewer
  DM_REQUIRE(kc % 32 == 0 && kc >= ZX_KC, DM_ERR_ARGUMENT, "Z X column chunk must be a multiple of 32, >= 4096");
  const i64 nk = (n + kc - 1) / kc, np = (n + 31) / 32 * 32;
  const int Rp = (R + ZX_BN - 1) / ZX_BN * ZX_BN;
  double2* part = reinterpret_cast<double2*>(ws);
  __bf16* Xp = reinterpret_cast<__bf16*>(reinterpret_cast<char*>(ws) +
                                         (((size_t)nk * R * n * sizeof(double2) + 63) & ~size_t(63)));
  const i64 xplane = np * Rp;
  k_split_x_bf16<<<(unsigned)((xplane + 255) / 256), 256, 0, s>>>(n, np, R, Rp, reinterpret_cast<const double2*>(X),
                                                                   ldx, Xp);
  const int nrb = (int)(ld / ZX_BM), ncb = Rp / ZX_BN;
  const i64 nwg = (i64)((nrb + 7) / 8) * 8 * ncb;
  DM_REQUIRE(nwg < (1LL << 31) && nk <= 65535 && ncb <= 65535, DM_ERR_ARGUMENT, "product too large for one launch");
  const i64 R2 = 2 * ld, ng = zx_bf16_groups(n);
  dim3 grid((unsigned)nwg, 1, (unsigned)nk);
  if (nterms == 6)
    k_zx_bf16<6><<<grid, 256, 0, s>>>(n, R2, nrb, ncb, reinterpret_cast<const __bf16*>(planes), ng * R2 * 8, Xp, np,
                                      xplane, R, (i64)kc, part);
  else
    k_zx_bf16<3><<<grid, 256, 0, s>>>(n, R2, nrb, ncb, reinterpret_cast<const __bf16*>(planes), ng * R2 * 8, Xp, np,
                                      xplane, R, (i64)kc, part);
  const i64 tot = n * (i64)R;
  k_zx_reduce<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(n, R, (int)nk, part, reinterpret_cast<double2*>(Y), ldy);
  HIP_OK(hipGetLastError());
}

}  // namespace dml
