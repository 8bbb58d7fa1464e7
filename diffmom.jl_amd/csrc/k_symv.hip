// k_symv.hip — y = Z x from packed symmetric storage (upper triangle of
// kPT x kPT tiles).  The EFIE matrix is exactly symmetric, so every stored
// off-diagonal tile Z_IJ serves two products, y_I += Z_IJ x_J and
// y_J += Z_IJ^T x_I, from one read: the product streams N^2 / 2 entries
// instead of N^2 (4 N^2 bytes for complex64).  Arithmetic is FP64 on
// widened entries.
//
// A workgroup walks a segment of up to kSeg consecutive tiles of one tile
// column J (contiguous in memory).  Wave w reads columns w, w + 4, ... of each
// tile, lane l the rows (2l, 2l+1) (complex64: one 16-byte load) or (l, l+64)
// (complex128: two 16-byte loads).  Row sums Z_IJ x_J are reduced across the
// four waves and stored per tile; column sums Z_IJ^T x_I are reduced across
// the wave and accumulated over the segment in LDS.  A second kernel adds, for
// every row block, its per-tile row sums and the column sums of the segments
// of its tile column, in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include "dm_common.h"
#include "k_linalg.h"

namespace {

constexpr int kSeg = 8;  // tiles per workgroup

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ inline double wsum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// number of segments of tile columns 0 .. J-1
__host__ __device__ inline i64 seg_base(i64 J) {
  i64 q = J / kSeg, rem = J % kSeg;
  return J + kSeg * q * (q - 1) / 2 + q * rem;
}

template <int F32>
__device__ inline void load_pair(const void* zp, i64 colbase, int lane, double2& za, double2& zb) {
  if constexpr (F32 != 0) {
    f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(zp) + colbase / 2 + lane);
    za = make_double2((double)v[0], (double)v[1]);
    zb = make_double2((double)v[2], (double)v[3]);
  } else {
    const double2* t = reinterpret_cast<const double2*>(zp) + colbase;
    za = t[lane];
    zb = t[lane + 64];
  }
}

template <int F32>
__global__ __launch_bounds__(256) void k_symv(i64 n, int nb, const void* __restrict__ zp,
                                              const double2* __restrict__ x, double2* __restrict__ rowpart,
                                              double2* __restrict__ colpart) {
  __shared__ double2 xs[2][kPT];
  __shared__ double2 colacc[kPT];
  __shared__ double2 rowred[4][kPT];
  __shared__ int segJ, segI0;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  if (tid == 0) {
    int b = blockIdx.x, J = 0;
    while (b >= J / kSeg + 1) {
      b -= J / kSeg + 1;
      ++J;
    }
    segJ = J;
    segI0 = b * kSeg;
  }
  __syncthreads();
  const int J = segJ, I0 = segI0;
  const int I1 = min(I0 + kSeg, J + 1);
  for (int c = tid; c < kPT; c += 256) {
    i64 g = (i64)J * kPT + c;
    xs[0][c] = g < n ? x[g] : make_double2(0.0, 0.0);
    colacc[c] = make_double2(0.0, 0.0);
  }
  const int ra = F32 ? 2 * lane : lane;
  const int rb = F32 ? 2 * lane + 1 : lane + 64;
  for (int I = I0; I < I1; ++I) {
    __syncthreads();
    for (int r = tid; r < kPT; r += 256) {
      i64 g = (i64)I * kPT + r;
      xs[1][r] = g < n ? x[g] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const double2 xa = xs[1][ra], xb = xs[1][rb];
    const i64 l = (i64)J * (J + 1) / 2 + I;
    const bool offdiag = I != J;
    double ar = 0, ai = 0, br = 0, bi = 0;
#pragma unroll 8
    for (int q = 0; q < kPT / 4; ++q) {
      const int c = w + 4 * q;
      double2 za, zb;
      load_pair<F32>(zp, (l * kPT + c) * kPT, lane, za, zb);
      const double2 xc = xs[0][c];
      ar += za.x * xc.x - za.y * xc.y;
      ai += za.x * xc.y + za.y * xc.x;
      br += zb.x * xc.x - zb.y * xc.y;
      bi += zb.x * xc.y + zb.y * xc.x;
      if (offdiag) {
        double tr = za.x * xa.x - za.y * xa.y + zb.x * xb.x - zb.y * xb.y;
        double ti = za.x * xa.y + za.y * xa.x + zb.x * xb.y + zb.y * xb.x;
        tr = wsum64(tr);
        ti = wsum64(ti);
        if (lane == 0) {
          colacc[c].x += tr;
          colacc[c].y += ti;
        }
      }
    }
    rowred[w][ra] = make_double2(ar, ai);
    rowred[w][rb] = make_double2(br, bi);
    __syncthreads();
    for (int r = tid; r < kPT; r += 256) {
      double sr = rowred[0][r].x + rowred[1][r].x + rowred[2][r].x + rowred[3][r].x;
      double si = rowred[0][r].y + rowred[1][r].y + rowred[2][r].y + rowred[3][r].y;
      rowpart[l * kPT + r] = make_double2(sr, si);
    }
  }
  __syncthreads();
  for (int c = tid; c < kPT; c += 256) colpart[(i64)blockIdx.x * kPT + c] = colacc[c];
}

// y[R kPT + r] = sum_{J >= R} rowpart[tile (R, J)][r] + sum_{segments s of column R} colpart[s][r]
__global__ void k_symv_reduce(i64 n, int nb, const double2* __restrict__ rowpart, const double2* __restrict__ colpart,
                              double2* __restrict__ y) {
  const i64 g = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int R = (int)(g / kPT), r = (int)(g % kPT);
  double sr = 0, si = 0;
  for (int J = R; J < nb; ++J) {
    double2 v = rowpart[((i64)J * (J + 1) / 2 + R) * kPT + r];
    sr += v.x;
    si += v.y;
  }
  const i64 sb = seg_base(R);
  for (int s = 0; s <= R / kSeg; ++s) {
    double2 v = colpart[(sb + s) * kPT + r];
    sr += v.x;
    si += v.y;
  }
  y[g] = make_double2(sr, si);
}

}  // namespace

namespace dml {

size_t symv_workspace_bytes(i64 n) {
  const i64 nb = (n + kPT - 1) / kPT;
  return (size_t)(packed_tiles(n) + seg_base(nb)) * kPT * sizeof(double2);
}

void symv_packed(i64 n, const ZRef& z, const double* x, double* y, void* ws, hipStream_t s) {
  DM_REQUIRE(z.packed, DM_ERR_ARGUMENT, "symmetric product needs packed storage");
  const i64 nb = (n + kPT - 1) / kPT;
  const i64 nseg = seg_base(nb);
  DM_REQUIRE(nseg < (1LL << 31), DM_ERR_ARGUMENT, "matrix too large for the packed product");
  double2* rowpart = reinterpret_cast<double2*>(ws);
  double2* colpart = rowpart + packed_tiles(n) * kPT;
  const double2* xv = reinterpret_cast<const double2*>(x);
  if (z.f32)
    k_symv<1><<<(unsigned)nseg, 256, 0, s>>>(n, (int)nb, z.p, xv, rowpart, colpart);
  else
    k_symv<0><<<(unsigned)nseg, 256, 0, s>>>(n, (int)nb, z.p, xv, rowpart, colpart);
  k_symv_reduce<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, (int)nb, rowpart, colpart, reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}

}  // namespace dml
