// k_symv.hip — y = Z x from packed symmetric storage (upper triangle of
// kPT x kPT tiles).  The EFIE matrix is exactly symmetric, so every stored
// off-diagonal tile Z_IJ serves two products, y_I += Z_IJ x_J and
// y_J += Z_IJ^T x_I, from one read: the product streams N^2 / 2 entries
// instead of N^2 (4 N^2 bytes for complex64).  Arithmetic is FP64 on
// widened entries.
//
// A workgroup walks a segment of up to kSeg consecutive tiles of one tile
// column J (contiguous in memory), staging each tile through LDS in chunks of
// columns.  Row sums Z_IJ x_J are stored per tile; column sums Z_IJ^T x_I are
// kept per thread over the whole segment and reduced through LDS once.  A
// second kernel adds, for every row block, its per-tile row sums and the
// column sums of the segments of its tile column, in a fixed order
// (deterministic).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "dm_common.h"
#include "k_linalg.h"

namespace {

constexpr int kSeg = 8;  // tiles per workgroup


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

// Chunks of CC columns x kPT rows are staged in LDS with coalesced 16-byte
// loads; rows are then summed by threads owning a row (Z_IJ x_J) and columns by
// threads owning a column (Z_IJ^T x_I), so neither product needs cross-lane
// reductions.  The next chunk's loads are issued into registers before the
// current chunk is consumed.
template <int F32>
struct SymvCfg {
  typedef typename std::conditional<F32 != 0, float2, double2>::type ET;
  static constexpr int CC = F32 ? 32 : 16;         // columns per chunk (32 KB)
  static constexpr int NCH = kPT / CC;             // chunks per tile
  static constexpr int LD = kPT + 1;               // padded LDS column
  static constexpr int G = 256 / CC;               // row groups of the column phase
  static constexpr int RG = kPT / G;               // rows per group
};

typedef float f4v __attribute__((ext_vector_type(4)));

template <int F32>
__device__ inline void chunk_load(const void* zp, i64 base, int tid, f4v (&v)[8]) {
  const f4v* src = reinterpret_cast<const f4v*>(zp) + base + tid;
#pragma unroll
  for (int k = 0; k < 8; ++k) v[k] = __builtin_nontemporal_load(src + 256 * k);
}

template <int F32>
__device__ inline void chunk_store(typename SymvCfg<F32>::ET* Zs, int tid, const f4v (&v)[8]) {
  using C = SymvCfg<F32>;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int e = tid + 256 * k;
    if constexpr (F32 != 0) {
      const int c = e / (kPT / 2), rp = e % (kPT / 2);
      Zs[c * C::LD + 2 * rp] = make_float2(v[k][0], v[k][1]);
      Zs[c * C::LD + 2 * rp + 1] = make_float2(v[k][2], v[k][3]);
    } else {
      const int c = e / kPT, r = e % kPT;
      double2 d;
      __builtin_memcpy(&d, &v[k], sizeof(d));
      Zs[c * C::LD + r] = d;
    }
  }
}

template <typename ET>
__device__ inline double2 widen(ET v) {
  return make_double2((double)v.x, (double)v.y);
}

template <int F32>
__global__ __launch_bounds__(256) void k_symv(i64 n, int nb, const void* __restrict__ zp,
                                              const double2* __restrict__ x, double2* __restrict__ rowpart,
                                              double2* __restrict__ colpart) {
  using C = SymvCfg<F32>;
  using ET = typename C::ET;
  __shared__ ET Zs[C::CC * C::LD];
  __shared__ double2 xs[2][kPT];
  __shared__ double2 red[2][kPT];
  __shared__ int segJ, segI0;
  const int tid = threadIdx.x;
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
  }
  // row phase: thread owns row rr of half hh of each chunk's columns
  const int rr = tid & (kPT - 1), hh = tid >> 7;
  // column phase: thread owns column cc of each chunk, rows of group gg
  const int cc = tid % C::CC, gg = tid / C::CC;
  double colr[C::NCH], coli[C::NCH];
#pragma unroll
  for (int k = 0; k < C::NCH; ++k) {
    colr[k] = 0.0;
    coli[k] = 0.0;
  }
  // 16-byte units per chunk: CC * kPT complex = 4096 (complex64) / 2048 (complex128) -> 2048 float4
  constexpr i64 kChunkF4 = (i64)C::CC * kPT * sizeof(ET) / 16;
  f4v v[8];
  chunk_load<F32>(zp, ((i64)J * (J + 1) / 2 + I0) * kPT * kPT * sizeof(ET) / 16, tid, v);
  for (int I = I0; I < I1; ++I) {
    const i64 l = (i64)J * (J + 1) / 2 + I;
    const bool offdiag = I != J;
    __syncthreads();
    for (int r = tid; r < kPT; r += 256) {
      i64 g = (i64)I * kPT + r;
      xs[1][r] = g < n ? x[g] : make_double2(0.0, 0.0);
    }
    double ar = 0, ai = 0;
    for (int ch = 0; ch < C::NCH; ++ch) {
      __syncthreads();
      chunk_store<F32>(Zs, tid, v);
      __syncthreads();
      // prefetch the next chunk (of this tile or the next one in the segment)
      const bool more = ch + 1 < C::NCH || I + 1 < I1;
      if (more) {
        const i64 nl = ch + 1 < C::NCH ? l : l + 1;
        const int nch = ch + 1 < C::NCH ? ch + 1 : 0;
        chunk_load<F32>(zp, nl * kPT * kPT * sizeof(ET) / 16 + nch * kChunkF4, tid, v);
      }
      const int c0 = ch * C::CC;
#pragma unroll 4
      for (int c = hh * (C::CC / 2); c < (hh + 1) * (C::CC / 2); ++c) {
        const double2 z = widen(Zs[c * C::LD + rr]);
        const double2 xc = xs[0][c0 + c];
        ar += z.x * xc.x - z.y * xc.y;
        ai += z.x * xc.y + z.y * xc.x;
      }
      if (offdiag) {
        double sr = 0, si = 0;
#pragma unroll 4
        for (int r = gg * C::RG; r < (gg + 1) * C::RG; ++r) {
          const double2 z = widen(Zs[cc * C::LD + r]);
          const double2 xr = xs[1][r];
          sr += z.x * xr.x - z.y * xr.y;
          si += z.x * xr.y + z.y * xr.x;
        }
        colr[ch] += sr;
        coli[ch] += si;
      }
    }
    red[hh][rr] = make_double2(ar, ai);
    __syncthreads();
    if (tid < kPT) {
      rowpart[l * kPT + tid] = make_double2(red[0][tid].x + red[1][tid].x, red[0][tid].y + red[1][tid].y);
    }
  }
  // column sums: reduce the G row groups of every column through LDS
  __syncthreads();
  double2* cred = reinterpret_cast<double2*>(Zs);  // G x kPT double2 fits in the chunk buffer
#pragma unroll
  for (int k = 0; k < C::NCH; ++k) cred[gg * kPT + k * C::CC + cc] = make_double2(colr[k], coli[k]);
  __syncthreads();
  if (tid < kPT) {
    double sr = 0, si = 0;
    for (int q = 0; q < C::G; ++q) {
      sr += cred[q * kPT + tid].x;
      si += cred[q * kPT + tid].y;
    }
    colpart[(i64)blockIdx.x * kPT + tid] = make_double2(sr, si);
  }
}

// y[R kPT + r] = sum_{J >= R} rowpart[tile (R, J)][r] + sum_{segments s of column R} colpart[s][r]
__global__ void k_symv_reduce(i64 n, int nb, const double2* __restrict__ rowpart, const double2* __restrict__ colpart,
                              double2* __restrict__ y) {
  const i64 g = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int R = (int)(g / kPT), r = (int)(g % kPT);
  double sr = 0, si = 0;
  // eight independent loads in flight per step (the sum order stays fixed)
  int J = R;
  for (; J + 8 <= nb; J += 8) {
    double2 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = rowpart[((i64)(J + u) * (J + u + 1) / 2 + R) * kPT + r];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      sr += v[u].x;
      si += v[u].y;
    }
  }
  for (; J < nb; ++J) {
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
