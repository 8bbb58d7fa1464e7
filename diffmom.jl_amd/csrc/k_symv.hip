// k_symv.hip — y = Z x from packed symmetric storage (upper triangle of
// kPT x kPT tiles).  The EFIE matrix is exactly symmetric, so every stored
// off-diagonal tile Z_IJ serves two products, y_I += Z_IJ x_J and
// y_J += Z_IJ^T x_I, from one read: the product streams N^2 / 2 entries
// instead of N^2 (4 N^2 bytes for complex64).  Arithmetic is FP64 on
// widened entries.
//
// A workgroup walks a segment of up to `seg` consecutive tiles of one tile
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

constexpr int kSegMin = 2;  // shortest segment (tiles per workgroup) the workspace allows


__device__ inline double wsum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// number of segments of tile columns 0 .. J-1 (column J has J / seg + 1)
__host__ __device__ inline i64 seg_base(i64 J, int seg) {
  i64 q = J / seg, rem = J % seg;
  return J + seg * q * (q - 1) / 2 + q * rem;
}

// Chunks of CC columns x kPT rows are read with coalesced 16-byte loads into
// registers.  Each thread's loads cover fixed rows, so the row sums (Z_IJ x_J)
// are formed from the registers; the chunk is then staged in LDS, where
// threads owning a column form the column sums (Z_IJ^T x_I) without cross-lane
// reductions.  Two chunks are in flight per workgroup (registers of chunk
// g + 2 are reloaded as soon as chunk g is in LDS).
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
__global__ __launch_bounds__(256) void k_symv(i64 n, int nb, int seg, const void* __restrict__ zp,
                                              const double2* __restrict__ x, double2* __restrict__ rowpart,
                                              double2* __restrict__ colpart) {
  using C = SymvCfg<F32>;
  using ET = typename C::ET;
  // thread t's 16-byte load k covers column t / RPL + (256 / RPL) k of the
  // chunk and rows (t % RPL) * EPL .. + EPL - 1 (EPL complex per 16 bytes)
  constexpr int EPL = F32 ? 2 : 1;          // complex entries per 16-byte load
  constexpr int RPL = kPT / EPL;            // threads covering one column
  constexpr int RGRP = 256 / RPL;           // thread groups sharing the same rows
  constexpr i64 kChunkF4 = (i64)C::CC * kPT * sizeof(ET) / 16;
  constexpr i64 kTileF4 = (i64)kPT * kPT * sizeof(ET) / 16;
  __shared__ ET Zs[C::CC * C::LD];
  __shared__ double2 xs[2][kPT];
  __shared__ double2 red[RGRP][kPT];
  __shared__ double2 cacc[C::G * kPT];  // column sums of (row group, column), one owner thread each
  __shared__ int segJ, segI0;
  const int tid = threadIdx.x;
  if (tid == 0) {
    int b = blockIdx.x, J = 0;
    while (b >= J / seg + 1) {
      b -= J / seg + 1;
      ++J;
    }
    segJ = J;
    segI0 = b * seg;
  }
  __syncthreads();
  const int J = segJ, I0 = segI0;
  const int I1 = min(I0 + seg, J + 1);
  const i64 l0 = (i64)J * (J + 1) / 2 + I0;  // first tile of the segment (tiles are contiguous)
  const int nchunk = (I1 - I0) * C::NCH;     // even: NCH is
  for (int c = tid; c < kPT; c += 256) {
    i64 g = (i64)J * kPT + c;
    xs[0][c] = g < n ? x[g] : make_double2(0.0, 0.0);
  }
  for (int e = tid; e < C::G * kPT; e += 256) cacc[e] = make_double2(0.0, 0.0);
  const int row0 = (tid % RPL) * EPL, rgrp = tid / RPL;
  const int cc = tid % C::CC, gg = tid / C::CC;
  double ar[EPL], ai[EPL];
  // two chunks in flight: A holds even chunks, B odd ones
  f4v va[8], vb[8];
  const i64 base = l0 * kTileF4;
  chunk_load<F32>(zp, base, tid, va);
  chunk_load<F32>(zp, base + kChunkF4, tid, vb);
  auto step = [&](int g, f4v(&v)[8]) {
    const int ch = g % C::NCH;
    const int I = I0 + g / C::NCH;
    const i64 l = l0 + g / C::NCH;
    const int c0 = ch * C::CC;
    if (ch == 0) {
      __syncthreads();
      for (int r = tid; r < kPT; r += 256) {
        i64 gr = (i64)I * kPT + r;
        xs[1][r] = gr < n ? x[gr] : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        ar[e] = 0.0;
        ai[e] = 0.0;
      }
    }
    // row sums straight from the loaded registers
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const double2 xc = xs[0][c0 + rgrp + RGRP * k];
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        double zr, zi;
        if constexpr (F32 != 0) {
          zr = (double)v[k][2 * e];
          zi = (double)v[k][2 * e + 1];
        } else {
          double2 d;
          __builtin_memcpy(&d, &v[k], sizeof(d));
          zr = d.x;
          zi = d.y;
        }
        ar[e] += zr * xc.x - zi * xc.y;
        ai[e] += zr * xc.y + zi * xc.x;
      }
    }
    __syncthreads();
    chunk_store<F32>(Zs, tid, v);
    __syncthreads();
    if (g + 2 < nchunk) chunk_load<F32>(zp, base + (i64)(g + 2) * kChunkF4, tid, v);
    if (I != J) {
      double sr = 0, si = 0;
#pragma unroll 4
      for (int r = gg * C::RG; r < (gg + 1) * C::RG; ++r) {
        const double2 z = widen(Zs[cc * C::LD + r]);
        const double2 xr = xs[1][r];
        sr += z.x * xr.x - z.y * xr.y;
        si += z.x * xr.y + z.y * xr.x;
      }
      double2& acc = cacc[gg * kPT + c0 + cc];
      acc.x += sr;
      acc.y += si;
    }
    if (ch == C::NCH - 1) {
#pragma unroll
      for (int e = 0; e < EPL; ++e) red[rgrp][row0 + e] = make_double2(ar[e], ai[e]);
      __syncthreads();
      if (tid < kPT) {
        double sr = 0, si = 0;
#pragma unroll
        for (int q = 0; q < RGRP; ++q) {
          sr += red[q][tid].x;
          si += red[q][tid].y;
        }
        rowpart[l * kPT + tid] = make_double2(sr, si);
      }
    }
  };
#pragma unroll 1
  for (int g = 0; g < nchunk; g += 2) {
    step(g, va);
    step(g + 1, vb);
  }
  // column sums: reduce the G row groups of every column
  __syncthreads();
  if (tid < kPT) {
    double sr = 0, si = 0;
    for (int q = 0; q < C::G; ++q) {
      sr += cacc[q * kPT + tid].x;
      si += cacc[q * kPT + tid].y;
    }
    colpart[(i64)blockIdx.x * kPT + tid] = make_double2(sr, si);
  }
}

// y[R kPT + r] = sum_{J >= R} rowpart[tile (R, J)][r] + sum_{segments s of column R} colpart[s][r]
__global__ void k_symv_reduce(i64 n, int nb, int seg, const double2* __restrict__ rowpart,
                              const double2* __restrict__ colpart, double2* __restrict__ y) {
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
  const i64 sb = seg_base(R, seg);
  for (int s = 0; s <= R / seg; ++s) {
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
  return (size_t)(packed_tiles(n) + seg_base(nb, kSegMin)) * kPT * sizeof(double2);
}

void symv_packed(i64 n, const ZRef& z, const double* x, double* y, void* ws, int seg, hipStream_t s) {
  DM_REQUIRE(z.packed, DM_ERR_ARGUMENT, "symmetric product needs packed storage");
  DM_REQUIRE(seg >= kSegMin && seg <= 64, DM_ERR_ARGUMENT, "segment length out of range");
  const i64 nb = (n + kPT - 1) / kPT;
  const i64 nseg = seg_base(nb, seg);
  DM_REQUIRE(nseg < (1LL << 31), DM_ERR_ARGUMENT, "matrix too large for the packed product");
  double2* rowpart = reinterpret_cast<double2*>(ws);
  double2* colpart = rowpart + packed_tiles(n) * kPT;
  const double2* xv = reinterpret_cast<const double2*>(x);
  if (z.f32)
    k_symv<1><<<(unsigned)nseg, 256, 0, s>>>(n, (int)nb, seg, z.p, xv, rowpart, colpart);
  else
    k_symv<0><<<(unsigned)nseg, 256, 0, s>>>(n, (int)nb, seg, z.p, xv, rowpart, colpart);
  k_symv_reduce<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, (int)nb, seg, rowpart, colpart,
                                                              reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}

}  // namespace dml
