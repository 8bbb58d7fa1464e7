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

#include <algorithm>
#include <type_traits>
#include <vector>

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

// ---------------------------------------------------------------- wave form
// Every wave (64 lanes) takes `tpw` consecutive stored tiles in memory order
// (crossing tile columns), so all waves carry the same number of bytes.  A
// tile is read column by column, one coalesced 1 KB (complex64) or 2 KB
// (complex128) column per step; lane l holds rows rows_e(l) of the column:
//   row sums    y_I += Z_IJ x_J : per lane, in registers, x_J[c] wave-uniform;
//   column sums y_J += Z_IJ^T x_I : per lane a partial over its two rows, then
//               a transposing butterfly over the 64 lanes for 8 columns at a
//               time (xor 32/16/8 halve the values each step, xor 4/2/1 finish
//               the last one), which leaves column 8 b + (l >> 3) of batch b
//               in lanes 8 (l >> 3) .. + 7; lane l keeps batches l & 7 and
//               (l & 7) + 8.
// Row sums are stored per tile.  Column sums are accumulated over the wave's
// run of tiles in one tile column and stored at the run's last tile (the end
// of the wave's range, or the diagonal tile, whose own column sums are not
// added: diagonal tiles are stored in full).
constexpr int kWaveShape = 1000;  // shape >= this: wave form with shape - 1000 tiles per wave

template <int F32>
__device__ inline int wrow(int lane, int e) {
  return F32 ? 2 * lane + e : lane + 64 * e;
}

__device__ inline void cadd(double2& a, double2 b) {
  a.x += b.x;
  a.y += b.y;
}

__device__ inline double2 shfl2(double2 v, int o) {
  return make_double2(__shfl_xor(v.x, o, 64), __shfl_xor(v.y, o, 64));
}

// Z column c of the tile (2 entries per lane, widened to FP64).
// Transposing butterfly over the 64 lanes for 8 per-lane partial sums: after
// xor 32 / 16 / 8 (each halving the values kept) and xor 4 / 2 / 1, p[0] of
// lane l holds the sum over all lanes of value (l >> 3).
__device__ inline void butterfly8(double2 (&p)[8], int lane) {
  const bool h5 = lane & 32, h4 = lane & 16, h3 = lane & 8;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double2 keep = h5 ? p[i + 4] : p[i], send = h5 ? p[i] : p[i + 4];
    cadd(keep, shfl2(send, 32));
    p[i] = keep;
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double2 keep = h4 ? p[i + 2] : p[i], send = h4 ? p[i] : p[i + 2];
    cadd(keep, shfl2(send, 16));
    p[i] = keep;
  }
  {
    double2 keep = h3 ? p[1] : p[0], send = h3 ? p[0] : p[1];
    cadd(keep, shfl2(send, 8));
    p[0] = keep;
  }
  cadd(p[0], shfl2(p[0], 4));
  cadd(p[0], shfl2(p[0], 2));
  cadd(p[0], shfl2(p[0], 1));
}

// The same for 4 values: p[0] of lane l holds the sum of value (l >> 4).
__device__ inline void butterfly4(double2 (&p)[4], int lane) {
  const bool h5 = lane & 32, h4 = lane & 16;
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    double2 keep = h5 ? p[i + 2] : p[i], send = h5 ? p[i] : p[i + 2];
    cadd(keep, shfl2(send, 32));
    p[i] = keep;
  }
  {
    double2 keep = h4 ? p[1] : p[0], send = h4 ? p[0] : p[1];
    cadd(keep, shfl2(send, 16));
    p[0] = keep;
  }
  cadd(p[0], shfl2(p[0], 8));
  cadd(p[0], shfl2(p[0], 4));
  cadd(p[0], shfl2(p[0], 2));
  cadd(p[0], shfl2(p[0], 1));
}

template <int F32>
__device__ inline void col_load(const f4v* tile, int c, int lane, f4v (&v)[2]) {
  if constexpr (F32 != 0) {
    v[0] = __builtin_nontemporal_load(tile + c * (kPT / 2) + lane);
  } else {
    v[0] = __builtin_nontemporal_load(tile + c * kPT + lane);
    v[1] = __builtin_nontemporal_load(tile + c * kPT + 64 + lane);
  }
}

template <int F32>
__device__ inline void col_entries(const f4v (&v)[2], double2 (&z)[2]) {
  if constexpr (F32 != 0) {
    z[0] = make_double2((double)v[0][0], (double)v[0][1]);
    z[1] = make_double2((double)v[0][2], (double)v[0][3]);
  } else {
    __builtin_memcpy(&z[0], &v[0], sizeof(double2));
    __builtin_memcpy(&z[1], &v[1], sizeof(double2));
  }
}

// x_J[col] from the lane that holds it (wave-uniform col)
__device__ inline double rl(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

template <int F32>
__device__ inline double2 bcast(const double2 (&xJ)[2], int col) {
  // row col is entry e of lane l with wrow(l, e) = col
  const int e = F32 ? (col & 1) : (col >> 6);
  const int l = F32 ? (col >> 1) : (col & 63);
  const double2 v = e ? xJ[1] : xJ[0];
  return make_double2(rl(v.x, l), rl(v.y, l));
}

// One dense stored tile: accR (this lane's rows) += Z x_J; for an
// off-diagonal tile also the column partials Z^T x_I, reduced by the
// butterfly into accC0 / accC1 (column 8 b + (lane >> 3) of batch b = lane & 7
// and (lane & 7) + 8).
template <int F32>
__device__ inline void dense_tile(const f4v* tile, int lane, bool offdiag, const double2 (&xI)[2],
                                  const double2 (&xJ)[2], double2 (&accR)[2], double2& accC0, double2& accC1) {
  auto batch = [&](int b, f4v(&v)[8][2]) {
    double2 p[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      double2 z[2];
      col_entries<F32>(v[c], z);
      const int col = 8 * b + c;
      const double2 xc = bcast<F32>(xJ, col);
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        accR[e].x += z[e].x * xc.x - z[e].y * xc.y;
        accR[e].y += z[e].x * xc.y + z[e].y * xc.x;
      }
      p[c].x = z[0].x * xI[0].x - z[0].y * xI[0].y + z[1].x * xI[1].x - z[1].y * xI[1].y;
      p[c].y = z[0].x * xI[0].y + z[0].y * xI[0].x + z[1].x * xI[1].y + z[1].y * xI[1].x;
    }
    if (offdiag) {
      butterfly8(p, lane);
      if ((b & 7) == (lane & 7)) {
        if (b < 8)
          cadd(accC0, p[0]);
        else
          cadd(accC1, p[0]);
      }
    }
  };
#pragma unroll 1
  for (int b = 0; b < kPT / 8; ++b) {
    f4v v[8][2];
#pragma unroll
    for (int c = 0; c < 8; ++c) col_load<F32>(tile, 8 * b + c, lane, v[c]);
    batch(b, v);
  }
}

// tile t0 = J (J + 1) / 2 + I
__device__ inline void tile_ij(i64 t0, int& I, int& J) {
  J = (int)((sqrt(8.0 * (double)t0 + 1.0) - 1.0) * 0.5);
  while ((i64)J * (J + 1) / 2 > t0) --J;
  while ((i64)(J + 1) * (J + 2) / 2 <= t0) ++J;
  I = (int)(t0 - (i64)J * (J + 1) / 2);
}

template <int F32>
__global__ __launch_bounds__(256) void k_symv_w(i64 n, int nb, int tpw, const void* __restrict__ zp,
                                                const double2* __restrict__ x, double2* __restrict__ rowpart,
                                                double2* __restrict__ colpart) {
  using ET = typename SymvCfg<F32>::ET;
  constexpr i64 kTileF4 = (i64)kPT * kPT * sizeof(ET) / 16;
  const int lane = threadIdx.x & 63;
  const i64 wave = (i64)blockIdx.x * 4 + (threadIdx.x >> 6);
  const i64 T = (i64)nb * (nb + 1) / 2;
  const i64 t0 = wave * tpw;
  if (t0 >= T) return;
  const i64 t1 = min(t0 + (i64)tpw, T);
  int I, J;
  tile_ij(t0, I, J);
  double2 accC0 = make_double2(0.0, 0.0), accC1 = make_double2(0.0, 0.0);
  for (i64 t = t0; t < t1; ++t) {
    const f4v* tile = reinterpret_cast<const f4v*>(zp) + t * kTileF4;
    // x_I and x_J for this lane's rows; x_J[c] is broadcast with readlane
    double2 xI[2], xJ[2], accR[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const i64 gi = (i64)I * kPT + wrow<F32>(lane, e), gj = (i64)J * kPT + wrow<F32>(lane, e);
      xI[e] = gi < n ? x[gi] : make_double2(0.0, 0.0);
      xJ[e] = gj < n ? x[gj] : make_double2(0.0, 0.0);
    }
    dense_tile<F32>(tile, lane, I != J, xI, xJ, accR, accC0, accC1);
#pragma unroll
    for (int e = 0; e < 2; ++e) rowpart[t * kPT + wrow<F32>(lane, e)] = accR[e];
    // end of this column run: store the accumulated column sums at tile t
    if (I == J || t + 1 == t1) {
      colpart[t * kPT + 8 * (lane & 7) + (lane >> 3)] = accC0;
      colpart[t * kPT + 64 + 8 * (lane & 7) + (lane >> 3)] = accC1;
      accC0 = accC1 = make_double2(0.0, 0.0);
    }
    if (I == J) {
      ++J;
      I = 0;
    } else {
      ++I;
    }
  }
}

// ------------------------------------------------ compressed (hierarchical) form
// Packed complex64 storage in which off-diagonal tiles may be replaced by a
// cross approximation Z_IJ ~ U V^T (rank[t] >= 0; U, V: 128 x rank, slot t of
// fac holds U then V, each rmax columns of 128 complex64).  Dense tiles
// (rank[t] < 0, always the diagonal ones) are read from the packed matrix.
// For a low-rank tile, with s = U^T x_I and r = V^T x_J:
//   y_I += U r,  y_J += V s.
// Waves take contiguous ranges of stored tiles that carry about the same
// number of bytes (wstart, from the tiles' ranks).  A wave's tiles form one
// stream of groups: four U and four V columns of a low-rank tile, or eight
// columns of a dense tile (1 KB each, one coalesced 16-byte load per lane).
// The next group's eight loads (possibly of the next tile) and the next
// tile's x rows are issued unconditionally before the current group is
// reduced, into the other of two register sets: no load sits behind a
// branch, so the compiler's wait counts let them stay in flight.  Columns
// past a tile's rank re-read its last column and drop out through the
// broadcast sums.  The eight dot products of a group (or the eight column
// sums of a dense group) are reduced by one transposing butterfly on the
// cross-half permlane swaps and DPP row moves, then broadcast with readlane.
// Row sums go to rowpart per tile; column sums are kept in this lane's row
// layout (columns 2 lane, 2 lane + 1) and flushed to colpart at the end of a
// column run (the diagonal tile) or of the wave's range.  Both partials are
// stored lane-major (row 2 l + e at position 64 e + l), so each store
// instruction writes 1 KB contiguously.
__device__ inline double2 cfma(double2 acc, double2 a, double2 b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
  return acc;
}

// acc += a b (complex), four FMAs
__device__ inline void cmac(double2& acc, double2 a, double2 b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

__device__ inline double2 bcast_lane(double2 v, int l) { return make_double2(rl(v.x, l), rl(v.y, l)); }

template <int CTRL>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <int CTRL>
__device__ inline void dpp_acc(double2& a) {
  a.x += dpp_f64<CTRL>(a.x);
  a.y += dpp_f64<CTRL>(a.y);
}

// v_permlane32_swap / v_permlane16_swap on one double: a keeps its lower
// half-wave (row pair) and takes b's, b takes a's upper one.
__device__ inline void swap32(double& a, double& b) {
  const long long A = __double_as_longlong(a), B = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)(A & 0xffffffffLL), (unsigned)(B & 0xffffffffLL), false,
                                                   false);
  const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(A >> 32), (unsigned)(B >> 32), false, false);
  a = __longlong_as_double(((long long)hi[0] << 32) | (unsigned long long)lo[0]);
  b = __longlong_as_double(((long long)hi[1] << 32) | (unsigned long long)lo[1]);
}
__device__ inline void swap16(double& a, double& b) {
  const long long A = __double_as_longlong(a), B = __double_as_longlong(b);
  const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)(A & 0xffffffffLL), (unsigned)(B & 0xffffffffLL), false,
                                                   false);
  const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(A >> 32), (unsigned)(B >> 32), false, false);
  a = __longlong_as_double(((long long)hi[0] << 32) | (unsigned long long)lo[0]);
  b = __longlong_as_double(((long long)hi[1] << 32) | (unsigned long long)lo[1]);
}
// lanes 0..31: a(l) + a(l + 32); lanes 32..63: b(l - 32) + b(l)
__device__ inline double2 half32(double2 a, double2 b) {
  swap32(a.x, b.x);
  swap32(a.y, b.y);
  return make_double2(a.x + b.x, a.y + b.y);
}
// lanes with bit 4 clear: a(l) + a(l ^ 16); the others: b(l) + b(l ^ 16)
__device__ inline double2 half16(double2 a, double2 b) {
  swap16(a.x, b.x);
  swap16(a.y, b.y);
  return make_double2(a.x + b.x, a.y + b.y);
}

// The same result as butterfly8 (p[0] of lane l: the sum over all lanes of
// value l >> 3) without LDS: the xor-32 and xor-16 stages on permlane swaps,
// xor 8 as a row rotation by 8, the last three on quad permutes and the
// half-row mirror (every lane of an 8-lane group ends with the group's sum).
__device__ inline void butterfly8x(double2 (&p)[8], int lane) {
#pragma unroll
  for (int i = 0; i < 4; ++i) p[i] = half32(p[i], p[i + 4]);
#pragma unroll
  for (int i = 0; i < 2; ++i) p[i] = half16(p[i], p[i + 2]);
  {
    const bool h3 = lane & 8;
    double2 keep = h3 ? p[1] : p[0];
    const double2 send = h3 ? p[0] : p[1];
    keep.x += dpp_f64<0x128>(send.x);  // row_ror 8
    keep.y += dpp_f64<0x128>(send.y);
    p[0] = keep;
  }
  dpp_acc<0xB1>(p[0]);   // quad_perm 1,0,3,2
  dpp_acc<0x4E>(p[0]);   // quad_perm 2,3,0,1
  dpp_acc<0x141>(p[0]);  // row_half_mirror
}

// Partial vectors of the products: FP64 (double2), or FP32 (float2) where the
// product may carry a larger error anyway (the compression's looser levels).
__device__ inline void put_part(double2* p, double2 v) { *p = v; }
__device__ inline void put_part(float2* p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }
__device__ inline double2 get_part(const double2* p) { return *p; }
__device__ inline double2 get_part(const float2* p) {
  const float2 v = *p;
  return make_double2((double)v.x, (double)v.y);
}

template <typename P>
__global__ __launch_bounds__(256, 2) void k_symv_h(i64 n, int nb, int W, const int* __restrict__ wstart,
                                                const void* __restrict__ zp, const int* __restrict__ rank,
                                                const float2* __restrict__ fac, int rmax,
                                                const double2* __restrict__ x, P* __restrict__ rowpart,
                                                P* __restrict__ colpart) {
  constexpr i64 kTileF4 = (i64)kPT * kPT * sizeof(float2) / 16;
  const int lane = threadIdx.x & 63;
  const int wave = (int)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (wave >= W) return;
  const int t0 = wstart[wave], t1 = wstart[wave + 1];
  if (t0 >= t1) return;
  const f4v* Z4 = reinterpret_cast<const f4v*>(zp);
  const f4v* F4 = reinterpret_cast<const f4v*>(fac);
  const i64 slot = (i64)rmax * kPT;  // 16-byte words per factor slot
  int I, J;
  tile_ij(t0, I, J);
  I = __builtin_amdgcn_readfirstlane(I);
  J = __builtin_amdgcn_readfirstlane(J);
  double2 accC0 = make_double2(0.0, 0.0), accC1 = make_double2(0.0, 0.0);
  double2 accJ[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  double2 accR[2] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
  double2 xI[2], xJ[2], xIn[2];
  // x rows 2 lane, 2 lane + 1 of row block i (zero past n; the load itself
  // is clamped so that it is issued unconditionally)
  auto xload = [&](int i, double2(&o)[2]) {
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const i64 g = (i64)i * kPT + 2 * lane + e;
      const double2 v = x[min(g, n - 1)];
      o[e] = g < n ? v : make_double2(0.0, 0.0);
    }
  };
  // the eight loads of group k of tile t (rank rk; -1: dense, columns
  // k .. k + 7; rank 0: four finite columns of the dense tile, all dropped)
  auto gload = [&](int t, int k, int rk, f4v(&a)[4], f4v(&b)[4]) {
    const f4v* pa = Z4 + (i64)t * kTileF4;
    const f4v* pb = pa;
    int ca = 0, cb = 0, cl = 3;
    if (rk > 0) {
      pa = F4 + (i64)t * slot;
      pb = pa + slot / 2;
      ca = cb = k;
      cl = rk - 1;
    } else if (rk < 0) {
      ca = k;
      cb = k + 4;
      cl = kPT - 1;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      a[c] = __builtin_nontemporal_load(pa + (i64)min(ca + c, cl) * (kPT / 2) + lane);
      b[c] = __builtin_nontemporal_load(pb + (i64)min(cb + c, cl) * (kPT / 2) + lane);
    }
  };
  // factor columns k .. k + 3 (cnt of them below the rank)
  auto lowrank = [&](const f4v(&u)[4], const f4v(&v)[4], int cnt) {
    double2 p[8];  // p[c]: U_c . x_I over this lane's rows; p[4 + c]: V_c . x_J
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      p[c] = p[4 + c] = make_double2(0.0, 0.0);
      cmac(p[c], make_double2(u[c][0], u[c][1]), xI[0]);
      cmac(p[c], make_double2(u[c][2], u[c][3]), xI[1]);
      cmac(p[4 + c], make_double2(v[c][0], v[c][1]), xJ[0]);
      cmac(p[4 + c], make_double2(v[c][2], v[c][3]), xJ[1]);
    }
    butterfly8x(p, lane);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      double2 sc = bcast_lane(p[0], 8 * c), rc = bcast_lane(p[0], 32 + 8 * c);
      if (c >= cnt) sc = rc = make_double2(0.0, 0.0);
      cmac(accR[0], make_double2(u[c][0], u[c][1]), rc);
      cmac(accR[1], make_double2(u[c][2], u[c][3]), rc);
      cmac(accJ[0], make_double2(v[c][0], v[c][1]), sc);
      cmac(accJ[1], make_double2(v[c][2], v[c][3]), sc);
    }
  };
  // dense columns 8 b .. 8 b + 7; the column sums end in lanes 8 c + (b & 7)
  // (column 8 b + c), accC0 for b < 8, accC1 after
  auto dense8 = [&](int b, bool offdiag, const f4v(&a)[4], const f4v(&c4)[4]) {
    double2 p[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const f4v& q = c < 4 ? a[c] : c4[c - 4];
      const double2 z0 = make_double2(q[0], q[1]), z1 = make_double2(q[2], q[3]);
      const double2 xc = bcast<1>(xJ, 8 * b + c);
      cmac(accR[0], z0, xc);
      cmac(accR[1], z1, xc);
      p[c] = make_double2(0.0, 0.0);
      cmac(p[c], z0, xI[0]);
      cmac(p[c], z1, xI[1]);
    }
    if (offdiag) {
      butterfly8x(p, lane);
      if ((lane & 7) == (b & 7)) {
        if (b < 8)
          cadd(accC0, p[0]);
        else
          cadd(accC1, p[0]);
      }
    }
  };
  auto finish = [&](int t, bool flush) {
#pragma unroll
    for (int e = 0; e < 2; ++e) put_part(rowpart + (i64)t * kPT + 64 * e + lane, accR[e]);
    if (flush) {
#pragma unroll
      for (int e = 0; e < 2; ++e) {
        const int c = 2 * lane + e;
        const int src = 8 * (c & 7) + ((c >> 3) & 7);
        const double2 g0 = make_double2(__shfl(accC0.x, src, 64), __shfl(accC0.y, src, 64));
        const double2 g1 = make_double2(__shfl(accC1.x, src, 64), __shfl(accC1.y, src, 64));
        double2 v = c < 64 ? g0 : g1;
        cadd(v, accJ[e]);
        put_part(colpart + (i64)t * kPT + 64 * e + lane, v);
      }
      accC0 = accC1 = make_double2(0.0, 0.0);
      accJ[0] = accJ[1] = make_double2(0.0, 0.0);
    }
  };

  int tc = t0, kc = 0;
  int rkc = __builtin_amdgcn_readfirstlane(rank[tc]);
  f4v Au[4], Av[4], Bu[4], Bv[4];
  xload(I, xI);
  xload(J, xJ);
  gload(tc, kc, rkc, Au, Av);
  bool done = false;
  // one group from (cu, cv); the next group's loads go to (nu, nv) first
  auto step = [&](const f4v(&cu)[4], const f4v(&cv)[4], f4v(&nu)[4], f4v(&nv)[4]) {
    const bool dense = rkc < 0;
    const int lim = dense ? kPT : max(rkc, 1);
    const int adv = dense ? 8 : 4;
    const bool newtile = kc + adv >= lim;
    const bool endr = newtile && tc + 1 >= t1;
    int tn = tc, kn = kc + adv, rkn = rkc;
    if (newtile) {
      kn = kc;  // at the end of the range: the same group again (loads unused)
      if (!endr) {
        tn = tc + 1;
        kn = 0;
        rkn = __builtin_amdgcn_readfirstlane(rank[tn]);
      }
    }
    int In = I, Jn = J;
    if (newtile && !endr) {
      if (I == J) {
        Jn = J + 1;
        In = 0;
      } else {
        In = I + 1;
      }
    }
    xload(In, xIn);
    gload(tn, kn, rkn, nu, nv);
    if (dense)
      dense8(kc >> 3, I != J, cu, cv);
    else
      lowrank(cu, cv, rkc - kc);
    if (newtile) {
      finish(tc, I == J || endr);
      accR[0] = accR[1] = make_double2(0.0, 0.0);
      const bool newcol = Jn != J;
      I = In;
      J = Jn;
      if (newcol) xload(J, xJ);
    }
    xI[0] = xIn[0];
    xI[1] = xIn[1];
    tc = tn;
    kc = kn;
    rkc = rkn;
    done = endr;
  };
#pragma unroll 1
  for (;;) {
    step(Au, Av, Bu, Bv);
    if (done) break;
    step(Bu, Bv, Au, Av);
    if (done) break;
  }
}

// y[R kPT + r] = sum_{J >= R} rowpart[tile (R, J)][r]
//              + sum over the column runs of tile column R of their colpart
// (a run ends at tile t = R (R + 1) / 2 + I when t + 1 is a multiple of tpw
// or I = R); fixed order, deterministic.
template <typename P>
__global__ __launch_bounds__(256) void k_symv_w_reduce(i64 n, int nb, int tpw, const unsigned char* __restrict__ rend,
                                                       int lanemajor, const P* __restrict__ rowpart,
                                                       const P* __restrict__ colpart, double2* __restrict__ y) {
  // 32 rows per block, 8 groups sharing the tiles of a row block round robin;
  // the groups' sums are added in a fixed order
  const int rl = threadIdx.x & 31, grp = threadIdx.x >> 5;
  const i64 g = blockIdx.x * (i64)32 + rl;
  const int R = (int)min((i64)(nb - 1), (blockIdx.x * (i64)32) / kPT);
  // partial of row r: at r, or at 64 (r & 1) + (r >> 1) in the lane-major layout
  const int r = lanemajor ? (int)(((g & 1) << 6) + ((g % kPT) >> 1)) : (int)(g % kPT);
  double sr = 0, si = 0;
  if (g < n) {
    int J = R + grp;
    for (; J + 24 < nb; J += 32) {
      double2 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const i64 Ju = J + 8 * u;
        v[u] = get_part(rowpart + (Ju * (Ju + 1) / 2 + R) * kPT + r);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        sr += v[u].x;
        si += v[u].y;
      }
    }
    for (; J < nb; J += 8) {
      const double2 v = get_part(rowpart + ((i64)J * (J + 1) / 2 + R) * kPT + r);
      sr += v.x;
      si += v.y;
    }
    const i64 c0 = (i64)R * (R + 1) / 2;
    int I = grp;
    // four tiles per step: their end flags, then the flagged partials, are
    // loaded together (an unflagged tile adds an exact zero, so the sum is
    // the same as tile by tile)
    for (; I + 24 <= R; I += 32) {
      bool f[4];
      double2 v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const i64 t = c0 + I + 8 * u;
        f[u] = rend ? rend[t] != 0 : (I + 8 * u == R || (t + 1) % tpw == 0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const i64 t = c0 + I + 8 * u;
        v[u] = f[u] ? get_part(colpart + t * kPT + r) : make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        sr += v[u].x;
        si += v[u].y;
      }
    }
    for (; I <= R; I += 8) {
      const i64 t = c0 + I;
      if (rend ? rend[t] != 0 : (I == R || (t + 1) % tpw == 0)) {
        const double2 v = get_part(colpart + t * kPT + r);
        sr += v.x;
        si += v.y;
      }
    }
  }
  __shared__ double2 part[8][32];
  part[grp][rl] = make_double2(sr, si);
  __syncthreads();
  if (grp == 0 && g < n) {
    double2 acc = part[0][rl];
#pragma unroll
    for (int q = 1; q < 8; ++q) {
      acc.x += part[q][rl].x;
      acc.y += part[q][rl].y;
    }
    y[g] = acc;
  }
}

// Bytes-balanced wave ranges of the compressed product.  cost(t): the bytes
// a product reads for stored tile t plus a fixed 4 KB for its x loads and
// partials; pre: inclusive prefix of cost (one block, chunked scan).
__device__ inline i64 hcost(int r) { return r < 0 ? (i64)kPT * kPT * 8 + 4096 : (i64)2 * kPT * r * 8 + 4096; }

__global__ __launch_bounds__(1024) void k_hscan(i64 T, const int* __restrict__ rank, i64* __restrict__ pre) {
  __shared__ i64 part[1024];
  const int tid = threadIdx.x;
  const i64 per = (T + 1023) / 1024;
  const i64 a = min(T, tid * per), b = min(T, a + per);
  i64 sum = 0;
  for (i64 t = a; t < b; ++t) sum += hcost(rank[t]);
  part[tid] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const i64 v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  i64 base = tid ? part[tid - 1] : 0;
  for (i64 t = a; t < b; ++t) {
    base += hcost(rank[t]);
    pre[t] = base;
  }
}

// rend[t] = 1 for diagonal tiles (the end of every tile column)
__global__ void k_hflags(i64 T, unsigned char* __restrict__ rend) {
  const i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (t >= T) return;
  int I, J;
  tile_ij(t, I, J);
  rend[t] = I == J ? 1 : 0;
}

// wstart[w] = first tile whose prefix cost exceeds w / W of the total (the
// tile holding that byte offset starts range w); the tile before each range
// start ends a column run.
__global__ void k_hranges(i64 T, int W, const i64* __restrict__ pre, int* __restrict__ wstart,
                          unsigned char* __restrict__ rend) {
  const i64 w = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (w > W) return;
  if (w == 0) {
    wstart[0] = 0;
    return;
  }
  if (w == W) {
    wstart[W] = (int)T;
    return;
  }
  const i64 target = (i64)((double)pre[T - 1] * (double)w / (double)W);
  i64 lo = 0, hi = T;  // first t with pre[t] > target
  while (lo < hi) {
    const i64 mid = (lo + hi) / 2;
    if (pre[mid] > target)
      hi = mid;
    else
      lo = mid + 1;
  }
  wstart[w] = (int)lo;
  if (lo > 0) rend[lo - 1] = 1;
}

// The three range kernels over several rank arrays at once (one per level):
// the level is blockIdx.x (scan) or blockIdx.y.
constexpr int kMaxLevels = 16;
struct LevelPtrs {
  const int* rank[kMaxLevels];
  i64* pre[kMaxLevels];
  unsigned char* rend[kMaxLevels];
  int* wstart[kMaxLevels];
};

__global__ __launch_bounds__(1024) void k_hscan_lv(i64 T, LevelPtrs lp) {
  const int l = blockIdx.x;
  const int* __restrict__ rank = lp.rank[l];
  i64* __restrict__ pre = lp.pre[l];
  __shared__ i64 part[1024];
  const int tid = threadIdx.x;
  const i64 per = (T + 1023) / 1024;
  const i64 a = min(T, tid * per), b = min(T, a + per);
  i64 sum = 0;
  for (i64 t = a; t < b; ++t) sum += hcost(rank[t]);
  part[tid] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const i64 v = tid >= o ? part[tid - o] : 0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  i64 base = tid ? part[tid - 1] : 0;
  for (i64 t = a; t < b; ++t) {
    base += hcost(rank[t]);
    pre[t] = base;
  }
}

__global__ void k_hflags_lv(i64 T, LevelPtrs lp) {
  const i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (t >= T) return;
  int I, J;
  tile_ij(t, I, J);
  lp.rend[blockIdx.y][t] = I == J ? 1 : 0;
}

__global__ void k_hranges_lv(i64 T, int W, LevelPtrs lp) {
  const i64 w = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  const int l = blockIdx.y;
  if (w > W) return;
  int* wstart = lp.wstart[l];
  if (w == 0) {
    wstart[0] = 0;
    return;
  }
  if (w == W) {
    wstart[W] = (int)T;
    return;
  }
  const i64* pre = lp.pre[l];
  const i64 target = (i64)((double)pre[T - 1] * (double)w / (double)W);
  i64 lo = 0, hi = T;  // first t with pre[t] > target
  while (lo < hi) {
    const i64 mid = (lo + hi) / 2;
    if (pre[mid] > target)
      hi = mid;
    else
      lo = mid + 1;
  }
  wstart[w] = (int)lo;
  if (lo > 0) lp.rend[l][lo - 1] = 1;
}

}  // namespace

namespace dml {

size_t hier_ranges_bytes(i64 n, int W) {
  return (size_t)packed_tiles(n) * (sizeof(i64) + 1) + sizeof(int) * ((size_t)W + 1) + 64;
}

int hier_wave_slots() {
  int dev = 0, ncu = 0, per_cu = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_symv_h<double2>, 256, 0));
  return std::max(per_cu, 1) * ncu * 4;
}

void hier_ranges(i64 n, const int* rank, int W, void* buf, HierRanges& hr, hipStream_t s) {
  const i64 T = packed_tiles(n);
  DM_REQUIRE(W >= 1 && W < (1 << 30), DM_ERR_ARGUMENT, "wave count out of range");
  i64* pre = reinterpret_cast<i64*>(buf);
  hr.rend = reinterpret_cast<unsigned char*>(pre + T);
  hr.wstart = reinterpret_cast<int*>(reinterpret_cast<uintptr_t>(hr.rend + T + 15) & ~uintptr_t(15));
  hr.W = W;
  k_hscan<<<1, 1024, 0, s>>>(T, rank, pre);
  k_hflags<<<(unsigned)((T + 255) / 256), 256, 0, s>>>(T, hr.rend);
  k_hranges<<<(unsigned)((W + 1 + 255) / 256), 256, 0, s>>>(T, W, pre, hr.wstart, hr.rend);
  HIP_OK(hipGetLastError());
}

void hier_ranges_levels(i64 n, int nl, const int* const* ranks, int W, void* const* bufs, HierRanges* hr,
                        hipStream_t s) {
  const i64 T = packed_tiles(n);
  DM_REQUIRE(W >= 1 && W < (1 << 30) && nl >= 1 && nl <= kMaxLevels, DM_ERR_ARGUMENT, "wave ranges out of range");
  LevelPtrs lp;
  for (int l = 0; l < nl; ++l) {
    i64* pre = reinterpret_cast<i64*>(bufs[l]);
    hr[l].rend = reinterpret_cast<unsigned char*>(pre + T);
    hr[l].wstart = reinterpret_cast<int*>(reinterpret_cast<uintptr_t>(hr[l].rend + T + 15) & ~uintptr_t(15));
    hr[l].W = W;
    lp.rank[l] = ranks[l];
    lp.pre[l] = pre;
    lp.rend[l] = hr[l].rend;
    lp.wstart[l] = hr[l].wstart;
  }
  k_hscan_lv<<<nl, 1024, 0, s>>>(T, lp);
  k_hflags_lv<<<dim3((unsigned)((T + 255) / 256), nl), 256, 0, s>>>(T, lp);
  k_hranges_lv<<<dim3((unsigned)((W + 1 + 255) / 256), nl), 256, 0, s>>>(T, W, lp);
  HIP_OK(hipGetLastError());
}

template <typename P>
static void symv_hier_launch(i64 n, const float* z32p, const int* rank, const float* fac, int rmax, const double* x,
                             double* y, void* ws, const HierRanges& hr, hipStream_t s) {
  const i64 nb = (n + kPT - 1) / kPT;
  const i64 T = packed_tiles(n);
  const i64 blocks = ((i64)hr.W + 3) / 4;
  DM_REQUIRE(hr.W >= 1 && blocks < (1LL << 31), DM_ERR_ARGUMENT, "wave count out of range");
  P* rowpart = reinterpret_cast<P*>(ws);
  P* colpart = rowpart + T * kPT;
  k_symv_h<P><<<(unsigned)blocks, 256, 0, s>>>(n, (int)nb, hr.W, hr.wstart, z32p, rank,
                                               reinterpret_cast<const float2*>(fac), rmax,
                                               reinterpret_cast<const double2*>(x), rowpart, colpart);
  k_symv_w_reduce<P><<<(unsigned)((n + 31) / 32), 256, 0, s>>>(n, (int)nb, 1, hr.rend, 1, rowpart, colpart,
                                                               reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}

void symv_hier(i64 n, const float* z32p, const int* rank, const float* fac, int rmax, const double* x, double* y,
               void* ws, const HierRanges& hr, hipStream_t s, bool f32part) {
  if (f32part)
    symv_hier_launch<float2>(n, z32p, rank, fac, rmax, x, y, ws, hr, s);
  else
    symv_hier_launch<double2>(n, z32p, rank, fac, rmax, x, y, ws, hr, s);
}

std::vector<int> symv_shapes(i64 n, bool f32) {
  std::vector<int> c = {4, 8, 16, kWaveShape + 1, kWaveShape + 2, kWaveShape + 4};
  // wave form sized so that every resident wave slot gets one or two equal ranges
  int dev = 0, ncu = 0, per_cu = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const void* fn = f32 ? (const void*)k_symv_w<1> : (const void*)k_symv_w<0>;
  HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0));
  const i64 slots = (i64)std::max(per_cu, 1) * ncu * 4;
  const i64 T = packed_tiles(n);
  for (int k = 1; k <= 2; ++k) {
    const i64 t = (T + k * slots - 1) / (k * slots);
    if (t > 4 && t <= 1024) c.push_back(kWaveShape + (int)t);
  }
  return c;
}

size_t symv_workspace_bytes(i64 n) {
  const i64 nb = (n + kPT - 1) / kPT;
  const i64 lds_form = packed_tiles(n) + seg_base(nb, kSegMin);
  const i64 wave_form = 2 * packed_tiles(n);
  return (size_t)std::max(lds_form, wave_form) * kPT * sizeof(double2);
}

void symv_packed(i64 n, const ZRef& z, const double* x, double* y, void* ws, int shape, hipStream_t s) {
  DM_REQUIRE(z.packed, DM_ERR_ARGUMENT, "symmetric product needs packed storage");
  const i64 nb = (n + kPT - 1) / kPT;
  if (shape >= kWaveShape) {
    const int tpw = shape - kWaveShape;
    DM_REQUIRE(tpw >= 1 && tpw <= 1024, DM_ERR_ARGUMENT, "tiles per wave out of range");
    const i64 T = packed_tiles(n);
    const i64 waves = (T + tpw - 1) / tpw;
    const i64 blocks = (waves + 3) / 4;
    DM_REQUIRE(blocks < (1LL << 31), DM_ERR_ARGUMENT, "matrix too large for the packed product");
    double2* rowpart = reinterpret_cast<double2*>(ws);
    double2* colpart = rowpart + T * kPT;
    const double2* xv = reinterpret_cast<const double2*>(x);
    if (z.f32)
      k_symv_w<1><<<(unsigned)blocks, 256, 0, s>>>(n, (int)nb, tpw, z.p, xv, rowpart, colpart);
    else
      k_symv_w<0><<<(unsigned)blocks, 256, 0, s>>>(n, (int)nb, tpw, z.p, xv, rowpart, colpart);
    k_symv_w_reduce<double2><<<(unsigned)((n + 31) / 32), 256, 0, s>>>(n, (int)nb, tpw, nullptr, 0, rowpart, colpart,
                                                              reinterpret_cast<double2*>(y));
    HIP_OK(hipGetLastError());
    return;
  }
  const int seg = shape;
  DM_REQUIRE(seg >= kSegMin && seg <= 64, DM_ERR_ARGUMENT, "segment length out of range");
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
