// k_aca.hip — compression of the packed complex64 matrix: every off-diagonal
// stored tile (128 x 128) whose numerical rank is small is replaced by a cross
// approximation Z_IJ ~ U V^T.
//
// The reference compresses far blocks with adaptive cross approximation over
// entries computed on demand (src/fast/ACA.jl:558-688).  Here the dense tile
// is already in HBM after assembly, so the cross approximation runs with full
// pivoting on the whole tile held in registers (512 threads x 32 entries):
// each step picks the largest residual entry (i*, j*) (to 2^-12 relative in
// |R|^2), takes U[:, k] = R[:, j*] and V[:, k] = R[i*, :] / R[i*, j*], and
// subtracts U[:, k] V[:, k]^T from the residual.  The Frobenius norm of the
// residual is summed over all its entries at every step (FP32 within a wave,
// FP64 across the waves: relative error ~1e-6), so the stopping test
// ||R||_F <= tau is a bound on the tile's approximation error (up to that
// and the FP32 rounding of the updates, ~1e-7 of the tile's norm), not an
// estimate.  With
// tau^2 = eps^2 D / (2 T_off) for T_off stored off-diagonal tiles, where
// D = sum of the diagonal tiles' ||Z_JJ||_F^2 <= ||Z||_F^2 (computed on the
// device, so the whole compression is queued without a host round trip; the
// diagonal tiles carry most of the EFIE matrix's norm, ~90 % on a sphere, and
// reading them is 1/(nb/2 + 1/2) of reading the whole matrix), the
// whole matrix (where each such tile appears twice) satisfies
// ||Z - Z_H||_F <= eps sqrt(D) <= eps ||Z||_F.
//
// The steps of one run are nested: the first k columns of the factors are the
// cross approximation a looser tolerance would have stopped at.  The kernel
// records, per tile, the first step at which ||R||_F fell below
// 10^(l/2) tau (l = 1 .. kHLevels, "level l"), so a product can read only
// that prefix of the factors (error bound eps 10^(l/2)) at no extra
// compression cost; tiles that stay dense at eps may be low-rank at a looser
// level.  Tiles that need more than rmax steps stay
// dense (rank -1), as do the diagonal tiles.
#include <hip/hip_runtime.h>

#include "dm_common.h"
#include "k_linalg.h"

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
constexpr int kAcaThreads = 512;

// Sum of |z|^2 over each diagonal stored tile (FP64), one workgroup per tile
// (workgroup J: stored tile J (J + 3) / 2).
__global__ __launch_bounds__(256) void k_tile_norm2(const float2* __restrict__ zp, double* __restrict__ out) {
  const i64 t = (i64)blockIdx.x * (blockIdx.x + 3) / 2;
  const f4v* tile = reinterpret_cast<const f4v*>(zp + t * kPT * kPT);
  double s = 0;
  for (int k = threadIdx.x; k < kPT * kPT / 2; k += 256) {
    const f4v v = __builtin_nontemporal_load(tile + k);
    s += (double)v[0] * v[0] + (double)v[1] * v[1] + (double)v[2] * v[2] + (double)v[3] * v[3];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ double ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
}

// tau^2 = eps^2 D / (2 T_off), D the sum of the nb diagonal tiles' norms;
// one workgroup, fixed order.
__global__ __launch_bounds__(256) void k_tau2(i64 T, int nb, const double* __restrict__ norms, double eps2,
                                              double* __restrict__ tau2) {
  double s = 0;
  for (int J = threadIdx.x; J < nb; J += 256) s += norms[J];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ double ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double d2 = (ws[0] + ws[1]) + (ws[2] + ws[3]);
    const i64 toff = T - nb;
    tau2[0] = toff > 0 ? eps2 * d2 / (2.0 * (double)toff) : 0.0;
  }
}

// Wave reductions through DPP row operations (no LDS traffic): quad
// permutes, half-row and row mirrors, then the row 15 / row 31 broadcasts;
// lane 63 ends with the result over all 64 lanes (the identity 0 enters
// where a broadcast leaves a row out).
__device__ inline unsigned wave_max_u32(unsigned x) {
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0xB1, 0xF, 0xF, false));   // quad_perm 1,0,3,2
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x4E, 0xF, 0xF, false));   // quad_perm 2,3,0,1
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x141, 0xF, 0xF, false));  // row_half_mirror
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x140, 0xF, 0xF, false));  // row_mirror
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false));  // row_bcast15
  x = max(x, (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false));  // row_bcast31
  return (unsigned)__builtin_amdgcn_readlane((int)x, 63);
}
template <int CTRL, int ROWS>
__device__ inline float dpp_add(float v) {
  return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
__device__ inline float wave_sum_f32(float x) {
  x = dpp_add<0xB1, 0xF>(x);
  x = dpp_add<0x4E, 0xF>(x);
  x = dpp_add<0x141, 0xF>(x);
  x = dpp_add<0x140, 0xF>(x);
  x = dpp_add<0x142, 0xA>(x);
  x = dpp_add<0x143, 0xC>(x);
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// Thread t owns column t % 128, rows 32 (t / 128) .. + 31 of the tile: the
// rows a wave holds form one quarter of the tile, so extracting the pivot row
// is work for the two waves of that quarter only.
__global__ __launch_bounds__(kAcaThreads, 4) void k_aca(int nb, const float2* __restrict__ zp,
                                                     const double* __restrict__ tau2p, int rmax, int* __restrict__ rank,
                                                     int* __restrict__ lvrank, i64 T, float2* __restrict__ fac) {
  // Workgroups take the tiles by distance from the diagonal, d = J - I
  // ascending (S(d) = d nb - d (d - 1) / 2 tiles are nearer): the tiles
  // near the diagonal, which need the most steps, start first and the cheap
  // far tiles fill the end of the launch.
  const double tau2 = tau2p[0];
  const i64 b = blockIdx.x;
  auto S = [&](i64 dd) { return dd * nb - dd * (dd - 1) / 2; };
  const double A = (double)nb + 0.5;
  int d = (int)(A - sqrt(A * A - 2.0 * (double)b));
  while (d > 0 && S(d) > b) --d;
  while (S(d + 1) <= b) ++d;
  const int I = (int)(b - S(d)), J = I + d;
  const i64 t = (i64)J * (J + 1) / 2 + I;
  const int tid = threadIdx.x;
  if (I == J) {
    if (tid == 0) {
      rank[t] = -1;
      if (lvrank)
        for (int l = 0; l < kHLevels; ++l) lvrank[l * T + t] = -1;
    }
    return;
  }
  // level l + 1 stops at ||R||^2 <= 10^(l + 1) tau^2 (uniform over the block)
  int lvk[kHLevels];
#pragma unroll
  for (int l = 0; l < kHLevels; ++l) lvk[l] = -1;
  const int c = tid & (kPT - 1), qd = tid >> 7, r0 = qd * 32, lane = tid & 63, w = tid >> 6;
  float Rr[32], Ri[32];
  {
    const f4v* src = reinterpret_cast<const f4v*>(zp + t * kPT * kPT + (i64)c * kPT + r0);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const f4v v = src[q];
      Rr[2 * q] = v[0];
      Ri[2 * q] = v[1];
      Rr[2 * q + 1] = v[2];
      Ri[2 * q + 1] = v[3];
    }
  }
  __shared__ float2 colb[kPT], rowb[kPT];
  __shared__ unsigned wkey[kAcaThreads / 64];
  __shared__ double wsum[kAcaThreads / 64];
  float2* U = fac + t * 2 * kPT * (i64)rmax;
  float2* V = U + (i64)rmax * kPT;
  // The wave's largest residual entry and its share of ||R||^2.  The maximum
  // is taken over keys: |R|^2 with the low 11 mantissa bits replaced by the
  // entry's place in the wave (lane, register), so one integer max finds the
  // pivot (the largest entry to 2^-12 relative in |R|^2; equal keys resolve
  // to the higher place).  The FP32 sums (relative error ~1e-6 of the wave's
  // share) are far more accurate than the stopping test needs.
  auto publish = [&]() {
    float sumf = 0.f;
    unsigned key = 0;
#pragma unroll
    for (int q = 0; q < 32; ++q) {
      const float a = Rr[q] * Rr[q] + Ri[q] * Ri[q];
      sumf += a;
      key = max(key, (__float_as_uint(a) & ~0x7ffu) | (unsigned)q);
    }
    key = wave_max_u32(key | ((unsigned)lane << 5));
    const float sum = wave_sum_f32(sumf);
    if (lane == 0) {
      wkey[w] = key;
      wsum[w] = (double)sum;
    }
  };
  // flat index (column * 128 + row) of the entry behind wave wv's key
  auto key_index = [](int wv, unsigned kk) {
    const int col = (wv & 1) * 64 + (int)((kk >> 5) & 63u), row = (wv >> 1) * 32 + (int)(kk & 31u);
    return col * kPT + row;
  };
  publish();
  int k = 0;
  for (;; ++k) {
    // two barriers per step: the waves' results are read between the first
    // and the second, and rewritten only after the second
    __syncthreads();
    // the waves' results in wave order (ties between waves: the lower wave)
    unsigned bk = wkey[0] & ~0x7ffu;
    int bw = 0;
    double tot = wsum[0];
#pragma unroll
    for (int q = 1; q < kAcaThreads / 64; ++q) {
      const unsigned kq = wkey[q] & ~0x7ffu;
      if (kq > bk) {
        bk = kq;
        bw = q;
      }
      tot += wsum[q];
    }
    const int piv_idx = key_index(bw, wkey[bw]);
    {
      double lt = tau2;
#pragma unroll
      for (int l = 0; l < kHLevels; ++l) {
        lt *= 10.0;
        if (lvk[l] < 0 && tot <= lt) lvk[l] = k;
      }
    }
    if (tot <= tau2) break;  // rank k reaches the tolerance
    if (k == rmax) {
      k = -1;  // not compressible within rmax: stays dense
      break;
    }
    const int pj = piv_idx / kPT, pi = piv_idx % kPT;
    if (c == pj) {
#pragma unroll
      for (int q = 0; q < 32; ++q) colb[r0 + q] = make_float2(Rr[q], Ri[q]);
    }
    if ((pi >> 5) == qd) {  // uniform over the wave
      const int qq = pi & 31;
      float vr = 0.f, vi = 0.f;
#pragma unroll
      for (int q = 0; q < 32; ++q)
        if (q == qq) {
          vr = Rr[q];
          vi = Ri[q];
        }
      rowb[c] = make_float2(vr, vi);
    }
    __syncthreads();
    // 1 / pivot once per step (the same value in every thread), then products
    const float2 piv = colb[pi];
    const float pd = 1.0f / (piv.x * piv.x + piv.y * piv.y);
    const float2 ip = make_float2(piv.x * pd, -piv.y * pd);
    auto over_piv = [&](float2 a) { return make_float2(a.x * ip.x - a.y * ip.y, a.x * ip.y + a.y * ip.x); };
    if (tid < kPT)
      U[(i64)k * kPT + tid] = colb[tid];
    else if (tid < 2 * kPT)
      V[(i64)k * kPT + tid - kPT] = over_piv(rowb[tid - kPT]);
    const float2 vc = over_piv(rowb[c]);
#pragma unroll
    for (int q = 0; q < 32; ++q) {
      const float2 u = colb[r0 + q];
      Rr[q] = fmaf(-u.x, vc.x, fmaf(u.y, vc.y, Rr[q]));
      Ri[q] = fmaf(-u.x, vc.y, fmaf(-u.y, vc.x, Ri[q]));
    }
    publish();
  }
  if (tid == 0) {
    rank[t] = k;
    if (lvrank)
      for (int l = 0; l < kHLevels; ++l) lvrank[l * T + t] = lvk[l];
  }
}

}  // namespace

namespace dml {

void aca_compress(i64 n, const float* z32p, double eps, int rmax, double* work, int* rank, int* lvrank, float* fac,
                  hipStream_t s) {
  const i64 nb = (n + kPT - 1) / kPT;
  const i64 T = packed_tiles(n);
  DM_REQUIRE(rmax >= 1 && rmax <= kPT / 2, DM_ERR_ARGUMENT, "cross approximation rank cap out of range");
  DM_REQUIRE(T < (1LL << 31), DM_ERR_ARGUMENT, "matrix too large");
  k_tile_norm2<<<(unsigned)nb, 256, 0, s>>>(reinterpret_cast<const float2*>(z32p), work);
  k_tau2<<<1, 256, 0, s>>>(T, (int)nb, work, eps * eps, work + T);
  k_aca<<<(unsigned)T, kAcaThreads, 0, s>>>((int)nb, reinterpret_cast<const float2*>(z32p), work + T, rmax, rank,
                                            lvrank, T, reinterpret_cast<float2*>(fac));
  HIP_OK(hipGetLastError());
}

}  // namespace dml
