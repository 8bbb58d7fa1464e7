// k_precond.hip — near-field sparse approximate inverse (SPAI) preconditioner.
//
// The reference preconditions GMRES with the LU factors of the near-field
// matrix Z_nf (entries whose RWG centres are within a cutoff,
// src/solver/NearFieldPreconditioner.jl:670-717).  A sparse LU is a serial,
// fill-in driven factorisation; the GPU form here instead minimises
// || Z M - I ||_F column by column over the same kind of geometric near-field
// pattern (Frobenius-norm minimisation, as used for EFIE by Alleon, Benzi and
// Giraud 1997 / Carpentieri et al. 2000):
//   m_j = argmin || Z[I_j, J_j] m - e_j ||_2,
// J_j = bases within cutJ of basis j (the columns' support), I_j = bases within
// cutI >= cutJ (the rows that carry the significant part of Z[:, J_j]).
// Each column is an independent small least-squares problem solved through the
// normal equations (G = A^H A, Cholesky in LDS, FP64) by one workgroup; the
// result is applied as a CSR SpMV, y_n = sum_{j in J(n)} M[n, j] x_j.
#include <hip/hip_runtime.h>

#include "dm_common.h"
#include "k_linalg.h"

namespace {

constexpr int SPAI_MAXJ = 80;   // columns per least-squares problem (LDS bound)
constexpr int SPAI_CHUNK = 8;   // rows of A staged per pass

// One column per workgroup.  LDS holds, for the widest support mJ of the
// pattern (not the cap), the Gram matrix's lower triangle packed by rows
// (entry (a, b), b <= a, at a (a + 1) / 2 + b), SPAI_CHUNK staged rows of A,
// the right-hand side and the support's column ids: about 27 KB at mJ = 49,
// so five columns share a CU (with at most 102 VGPRs: five waves per SIMD).  The four waves form the Gram matrix (each
// thread owns 2 x 2 blocks of it: one LDS read per complex multiply-add, two
// barriers per chunk of rows); the Cholesky factorization (left-looking)
// and the two triangular solves (right-hand side in registers, each solved
// entry broadcast by a lane read) are then wave 0's alone, ordered by
// wave-level fences instead of workgroup barriers (the other waves return).
__device__ inline int tri(int a, int b) { return a * (a + 1) / 2 + b; }

// (a, b) with b <= a from the packed lower-triangle index e = a (a + 1) / 2 + b
__device__ inline void untri(int e, int& a, int& b) {
  a = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
  while (a * (a + 1) / 2 > e) --a;
  while ((a + 1) * (a + 2) / 2 <= e) ++a;
  b = e - a * (a + 1) / 2;
}

// The lanes of one wave see each other's earlier LDS writes after this point
// (a code-motion barrier between wavefront-scope release and acquire fences).
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kSpaiThreads = 256;
constexpr int kAchLd = SPAI_CHUNK + 1;  // staged column stride (odd: conflict-free LDS reads)
constexpr int kSpaiPre = (SPAI_CHUNK * SPAI_MAXJ + kSpaiThreads - 1) / kSpaiThreads;  // staged entries per thread
static_assert(SPAI_MAXJ < 128, "a lane holds at most two rows of a column's system");

// v of lane src (uniform) in every lane
__device__ inline double bcast(double v, int src) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffLL), src);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), src);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline double2 bcast2(double2 v, int src) { return make_double2(bcast(v.x, src), bcast(v.y, src)); }

__global__ __launch_bounds__(kSpaiThreads, 5) void k_spai_build(int n, int per, ZRef z, int mJ,
                                                             const int* __restrict__ jptr,
                                                             const int* __restrict__ jidx,
                                                             const int* __restrict__ iptr,
                                                             const int* __restrict__ iidx,
                                                             double2* __restrict__ vals, int* __restrict__ status) {
  // XCD-aware order: workgroup w runs on XCD w % 8, which takes the columns
  // [x per, (x + 1) per) in order, so the columns in flight on one XCD are
  // neighbours whose rows of Z overlap and are read through the same L2.
  const int j = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (j >= n) return;
  const int j0 = jptr[j], nJ = jptr[j + 1] - j0;
  const int i0 = iptr[j], nI = iptr[j + 1] - i0;
  extern __shared__ __align__(16) char smem[];
  double2* G = reinterpret_cast<double2*>(smem);
  double2* Ach = G + mJ * (mJ + 1) / 2;
  double2* rhs = Ach + kAchLd * mJ;
  int* cols = reinterpret_cast<int*>(rhs + mJ);
  const int tid = threadIdx.x;
  for (int a = tid; a < nJ; a += kSpaiThreads) {
    const int c = jidx[j0 + a];
    cols[a] = c;
    // rhs = A^H e_j = conj(Z[j, J]) (row j of Z restricted to the pattern)
    const double2 v = zat(z, j, c);
    rhs[a] = make_double2(v.x, -v.y);
  }
  for (int e = tid; e < nJ * (nJ + 1) / 2; e += kSpaiThreads) G[e] = make_double2(0.0, 0.0);
  // G = A^H A accumulated over chunks of SPAI_CHUNK rows staged column by
  // column (Ach[a * kAchLd + r]; rows past the end of A are zero).  Thread t
  // owns the 2 x 2 blocks 255 - t, 511 - t, ... of the block lower triangle
  // (block rows A >= B; a second block falls to the last wave, not to wave 0,
  // which goes on to the factorization) and adds each chunk's sums to them.
  const int nb = (nJ + 1) / 2, nblk = nb * (nb + 1) / 2;
  // each thread stages the same (row, column) places of every chunk; the
  // next chunk's entries are loaded into registers before this chunk's sums
  // are formed, so their latency is hidden behind the arithmetic
  int pr[kSpaiPre], pa[kSpaiPre];
  double2 pre[kSpaiPre];
#pragma unroll
  for (int q = 0; q < kSpaiPre; ++q) {
    const int e = tid + q * kSpaiThreads;
    pr[q] = e < SPAI_CHUNK * nJ ? e / nJ : SPAI_CHUNK;
    pa[q] = e < SPAI_CHUNK * nJ ? e - pr[q] * nJ : 0;
  }
  auto fetch = [&](int r0) {
#pragma unroll
    for (int q = 0; q < kSpaiPre; ++q)
      pre[q] = pr[q] < SPAI_CHUNK && r0 + pr[q] < nI ? zat(z, iidx[i0 + r0 + pr[q]], cols[pa[q]])
                                                     : make_double2(0.0, 0.0);
  };
  __syncthreads();  // cols
  fetch(0);
  for (int r0 = 0; r0 < nI; r0 += SPAI_CHUNK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kSpaiPre; ++q)
      if (pr[q] < SPAI_CHUNK) Ach[pa[q] * kAchLd + pr[q]] = pre[q];
    __syncthreads();
    if (r0 + SPAI_CHUNK < nI) fetch(r0 + SPAI_CHUNK);
    for (int e = kSpaiThreads - 1 - tid; e < nblk; e += kSpaiThreads) {
      int A, B;
      untri(e, A, B);
      // columns past nJ - 1 (odd nJ) read column nJ - 1 and are not stored
      const double2* ca0 = Ach + 2 * A * kAchLd;
      const double2* ca1 = Ach + min(2 * A + 1, nJ - 1) * kAchLd;
      const double2* cb0 = Ach + 2 * B * kAchLd;
      const double2* cb1 = Ach + min(2 * B + 1, nJ - 1) * kAchLd;
      double sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
      auto mac = [&](int u, double2 p, double2 q) {
        // conj(p) q
        sr[u] = fma(p.x, q.x, sr[u]);
        sr[u] = fma(p.y, q.y, sr[u]);
        si[u] = fma(p.x, q.y, si[u]);
        si[u] = fma(-p.y, q.x, si[u]);
      };
#pragma unroll 4
      for (int ii = 0; ii < SPAI_CHUNK; ++ii) {
        const double2 p0 = ca0[ii], p1 = ca1[ii], q0 = cb0[ii], q1 = cb1[ii];
        mac(0, p0, q0);
        mac(1, p0, q1);
        mac(2, p1, q0);
        mac(3, p1, q1);
      }
      const int a0 = 2 * A, b0 = 2 * B;
      auto add = [&](int a, int b, int u) {
        G[tri(a, b)].x += sr[u];
        G[tri(a, b)].y += si[u];
      };
      add(a0, b0, 0);
      if (b0 + 1 <= a0) add(a0, b0 + 1, 1);
      if (a0 + 1 < nJ) {
        add(a0 + 1, b0, 2);
        if (b0 + 1 < nJ) add(a0 + 1, b0 + 1, 3);
      }
    }
  }
  __syncthreads();
  if (tid >= 64) return;
  const int lane = tid;
  bool bad = false;
  // the right-hand side in registers: lane l holds rows l and l + 64
  // (nJ <= SPAI_MAXJ < 128); its LDS place then keeps 1 / L[k][k]
  double2 b[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = lane + 64 * h;
    b[h] = i < nJ ? rhs[i] : make_double2(0.0, 0.0);
  }
  double* dinv = reinterpret_cast<double*>(rhs);
  wave_sync();
  // Cholesky G = L L^H in place (lower triangle), left-looking: column k of
  // L from the dot products of rows i >= k with row k over the columns
  // p < k (independent LDS reads, no round trip per update).  Lane l owns
  // rows l and l + 64.
  for (int k = 0; k < nJ; ++k) {
    double2 s[2];
    double dk = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = lane + 64 * h;
      s[h] = make_double2(0.0, 0.0);
      if (i < k || i >= nJ) continue;
      double2 v = G[tri(i, k)];
      for (int p = 0; p < k; ++p) {
        const double2 a = G[tri(i, p)], b = G[tri(k, p)];
        // v -= L[i][p] conj(L[k][p])
        v.x = fma(-a.x, b.x, v.x);
        v.x = fma(-a.y, b.y, v.x);
        v.y = fma(-a.y, b.x, v.y);
        v.y = fma(a.x, b.y, v.y);
      }
      s[h] = v;
      if (i == k) dk = v.x;
    }
    double d = bcast(dk, k & 63);
    if (!(d > 0.0)) {
      bad = true;
      d = 1e-300;
    }
    const double lkk = sqrt(d), il = 1.0 / lkk;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = lane + 64 * h;
      if (i < k || i >= nJ) continue;
      G[tri(i, k)] = i == k ? make_double2(lkk, 0.0) : make_double2(s[h].x * il, s[h].y * il);
    }
    if (lane == 0) dinv[k] = il;
    wave_sync();
  }
  // forward: L y = rhs (y_k reaches every lane through a lane read)
  for (int k = 0; k < nJ; ++k) {
    const double il = dinv[k];
    const double2 bk = bcast2(k < 64 ? b[0] : b[1], k & 63);
    const double2 yk = make_double2(bk.x * il, bk.y * il);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = lane + 64 * h;
      if (i == k) b[h] = yk;
      if (i <= k || i >= nJ) continue;
      const double2 l = G[tri(i, k)];
      b[h].x -= l.x * yk.x - l.y * yk.y;
      b[h].y -= l.x * yk.y + l.y * yk.x;
    }
  }
  // backward: L^H m = y
  for (int k = nJ - 1; k >= 0; --k) {
    const double il = dinv[k];
    const double2 bk = bcast2(k < 64 ? b[0] : b[1], k & 63);
    const double2 mk = make_double2(bk.x * il, bk.y * il);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = lane + 64 * h;
      if (i == k) b[h] = mk;
      if (i >= k) continue;
      const double2 l = G[tri(k, i)];  // (L^H)[i][k] = conj(L[k][i])
      b[h].x -= l.x * mk.x + l.y * mk.y;
      b[h].y -= l.x * mk.y - l.y * mk.x;
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = lane + 64 * h;
    if (i < nJ) vals[j0 + i] = b[h];
  }
  if (lane == 0 && bad) status[0] = j + 1;
}

// y = M x with M in CSR form (row n: columns colidx[rowptr[n]..], values csrval).
// Eight lanes per row: lane j of a row's group takes nonzeros j, j + 8, ...
// (coalesced 128-byte reads of the row's values), then a 3-step xor reduction
// inside the group.  blockIdx.y selects one of several right-hand sides.
constexpr int kSpmvLanes = 8;

__global__ __launch_bounds__(256) void k_spmv(i64 n, const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                              const double2* __restrict__ val, const double2* __restrict__ x,
                                              double2* __restrict__ y) {
  x += (i64)blockIdx.y * n;
  y += (i64)blockIdx.y * n;
  const i64 r = (blockIdx.x * (i64)blockDim.x + threadIdx.x) / kSpmvLanes;
  const int j = threadIdx.x % kSpmvLanes;
  double ar = 0, ai = 0;
  if (r < n) {
    const int p1 = rowptr[r + 1];
#pragma unroll 4
    for (int p = rowptr[r] + j; p < p1; p += kSpmvLanes) {
      double2 a = val[p], v = x[colidx[p]];
      ar += a.x * v.x - a.y * v.y;
      ai += a.x * v.y + a.y * v.x;
    }
  }
#pragma unroll
  for (int o = kSpmvLanes / 2; o > 0; o >>= 1) {
    ar += __shfl_xor(ar, o, kSpmvLanes);
    ai += __shfl_xor(ai, o, kSpmvLanes);
  }
  if (r < n && j == 0) y[r] = make_double2(ar, ai);
}

// Y[:, c] = M X[:, c] for kSpmmCols columns per workgroup row (blockIdx.y
// selects the column group): the row's values and column indices are read once
// per group instead of once per column.  Lanes take nonzeros as in k_spmv; the
// 8-lane reduction is a reduce-scatter (offsets 4, 2, 1, each lane keeping half
// of its remaining columns), which adds the same partial sums in the same
// order as k_spmv's butterfly and leaves lane j with columns 2j and 2j + 1.
constexpr int kSpmmCols = 2 * kSpmvLanes;

__global__ __launch_bounds__(256) void k_spmm(i64 n, int R, const int* __restrict__ rowptr,
                                              const int* __restrict__ colidx, const double2* __restrict__ val,
                                              const double2* __restrict__ X, double2* __restrict__ Y) {
  const int c0 = blockIdx.y * kSpmmCols;
  const int nc = min(kSpmmCols, R - c0);
  const i64 r = (blockIdx.x * (i64)blockDim.x + threadIdx.x) / kSpmvLanes;
  const int j = threadIdx.x % kSpmvLanes;
  const double2* xb = X + (i64)c0 * n;
  double ar[kSpmmCols], ai[kSpmmCols];
#pragma unroll
  for (int c = 0; c < kSpmmCols; ++c) ar[c] = ai[c] = 0.0;
  if (r < n) {
    const int p1 = rowptr[r + 1];
    for (int p = rowptr[r] + j; p < p1; p += kSpmvLanes) {
      const double2 a = val[p];
      const double2* xc = xb + colidx[p];
#pragma unroll
      for (int c = 0; c < kSpmmCols; ++c) {
        if (c < nc) {
          const double2 v = xc[(i64)c * n];
          ar[c] += a.x * v.x - a.y * v.y;
          ai[c] += a.x * v.y + a.y * v.x;
        }
      }
    }
  }
#pragma unroll
  for (int o = kSpmvLanes / 2, m = kSpmmCols; o > 0; o >>= 1, m >>= 1) {
    const bool hi = (j & o) != 0;
#pragma unroll
    for (int i = 0; i < m / 2; ++i) {
      const double sr = hi ? ar[i] : ar[i + m / 2], si = hi ? ai[i] : ai[i + m / 2];
      const double kr = hi ? ar[i + m / 2] : ar[i], ki = hi ? ai[i + m / 2] : ai[i];
      ar[i] = kr + __shfl_xor(sr, o, kSpmvLanes);
      ai[i] = ki + __shfl_xor(si, o, kSpmvLanes);
    }
  }
  if (r < n) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int c = 2 * j + i;
      if (c < nc) Y[(i64)(c0 + c) * n + r] = make_double2(ar[i], ai[i]);
    }
  }
}

__global__ void k_gather_vals(i64 nnz, const int* __restrict__ map, const double2* __restrict__ src,
                              double2* __restrict__ dst) {
  i64 p = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (p < nnz) dst[p] = src[map[p]];
}

}  // namespace

namespace dmp {

int spai_max_cols() { return SPAI_MAXJ; }

void spai_build(i64 n, const ZRef& z, int maxJ, const int* jptr, const int* jidx, const int* iptr, const int* iidx,
                double* vals, int* status, hipStream_t s) {
  DM_REQUIRE(maxJ >= 1 && maxJ <= SPAI_MAXJ, DM_ERR_ARGUMENT, "near-field preconditioner: support width out of range");
  const size_t smem = sizeof(double2) * ((size_t)maxJ * (maxJ + 1) / 2 + (size_t)kAchLd * maxJ + maxJ) +
                      sizeof(int) * maxJ;
  HIP_OK(hipFuncSetAttribute((const void*)k_spai_build, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  DM_REQUIRE(n < (1LL << 31) - 8, DM_ERR_ARGUMENT, "near-field preconditioner: too many bases for one launch");
  const int per = (int)((n + 7) / 8);
  k_spai_build<<<8u * (unsigned)per, kSpaiThreads, smem, s>>>((int)n, per, z, maxJ, jptr, jidx, iptr, iidx,
                                                              reinterpret_cast<double2*>(vals), status);
  HIP_OK(hipGetLastError());
}

void csr_from_csc(i64 nnz, const int* map, const double* cscval, double* csrval, hipStream_t s) {
  k_gather_vals<<<(unsigned)((nnz + 255) / 256), 256, 0, s>>>(nnz, map, reinterpret_cast<const double2*>(cscval),
                                                               reinterpret_cast<double2*>(csrval));
  HIP_OK(hipGetLastError());
}

void spmv(i64 n, const int* rowptr, const int* colidx, const double* val, const double* x, double* y, hipStream_t s) {
  k_spmv<<<(unsigned)((n * kSpmvLanes + 255) / 256), 256, 0, s>>>(
      n, rowptr, colidx, reinterpret_cast<const double2*>(val), reinterpret_cast<const double2*>(x),
      reinterpret_cast<double2*>(y));
  HIP_OK(hipGetLastError());
}

void spmm(i64 n, int R, const int* rowptr, const int* colidx, const double* val, const double* X, double* Y,
          hipStream_t s) {
  if (R <= 0) return;
  DM_REQUIRE(R <= 65535, DM_ERR_ARGUMENT, "too many right-hand sides for one sparse product");
  dim3 g((unsigned)((n * kSpmvLanes + 255) / 256), (unsigned)((R + kSpmmCols - 1) / kSpmmCols));
  k_spmm<<<g, 256, 0, s>>>(n, R, rowptr, colidx, reinterpret_cast<const double2*>(val),
                           reinterpret_cast<const double2*>(X), reinterpret_cast<double2*>(Y));
  HIP_OK(hipGetLastError());
}

}  // namespace dmp
