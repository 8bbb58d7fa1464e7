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
constexpr int SPAI_CHUNK = 16;  // rows of A staged per pass

// One column per workgroup.  LDS holds, for the widest support mJ of the
// pattern (not the cap), the Gram matrix's lower triangle packed by rows
// (entry (a, b), b <= a, at a (a + 1) / 2 + b), SPAI_CHUNK staged rows of A,
// the right-hand side and the support's column ids: about 28 KB at mJ = 44,
// so several columns share a CU.
__device__ inline int tri(int a, int b) { return a * (a + 1) / 2 + b; }

__global__ __launch_bounds__(256) void k_spai_build(ZRef z, int mJ, const int* __restrict__ jptr,
                                                    const int* __restrict__ jidx, const int* __restrict__ iptr,
                                                    const int* __restrict__ iidx, double2* __restrict__ vals,
                                                    int* __restrict__ status) {
  const int j = blockIdx.x;
  const int j0 = jptr[j], nJ = jptr[j + 1] - j0;
  const int i0 = iptr[j], nI = iptr[j + 1] - i0;
  extern __shared__ __align__(16) char smem[];
  double2* G = reinterpret_cast<double2*>(smem);
  double2* Ach = G + mJ * (mJ + 1) / 2;
  double2* rhs = Ach + SPAI_CHUNK * mJ;
  int* cols = reinterpret_cast<int*>(rhs + mJ);
  __shared__ int bad;
  const int tid = threadIdx.x;
  if (tid == 0) bad = 0;
  for (int a = tid; a < nJ; a += 256) cols[a] = jidx[j0 + a];
  for (int e = tid; e < nJ * (nJ + 1) / 2; e += 256) G[e] = make_double2(0.0, 0.0);
  __syncthreads();
  // rhs = A^H e_j = conj(Z[j, J]) (row j of Z restricted to the pattern)
  for (int a = tid; a < nJ; a += 256) {
    double2 v = zat(z, j, cols[a]);
    rhs[a] = make_double2(v.x, -v.y);
  }
  // G = A^H A accumulated over row chunks (lower triangle incl. diagonal)
  const int ntri = nJ * (nJ + 1) / 2;
  for (int r0 = 0; r0 < nI; r0 += SPAI_CHUNK) {
    const int nr = min(SPAI_CHUNK, nI - r0);
    __syncthreads();
    for (int e = tid; e < nr * nJ; e += 256) {
      int ii = e / nJ, a = e - ii * nJ;
      Ach[ii * mJ + a] = zat(z, iidx[i0 + r0 + ii], cols[a]);
    }
    __syncthreads();
    for (int e = tid; e < ntri; e += 256) {
      // e = a (a + 1) / 2 + b
      int a = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
      while (a * (a + 1) / 2 > e) --a;
      while ((a + 1) * (a + 2) / 2 <= e) ++a;
      const int b = e - a * (a + 1) / 2;
      double sr = 0, si = 0;
      for (int ii = 0; ii < nr; ++ii) {
        double2 p = Ach[ii * mJ + a], q = Ach[ii * mJ + b];
        // conj(p) * q
        sr += p.x * q.x + p.y * q.y;
        si += p.x * q.y - p.y * q.x;
      }
      G[e].x += sr;
      G[e].y += si;
    }
  }
  __syncthreads();
  // Cholesky G = L L^H in place (lower triangle), right-looking
  for (int k = 0; k < nJ; ++k) {
    if (tid == 0) {
      double d = G[tri(k, k)].x;
      if (!(d > 0.0)) {
        bad = 1;
        d = 1e-300;
      }
      G[tri(k, k)] = make_double2(sqrt(d), 0.0);
    }
    __syncthreads();
    const double lkk = G[tri(k, k)].x;
    for (int i = k + 1 + tid; i < nJ; i += 256) {
      double2 v = G[tri(i, k)];
      G[tri(i, k)] = make_double2(v.x / lkk, v.y / lkk);
    }
    __syncthreads();
    const int m = nJ - k - 1;
    for (int e = tid; e < m * m; e += 256) {
      int i = k + 1 + e / m, l = k + 1 + e % m;
      if (l > i) continue;
      double2 a = G[tri(i, k)], b = G[tri(l, k)];
      // G[i][l] -= L[i][k] conj(L[l][k])
      G[tri(i, l)].x -= a.x * b.x + a.y * b.y;
      G[tri(i, l)].y -= a.y * b.x - a.x * b.y;
    }
    __syncthreads();
  }
  // forward: L y = rhs
  for (int k = 0; k < nJ; ++k) {
    if (tid == 0) {
      double d = G[tri(k, k)].x;
      rhs[k] = make_double2(rhs[k].x / d, rhs[k].y / d);
    }
    __syncthreads();
    double2 yk = rhs[k];
    for (int i = k + 1 + tid; i < nJ; i += 256) {
      double2 l = G[tri(i, k)];
      rhs[i].x -= l.x * yk.x - l.y * yk.y;
      rhs[i].y -= l.x * yk.y + l.y * yk.x;
    }
    __syncthreads();
  }
  // backward: L^H m = y
  for (int k = nJ - 1; k >= 0; --k) {
    if (tid == 0) {
      double d = G[tri(k, k)].x;
      rhs[k] = make_double2(rhs[k].x / d, rhs[k].y / d);
    }
    __syncthreads();
    double2 mk = rhs[k];
    for (int i = tid; i < k; i += 256) {
      double2 l = G[tri(k, i)];  // (L^H)[i][k] = conj(L[k][i])
      rhs[i].x -= l.x * mk.x + l.y * mk.y;
      rhs[i].y -= l.x * mk.y - l.y * mk.x;
    }
    __syncthreads();
  }
  for (int a = tid; a < nJ; a += 256) vals[j0 + a] = rhs[a];
  if (tid == 0 && bad) status[0] = j + 1;
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
  const size_t smem = sizeof(double2) * ((size_t)maxJ * (maxJ + 1) / 2 + (size_t)SPAI_CHUNK * maxJ + maxJ) +
                      sizeof(int) * maxJ;
  HIP_OK(hipFuncSetAttribute((const void*)k_spai_build, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  k_spai_build<<<(unsigned)n, 256, smem, s>>>(z, maxJ, jptr, jidx, iptr, iidx, reinterpret_cast<double2*>(vals),
                                              status);
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
