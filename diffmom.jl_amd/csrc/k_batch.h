// k_batch.h — launchers of the lockstep-batch Arnoldi kernels (k_batch.hip).
// Blocks are n x R complex128 (column r at r * n); bases are sequences of
// blocks (block i at i * n * R); h / y arrays are R x kcap complex.
#pragma once
#include "dm_common.h"

namespace dmb {
int max_basis();               // largest k supported by dots / update / combine
int dot_chunks(i64 n);         // row chunks of dots: part holds R * kcap * dot_chunks(n) complex
// h[r, i] = V_i[r]^H W[r] for i < k
void dots(i64 n, int R, int k, int kcap, const double* V, const double* W, const int* act, double* part, double* h,
          hipStream_t s);
// W[r] -= sum_{i<k} V_i[r] h[r, i]
void update(i64 n, int R, int k, int kcap, const double* V, const double* h, const int* act, double* W,
            hipStream_t s);
// X[r] += sum_{i<cnt[r]} V_i[r] y[r, i]
void combine(i64 n, int R, int kcap, const double* V, const double* y, const int* cnt, double* X, hipStream_t s);
// out[r] = ||W[r]||^2 (act may be null: every column)
void norm2(i64 n, int R, const double* W, const int* act, double* out, hipStream_t s);
// dst[r] = sc[r] * src[r]
void scale(i64 n, int R, const double* src, const double* sc, double* dst, hipStream_t s);
// Device Givens steps of R columns with bases of up to m vectors: init (g_r =
// beta_r e_1 for active columns, sc = 1 / beta, counters cleared), step j
// (deactivates converged columns; est_out, if given: each active column's
// residual estimate |g[j+1]|), back substitution into y (pitch kcap).
void givens_init(int R, int m, const double* beta, const int* act, double* g, double* sc, int* jcnt, int* its,
                 hipStream_t s);
void givens(int R, int j, int m, int kcap, const double* h1, const double* h2, const double* nr, double* H,
            double* cs, double* sn, double* g, double* sc, int* act, const double* eps, int* jcnt, int* its,
            int* solved, hipStream_t s, double* est_out = nullptr);
void backsolve(int R, int m, int kcap, const double* H, const double* g, const int* jcnt, double* y, hipStream_t s);
}  // namespace dmb
