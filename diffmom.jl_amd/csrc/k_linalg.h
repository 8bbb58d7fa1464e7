// k_linalg.h — launchers of the Krylov / preconditioner / LU kernels.
#pragma once
#include <vector>

#include "dm_common.h"

namespace dml {
// One CGS pass: dots writes the row-chunk partials of V_i^H w (i < k) to part
// (dots_part_count(n, k) complex); update adds them into h[i] and sets
// w -= V h.  With nrm (norm_part_count(n) doubles) update also leaves the
// per-workgroup partials of ||w||^2, which givens adds up.
size_t dots_part_count(i64 n, int k);
int norm_part_count(i64 n);
void dots(i64 n, i64 ldv, int k, const double* V, const double* w, double* part, hipStream_t s);
void update(i64 n, i64 ldv, int k, const double* V, const double* part, double* h, double* w, double* nrm,
            hipStream_t s);
void combine(i64 n, i64 ldv, int k, const double* V, const double* y, double* x, hipStream_t s);
void norm2(i64 n, const double* w, double* out, hipStream_t s);
void axpby(i64 n, double ar, double ai, const double* x, double br, double bi, double* y, hipStream_t s);
void extract_blocks(i64 n, const ZRef& z, int bs, int nb, double* blocks, hipStream_t s);
void block_inverse(int bs, int nb, double* blocks, int* status, hipStream_t s);
void block_apply(i64 n, int bs, int nb, const double* inv, const double* x, double* y, hipStream_t s);
void diag_inv(i64 n, const ZRef& z, double* d, hipStream_t s);
void diag_apply(i64 n, const double* d, const double* x, double* y, hipStream_t s);
void lu_factor(i64 n, i64 ld, double* A, int* ipiv, int* status, hipStream_t s);
void lu_solve(i64 n, i64 ld, const double* A, const int* ipiv, double* B, i64 ldb, int nrhs, hipStream_t s);
// entries of one stored form; rows / columns of out in caller order via iofe
void unpermute(i64 n, const ZRef& z, i64 c0, i64 nc, const int* iofe, double* out, hipStream_t s);
void entries(const ZRef& z, i64 count, const int* iofe, const i64* m, const i64* n, double* out, hipStream_t s);
void conj(i64 n, double* v, hipStream_t s);
// GMRES on the device: Givens step of Arnoldi column j (hbuf: the two CGS
// passes' h; nrm: nb partials of ||w||^2 from update; vnext = w / ||w||, n
// rows; H: ldh x m complex, column-major; cs, sn, g: rotations and rotated
// right-hand side; st[2j], st[2j+1] = residual estimate, subdiagonal; sc[0] =
// 1 / subdiagonal), triangular solve y = R^{-1} g of the leading j columns.
void givens(int j, int ldh, const double* hbuf, const double* nrm, int nb, i64 n, const double* w, double* vnext,
            double* H, double* cs, double* sn, double* g, double* st, double* sc, hipStream_t s);
void backsolve(int j, int ldh, const double* H, const double* g, double* y, hipStream_t s);
// The same steps for a captured (replayable) GMRES step: the step index j is
// read from device memory (*jd), the grids are sized for k = j + 1 <= kmax
// (givens: j <= jmax), givens writes v_{j+1} into column j + 1 of V and into
// vcur, copy_col_dev writes src into column *jd of base, step_inc adds 1 to *jd.
void dots_dev(i64 n, i64 ldv, int kmax, const int* jd, const double* V, const double* w, double* part,
              hipStream_t s);
void update_dev(i64 n, i64 ldv, const int* jd, int hsel, const double* V, const double* part, double* h, double* w,
                double* nrm, hipStream_t s);
void givens_dev(int jmax, const int* jd, int ldh, const double* hbuf, const double* nrm, int nb, i64 n,
                const double* w, double* V, i64 ldv, double* vcur, double* H, double* cs, double* sn, double* g,
                double* st, double* sc, hipStream_t s);
void copy_col_dev(i64 n, const double* src, double* base, i64 ldv, const int* jd, hipStream_t s);
void step_inc(int* jd, hipStream_t s);
// y = Z x from packed symmetric storage; ws >= symv_workspace_bytes(n)
size_t symv_workspace_bytes(i64 n);
// shape: 2..64 = LDS-staged form with that many tiles per workgroup segment;
// 1000 + t (t = 1..1024) = wave form with t stored tiles per wave (chosen by
// the tuner)
void symv_packed(i64 n, const ZRef& z, const double* x, double* y, void* ws, int shape, hipStream_t s);
// launch shapes worth timing for this size and precision
std::vector<int> symv_shapes(i64 n, bool f32);
// Compressed packed complex64 form: the cross approximation of the
// off-diagonal tiles to ||Z - Z_H||_F <= eps ||Z||_F (rank[t] = -1: dense;
// lvrank (kHLevels x T, may be null): the ranks at the looser levels; fac:
// per tile U then V, rmax columns of 128 complex64 each; work: T + 1 doubles)
// and its symmetric product (workspace: symv_workspace_bytes(n)).
void aca_compress(i64 n, const float* z32p, double eps, int rmax, double* work, int* rank, int* lvrank, float* fac,
                  hipStream_t s);
// Wave ranges of the compressed product: W waves, wstart[W + 1] tile
// boundaries carrying about equal bytes, rend[T] = 1 where a column run ends.
struct HierRanges {
  int W = 0;
  int* wstart = nullptr;
  unsigned char* rend = nullptr;
};
size_t hier_ranges_bytes(i64 n, int W);
int hier_wave_slots();  // resident waves of the product kernel on this device
void hier_ranges(i64 n, const int* rank, int W, void* buf, HierRanges& hr, hipStream_t s);
// The same for nl rank arrays at once (buffers of hier_ranges_bytes(n, W) each).
void hier_ranges_levels(i64 n, int nl, const int* const* ranks, int W, void* const* bufs, HierRanges* hr,
                        hipStream_t s);
// f32part: the row and column partial vectors in FP32 (half the partial
// bytes; rounding ~3e-8 relative per partial, for the compression's looser
// levels, whose own error bound is 5e-7 and more).
void symv_hier(i64 n, const float* z32p, const int* rank, const float* fac, int rmax, const double* x, double* y,
               void* ws, const HierRanges& hr, hipStream_t s, bool f32part = false);
// Workspace of either batched product (complex64 or complex128) for R columns.
size_t zx_workspace_bytes(i64 n, int R);
// The complex64 Z X on the bf16 matrix cores at FP32 accuracy (each FP32
// value split into three bf16 parts, six part products): planes of Z32 built
// once by zx_bf16_split (zx_bf16_planes_elems(ld, n) bf16), workspace
// zx_bf16_workspace_bytes(n, R) per product.
size_t zx_bf16_planes_elems(i64 ld, i64 n);
void zx_bf16_split(i64 n, i64 ld, const float* z32, void* planes, hipStream_t s);
size_t zx_bf16_workspace_bytes(i64 n, int R);
// kc: columns per FP32 accumulation chunk (0: 4096; one of zx_bf16_chunks());
// nterms: 6 (FP32 accuracy) or 3 (relative error ~1e-5, relaxed products).
std::vector<int> zx_bf16_chunks();
void zx_bf16(i64 n, i64 ld, const void* planes, const double* X, i64 ldx, int R, double* Y, i64 ldy, void* ws,
             hipStream_t s, int kc = 0, int nterms = 6);
// Y (mr x R) = A (mr x n, lda) X (n x R) (subtract: Y -= A X), complex128 on
// FP64 MFMA; ws: gemm_c128_workspace_bytes(n, R).
size_t gemm_c128_workspace_bytes(i64 n, int R);
void gemm_c128(i64 mr, i64 n, const double* A, i64 lda, const double* X, i64 ldx, int R, double* Y, i64 ldy,
               void* ws, hipStream_t s, bool subtract = false);
void zx_mfma64(i64 n, i64 ld, const double* z64, const double* X, i64 ldx, int R, double* Y, i64 ldy, void* ws,
               hipStream_t s);
}  // namespace dml

namespace dmf {
// Plane-wave RHS for M waves into V (n x M, ldv), entry of internal basis i
// written at row out_idx[i] (identity or caller order).
void planewave(const dm_rwg& r, const TriTables& t, const double* kvec, const double* pol, const double* E0, int M,
               const int* out_idx, double* V, i64 ldv, hipStream_t s);
void radiation_vectors(const dm_rwg& r, const TriTables& t, const double* rhat, i64 nomega, double k, double eta0,
                       double* G, hipStream_t s);
}  // namespace dmf
