/*
 * diffmom.h — C ABI of libdiffmom, the MI355X (gfx950) implementation of
 * DiffMoM.jl's surface-EFIE path (RWG assembly, EFIE products, dense / GMRES
 * solves, plane-wave excitation, radiation vectors).
 *
 * The reference exposes no FFI: its "operator API" is Julia's AbstractMatrix
 * protocol plus the functions cited below.  Each entry point names the
 * reference function it replaces (path:line in DiffMoM.jl).  A Julia host binds
 * these with `ccall` (one `ccall` per function).
 *
 * Conventions (identical to Julia's in-memory layout, so Julia arrays pass
 * through unchanged):
 *   - indices are 1-based int64 (vertex ids in `tri`, triangle ids in RWG arrays)
 *     except where a function says 0-based;
 *   - matrices are column-major; `xyz` is 3 x Nv, `tri` 3 x Nt;
 *   - complex numbers are interleaved (re, im) doubles = Julia ComplexF64
 *     = C `double _Complex`;
 *   - host buffers are only read or written during the call, never retained.
 * Every function returns 0 on success or a DM_ERR_* code; `dm_last_error()`
 * (thread-local) gives the message.  Codes map one-to-one onto the reference's
 * exception classes so a Julia shim can rethrow the same type.
 * Thread safety: calls on one context are serialised by a per-context lock;
 * distinct contexts run concurrently.  Handles from one context must not be
 * used with another.
 */
#ifndef DIFFMOM_H
#define DIFFMOM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DM_OK 0
#define DM_ERR_ARGUMENT 1    /* Julia ArgumentError */
#define DM_ERR_OVERFLOW 2    /* Julia OverflowError */
#define DM_ERR_DIMENSION 3   /* Julia DimensionMismatch */
#define DM_ERR_CONVERGENCE 4 /* Julia ErrorException: iterative solve did not converge */
#define DM_ERR_DEVICE 5      /* HIP runtime failure */
#define DM_ERR_MEMORY 6      /* device allocation failed */
#define DM_ERR_DOMAIN 7      /* Julia DomainError (e.g. zero-length RWG edge) */

/* Which stored forms of Z an assembly produces (bit mask). */
#define DM_Z64 1 /* complex128 from FP64 quadrature (the reference's arithmetic) */
#define DM_Z32 2 /* complex64 from FP32 far-pair quadrature + FP64 self/adjacent terms */
/* The same two matrices in packed symmetric storage: only the upper triangle of
 * 128 x 128 tiles is kept (Z = Z^T exactly), half the bytes of the full forms;
 * products stream each stored tile once and use it for Z x and Z^T x. */
#define DM_Z64P 4
#define DM_Z32P 8
/* complex64 packed storage whose off-diagonal tiles may be cross-approximated
 * (dm_zmat_compress); usable for products and solves, not for entry reads */
#define DM_Z32H 16
/* dm_assemble_z flag: the complex64 forms are the rounding of the FP64
 * assembly (one launch writes both) instead of an FP32-quadrature assembly;
 * needs the complex128 form of the same storage in the mask.  The rounded
 * matrix carries only FP32 rounding noise, so it compresses much better. */
#define DM_Z32_ROUNDED 32

typedef struct dm_ctx dm_ctx;
typedef struct dm_mesh dm_mesh;
typedef struct dm_rwg dm_rwg;
typedef struct dm_zmat dm_zmat;
typedef struct dm_mfop dm_mfop;

/* ------------------------------------------------------------------ context */
/* A context keeps device blocks released by its calls for reuse by later
 * calls of the same size (at most min(64 GB, a quarter of the device));
 * dm_destroy frees them. */
int dm_create(int device, dm_ctx** out);
int dm_destroy(dm_ctx* ctx);
const char* dm_last_error(void);
int dm_version(void);
/* Device synchronisation (for host timing around asynchronous work). */
int dm_sync(dm_ctx* ctx);

/* Kernel timing on the library's own streams (no reference counterpart; for
 * benchmarks): tracked launches ("assemble_z64", "assemble_z32",
 * "near_tables", "gemv_z64", "symv_z32h", "arnoldi", "precond", ...) are
 * bracketed with HIP events; dm_profile_read sums the elapsed times of one
 * name.  on: 0 off, 1 every tracked launch, 2 only the operator products
 * (names "symv*" and "gemv*"; fewer events in a timed solve). */
int dm_profile_enable(dm_ctx* ctx, int on);
int dm_profile_reset(dm_ctx* ctx);
int dm_profile_read(dm_ctx* ctx, const char* name, int64_t* count, double* total_ms);

/* --------------------------------------------------------------------- mesh */
/* TriMesh(xyz, tri): src/Types.jl:123-129.  Copies the arrays. */
int dm_mesh_create(dm_ctx* ctx, const double* xyz, int64_t nv, const int64_t* tri, int64_t nt,
                   dm_mesh** out);
int dm_mesh_free(dm_mesh* mesh);
/* assert_mesh_quality: src/geometry/Mesh.jl:1093-1255 (DM_ERR_ARGUMENT with the
 * report in the message when the mesh fails). */
int dm_mesh_check(const dm_mesh* mesh, int allow_boundary, int require_closed,
                  double area_tol_rel);
/* triangle_area for all triangles: src/geometry/Mesh.jl:3205-3232 */
int dm_mesh_areas(const dm_mesh* mesh, double* areas);

/* ---------------------------------------------------------------------- RWG */
/* build_rwg(mesh; precheck, allow_boundary, require_closed, area_tol_rel):
 * src/basis/RWG.jl:102-143.  T+ is the lower-index triangle of each interior
 * edge.  Basis order: interior edges sorted by their (v1 < v2) vertex key (the
 * reference's order is Julia Dict iteration order, which no other language
 * reproduces; callers holding a Julia RWGData use dm_rwg_create instead). */
int dm_build_rwg(const dm_mesh* mesh, int precheck, int allow_boundary, int require_closed,
                 double area_tol_rel, dm_rwg** out);
/* RWGData as the caller holds it (src/Types.jl:138-152), any basis order.
 * evert may be NULL; coeff_plus / coeff_minus may be NULL (unit coefficients). */
int dm_rwg_create(const dm_mesh* mesh, int64_t n, const int64_t* tplus, const int64_t* tminus,
                  const int64_t* evert, const int64_t* vplus_opp, const int64_t* vminus_opp,
                  const double* len, const double* area_plus, const double* area_minus,
                  const double* coeff_plus, const double* coeff_minus, dm_rwg** out);
int dm_rwg_size(const dm_rwg* rwg, int64_t* n);
/* Copies the basis arrays back (evert is 2 x n). */
int dm_rwg_export(const dm_rwg* rwg, int64_t* tplus, int64_t* tminus, int64_t* evert,
                  int64_t* vplus_opp, int64_t* vminus_opp, double* len, double* area_plus,
                  double* area_minus);
int dm_rwg_free(dm_rwg* rwg);
/* Number of edge-adjacent triangle pairs (_build_triangle_adjacency,
 * src/assembly/EFIE.jl:140-231). */
int dm_rwg_adjacent_pairs(const dm_rwg* rwg, int64_t* count);

/* ----------------------------------------------------------------- assembly */
/* assemble_Z_efie(mesh, rwg, k; quad_order, eta0, max_output_bytes):
 * src/assembly/EFIE.jl:699-753.  The matrix stays on the device.  `forms` is a
 * mask of DM_Z64, DM_Z32, DM_Z64P, DM_Z32P.  k may be complex (k_im != 0).  max_output_bytes bounds
 * N*N*16 as the reference's byte gate does (<= 0: no gate). */
int dm_assemble_z(dm_ctx* ctx, const dm_rwg* rwg, double k_re, double k_im, double eta0,
                  int quad_order, int forms, int64_t max_output_bytes, dm_zmat** out);
/* Many meshes in one persistent launch per stored form (no reference
 * counterpart; batches of small scatterers): every workgroup pulls the next
 * tile from one queue over all meshes' tiles, largest meshes first.  outs[m]
 * receives mesh m's matrix, exactly what dm_assemble_z would produce. */
int dm_assemble_z_multi(dm_ctx* ctx, int64_t nmesh, const dm_rwg* const* rwgs, double k_re, double k_im,
                        double eta0, int quad_order, int forms, dm_zmat** outs);
int dm_zmat_size(const dm_zmat* z, int64_t* n, int* forms);
/* Copy Z (caller basis order, column-major N x N, complex128) to host.
 * form = one stored form (complex64 values widened). */
int dm_zmat_get(const dm_zmat* z, int form, double* out);
/* Entries Z[m_i, n_i] (0-based caller indices) of one stored form. */
int dm_zmat_entries(const dm_zmat* z, int form, const int64_t* m, const int64_t* n,
                    int64_t count, double* out);
int dm_zmat_free(dm_zmat* z);

/* Y = op(Z) X for nrhs columns (mul!(y, Z, x) / zgemv inside Krylov.gmres,
 * src/solver/IterativeSolve.jl:364-384).  op = Z (adjoint = 0) or Z^H (1).
 * Host buffers, caller basis order, complex128; form as in dm_zmat_get. */
int dm_apply(dm_ctx* ctx, const dm_zmat* z, int form, int adjoint, const double* X, int64_t ldx,
             int64_t nrhs, double* Y, int64_t ldy);

/* No reference counterpart: an allocation-free form of dm_apply for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  y = op(Z) x for one
 * right-hand side; x and y are DEVICE pointers to N complex128 values in the
 * caller's basis order (x == y allowed).  form: DM_Z64, DM_Z32 (full storage)
 * or DM_Z64P, DM_Z32P (packed, with the launch shape the tuner chose for this
 * size or the default; never timed here).  The work is queued on the context's
 * stream and the call returns without waiting (dm_sync waits).  It calls no
 * hipMalloc / hipFree: its intermediates live in one 16 MiB device array per
 * GPU, set up by dm_create, so N is limited to what fits (full storage:
 * 32 N bytes, N <= 524 288; packed: 32 N bytes plus the symmetric product's
 * partials, N <= 11 392); larger N is refused with DM_ERR_ARGUMENT.  Every call
 * on one GPU shares that array: calls that overlap in time on one GPU (two
 * contexts' streams, or two host threads with different contexts) race and
 * must be ordered by the caller. */
int dm_apply_noalloc(dm_ctx* ctx, const dm_zmat* z, int form, int adjoint, const double* x, double* y);

/* ------------------------------------------------------- matrix-free operator */
/* matrixfree_efie_operator / mul! / adjoint: src/assembly/EFIE.jl:759-1007.
 * Each product recomputes every entry on the device (FP64), never storing Z. */
int dm_mf_create(dm_ctx* ctx, const dm_rwg* rwg, double k_re, double k_im, double eta0,
                 int quad_order, dm_mfop** out);
int dm_mf_apply(dm_ctx* ctx, const dm_mfop* op, int adjoint, const double* x, double* y);
int dm_mf_free(dm_mfop* op);

/* --------------------------------------------------------------- excitation */
/* assemble_plane_wave / assemble_multiple_excitations for plane waves:
 * src/assembly/Excitation.jl:2738-2797,2866-2896.  kvec, pol: 3 x M, E0: M.
 * Output V: N x M complex128 column-major (caller basis order).  Validates
 * |pol| = 1 and pol . k = 0 as src/assembly/Excitation.jl:1549-1598 does. */
int dm_excitation_planewave(dm_ctx* ctx, const dm_rwg* rwg, const double* kvec, const double* pol,
                            const double* E0, int64_t M, int quad_order, double* V);

/* ------------------------------------------------------------------- solves */
#define DM_METHOD_LU 0        /* dense LU (solve_forward :direct, src/solver/Solve.jl:1124-1198) */
#define DM_METHOD_GMRES 1     /* FP64 GMRES on Z64 (solve_gmres, src/solver/IterativeSolve.jl:341-392) */
#define DM_METHOD_MIXED 2     /* FGMRES on Z32 with FP64 Arnoldi + FP64 iterative refinement on Z64 */

#define DM_PRECOND_NONE 0
#define DM_PRECOND_DIAG 1     /* Jacobi (factorization=:diag, src/solver/NearFieldPreconditioner.jl:623-668) */
#define DM_PRECOND_NEARFIELD 2 /* near-field sparse approximate inverse over the cutoff pattern */
#define DM_PRECOND_LEARNED 3  /* message-passing network, flexible (needs a trained model) */
#define DM_PRECOND_BLOCK 4    /* block Jacobi over spatially consecutive bases */

typedef struct dm_mpnn dm_mpnn; /* trained learned-preconditioner network */

typedef struct dm_solve_opts {
  int method;          /* DM_METHOD_* */
  int precond;         /* DM_PRECOND_* */
  int precond_side;    /* 0 = left, 1 = right (solve_gmres precond_side) */
  int maxiter;         /* total Krylov iterations allowed */
  double tol;          /* relative tolerance on the (true FP64) residual */
  double nf_cutoff;    /* near-field cutoff in metres (<= 0: 2.5 x mean edge length) */
  int restart;         /* Krylov basis size before restart (0 = no restart) */
  int check_convergence; /* 1: DM_ERR_CONVERGENCE when not converged */
  double inner_tol;    /* MIXED: relative tolerance of the first inner solve (<= 0: tol/2) */
  int max_outer;       /* MIXED: refinement steps allowed */
  int block_size;      /* DM_PRECOND_BLOCK: bases per spatial block (0: 128) */
  int batch;           /* several right-hand sides: 0 = lockstep batches when nrhs >= 8,
                          1 = always, -1 = one right-hand side at a time */
  const dm_mpnn* learned; /* DM_PRECOND_LEARNED: the network (SPAI over nf_cutoff + network) */
  int adjoint;         /* dm_solve: 1 solves Z^H X = B (solve_gmres_adjoint,
                          src/solver/IterativeSolve.jl:409-460); Z = Z^T, so Z^H x = b
                          is solved as Z conj(x) = conj(b) */
  double compress_tol; /* MIXED, dm_solve_scattering_pw, one right-hand side at a time:
                          > 0 compresses the complex64 matrix (dm_zmat_compress with
                          this eps) for the inner solve; the refinement residual
                          stays the dense complex128 product.  0 = off */
  int storage;         /* dm_solve_scattering_pw, GMRES one right-hand side at a time:
                          0 = packed symmetric complex128 storage (half the bytes per
                          product), 1 = full N x N storage (the reference's dense
                          Matrix{ComplexF64}).  Batches and dense LU always use full
                          storage; MIXED always packs its one-column forms. */
  int graph;           /* one right-hand side at a time: 1 captures one Krylov step
                          (preconditioner, product, CGS2, Givens) in a HIP graph and
                          replays it, the step index read from device memory; 0 =
                          launches each kernel (default). */
  double relax;        /* MIXED: > 0 relaxes the inner products as the residual falls
                          (inexact Krylov): at each check point the product may
                          carry an error bound of at most relax * inner_tol *
                          ||r0|| / (residual estimate) -- with a compressed form
                          the compression's looser levels; on the dense complex64
                          form in lockstep batches, three instead of six bf16
                          part products of the Z X (bound 5e-5).  In a lockstep
                          batch the most demanding active column decides.  The
                          refinement residual stays the dense complex128
                          product.  0 = off (default). */
} dm_solve_opts;

typedef struct dm_solve_stats {
  int iterations;        /* Krylov iterations (sum over refinement steps) */
  int outer_iterations;  /* refinement steps (MIXED) */
  int solved;            /* 1 if the tolerance was met */
  double residual;       /* true FP64 relative residual ||b - Z64 x|| / ||b|| */
  double precond_time_s;
  double solve_time_s;
} dm_solve_stats;

void dm_solve_opts_default(dm_solve_opts* opts);

/* Preconditioner as an operator: build_nearfield_preconditioner +
 * NearFieldOperator's mul! (src/solver/NearFieldPreconditioner.jl:732-759,
 * 1034-1060).  kind: DM_PRECOND_NEARFIELD (sparse approximate inverse of the
 * near field within nf_cutoff metres, <= 0: 2.5 x mean edge), DM_PRECOND_DIAG or
 * DM_PRECOND_BLOCK; built from the stored form `form` of Z.  apply: y = M x
 * (host vectors, caller basis order). */
typedef struct dm_precond dm_precond;
int dm_precond_create(dm_ctx* ctx, const dm_zmat* z, int kind, double nf_cutoff, int block_size, int form,
                      dm_precond** out);
int dm_precond_apply(dm_ctx* ctx, const dm_precond* p, const double* x, double* y);
/* The near-field pattern the sparse approximate inverse uses for a cutoff
 * (nf_cutoff metres, <= 0: 2.5 x mean edge; rwg_centers and the cutoff
 * search, src/solver/NearFieldPreconditioner.jl:247-338): its nonzeros, the
 * mean radius its columns actually cover, the column cap (each column keeps
 * the nearest max_cols bases within the cutoff) and how many columns the cap
 * cut short of the cutoff.  A cutoff wider than the cap reaches is reported
 * here, not silently changed.  Any output pointer may be NULL. */
int dm_nearfield_pattern_info(const dm_rwg* rwg, double nf_cutoff, int64_t* nnz, double* mean_radius,
                              int* max_cols, int64_t* capped_cols);

int dm_precond_free(dm_precond* p);
/* Solve Z X = B for nrhs right-hand sides (host buffers, caller order). */
int dm_solve(dm_ctx* ctx, const dm_zmat* z, const double* B, int64_t ldb, int64_t nrhs,
             const dm_solve_opts* opts, double* X, int64_t ldx, dm_solve_stats* stats);

/* ------------------------------------------------------------ compression */
/* No reference counterpart as an entry point; the reference's ACA
 * (src/fast/ACA.jl:558-688) compresses far blocks from entries computed on
 * demand.  Here the assembled packed complex64 form (DM_Z32P, required) is
 * compressed tile by tile: every off-diagonal 128 x 128 tile whose cross
 * approximation reaches the per-tile bound within rmax terms (1..64) is stored
 * as U V^T, so that ||Z32P - Z32H||_F <= eps sqrt(D) <= eps ||Z32P||_F with
 * D the diagonal tiles' share of ||Z32P||_F^2 (plus the FP32 rounding of the
 * residual updates, a few 1e-8 ||Z||_F).  Adds DM_Z32H to the
 * forms.  stats (may be NULL): [0] stored tiles, [1] low-rank tiles, [2] mean
 * rank of the low-rank tiles, [3] bytes a product reads relative to DM_Z32P. */
int dm_zmat_compress(dm_zmat* z, double eps, int rmax, double* stats);
/* No reference counterpart.  The compression's looser levels: the steps of
 * one cross approximation are nested, so level l (1..10) reuses the prefix of
 * each tile's factors at which its residual fell below 10^(l/2) times the
 * tile tolerance (error bound eps 10^(l/2); some tiles dense at eps are
 * low-rank there).  Sets the level that dm_apply products on DM_Z32H use (0 = eps) and,
 * with stats (5 doubles, may be NULL), reports it: the four dm_zmat_compress
 * statistics at that level, then its error bound eps 10^(l/2).  Solves choose
 * their own levels (dm_solve_opts.relax). */
int dm_zmat_hlevel(dm_zmat* z, int level, double* stats);

/* ------------------------------------------------------------------ tuning */
/* No reference counterpart.  Launch shapes (the symmetric product's tiles per
 * workgroup, ...) are timed on first use per shape and kept in the context.
 * dm_tune_get writes them as a flat JSON object {"key": value, ...} (len: its
 * length; buf may be NULL to query it); dm_tune_set loads such an object, so a
 * caller can persist the choices across processes.  DM_TUNE=0 in the
 * environment skips timing and uses defaults. */
int dm_tune_get(dm_ctx* ctx, char* buf, int64_t cap, int64_t* len);
int dm_tune_set(dm_ctx* ctx, const char* json);

/* ------------------------------------------------- learned preconditioner */
/* No reference counterpart: a 4-layer message-passing network over the RWG
 * edge-adjacency graph (bases sharing a triangle), hidden width 64, bfloat16
 * weights with FP32 accumulation, that corrects the near-field SPAI S inside
 * FGMRES: M r = S (r + d(S r, r)) with d the network's correction
 * (opts.precond = DM_PRECOND_LEARNED, opts.learned = the network).  Trained
 * on the GPU by the library's own forward / backward kernels and Adam on the
 * relative residual one application of M leaves, ||Z M r - r||^2 / ||r||^2,
 * over residuals of the library's GMRES iterates on seeded meshes. */
int64_t dm_mpnn_param_count(void);
int dm_mpnn_create(dm_ctx* ctx, const float* params, dm_mpnn** out);
int dm_mpnn_get_params(const dm_mpnn* net, float* out);
int dm_mpnn_free(dm_mpnn* net);
/* The network's graph and geometry features of one mesh in caller order:
 * nbr (N x 4, 0-based, -1 padded), sgn (N x 4), geom (N x 2). */
int dm_rwg_graph(const dm_rwg* rwg, double k, int32_t* nbr, float* sgn, float* geom);
/* Training samples resident on the device: nsample meshes' nodes concatenated
 * (sample s owns nodes sample_ptr[s] .. sample_ptr[s+1]-1; nbr holds global
 * node ids), residual r and z = S r (complex128, N). */
typedef struct dm_mpnn_data dm_mpnn_data;
int dm_mpnn_data_create(dm_ctx* ctx, int64_t nnode, const int32_t* nbr, const float* sgn, const float* geom,
                        const double* z, const double* r, int64_t nsample, const int64_t* sample_ptr,
                        dm_mpnn_data** out);
int dm_mpnn_data_free(dm_mpnn_data* data);
/* The operator A = Z S of one mesh in caller order (N x N complex64,
 * column-major, host memory): S the preconditioner p built on z (which must
 * hold DM_Z64), N <= 16384. */
int dm_mpnn_residual_operator(dm_ctx* ctx, const dm_zmat* z, const dm_precond* p, float* out);
/* The operator A (N x N complex64, caller order, host memory) of samples
 * [s0, s1), all of one mesh with N nodes: the device keeps a copy and
 * p = r - A r for each of those samples.  Every sample needs one before
 * dm_mpnn_loss_grad or dm_mpnn_train. */
int dm_mpnn_data_set_matrix(dm_mpnn_data* data, int64_t s0, int64_t s1, const float* a, int64_t n);
/* Mean objective over samples [s0, s1) and its gradient w.r.t. the
 * bfloat16-rounded parameters (grad may be NULL). */
int dm_mpnn_loss_grad(dm_ctx* ctx, dm_mpnn* net, const dm_mpnn_data* data, int64_t s0, int64_t s1,
                      double* loss, float* grad);
/* Adam training: `epochs` passes over mini-batches of batch_samples samples
 * in a seeded shuffled order; epoch_loss (may be NULL) receives the mean
 * objective of each epoch. */
int dm_mpnn_train(dm_ctx* ctx, dm_mpnn* net, const dm_mpnn_data* data, int epochs, int64_t batch_samples,
                  double lr, uint64_t seed, double* epoch_loss);
/* The network's correction d for one mesh (caller order, complex128, in the
 * units of r): the preconditioner returns z + S d. */
int dm_mpnn_apply(dm_ctx* ctx, const dm_mpnn* net, const dm_rwg* rwg, double k, const double* z,
                  const double* r, double* out);

/* solve_scattering(mesh, freq, excitation; method=:dense_direct|:dense_gmres)
 * for plane waves, end to end on the device (src/Workflow.jl:67-317): assembly,
 * excitation, preconditioner and solve of nrhs waves (kvec, pol: 3 x nrhs).
 * opts->method DM_METHOD_LU | DM_METHOD_GMRES | DM_METHOD_MIXED; with MIXED one
 * FP64 assembly writes the complex128 matrix (for the refinement residual) and
 * its complex64 rounding (for the inner solve), which opts->compress_tol > 0
 * compresses on a second stream beside the preconditioner build.  X: N x nrhs
 * complex128, caller basis order.  assembly_time_s (optional): wall time until
 * the matrix the solver iterates on is ready.  zout (optional): receives the
 * matrix; without it the RWG handle keeps the matrix buffers for its next
 * solve (freed by dm_rwg_free). */
int dm_solve_scattering_pw(dm_ctx* ctx, const dm_rwg* rwg, double k_re, double k_im, double eta0, int quad_order,
                           const double* kvec, const double* pol, const double* E0, int64_t nrhs,
                           const dm_solve_opts* opts, double* X, dm_solve_stats* stats, double* assembly_time_s,
                           dm_zmat** zout);

/* ---------------------------------------------------------------- far field */
/* radiation_vectors(mesh, rwg, grid, k; quad_order, eta0):
 * src/postprocessing/FarField.jl:425-574.  rhat: 3 x NΩ; G: (3 NΩ) x N. */
int dm_radiation_vectors(dm_ctx* ctx, const dm_rwg* rwg, const double* rhat, int64_t nomega,
                         double k, double eta0, int quad_order, double* G);

/* Far fields of many currents: E (3 NΩ x ncur, column-major, leading
 * dimension lde) = G I with G the radiation vectors above and I (N x ncur,
 * caller basis order, leading dimension ldi) -- compute_farfield
 * (src/postprocessing/FarField.jl:584-605) for a batch of solutions, without
 * G leaving the device: G is formed for chunks of directions and multiplied
 * on the FP64 matrix cores. */
int dm_farfield(dm_ctx* ctx, const dm_rwg* rwg, const double* rhat, int64_t nomega, double k, double eta0,
                int quad_order, const double* I, int64_t ldi, int64_t ncur, double* E, int64_t lde);

#ifdef __cplusplus
}
#endif
#endif /* DIFFMOM_H */
