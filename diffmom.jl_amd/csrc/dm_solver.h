// dm_solver.h — device operators, preconditioners and the Krylov solvers behind
// dm_solve / dm_solve_scattering_pw: one right-hand side at a time (GEMV
// products) or a lockstep batch of right-hand sides (Z X products on MFMA).
#pragma once
#include <cmath>
#include <complex>
#include <functional>
#include <string>
#include <memory>
#include <vector>

#include "dm_common.h"
#include "k_linalg.h"

namespace dmt {
// The fastest of cands for key: timed once per context with run(c) on stream
// s (one warm-up and three timed runs each), then kept in the context and
// exported by dm_tune_get; DM_TUNE=0 returns fallback without timing.  run
// must leave the same results whichever candidate it is given.
int tune_choice(dm_ctx* ctx, const std::string& key, const std::vector<int>& cands, int fallback,
                const std::function<void(int)>& run, hipStream_t s);
}  // namespace dmt

namespace dms {

typedef std::complex<double> cd;

double now();

// Products switch from one GEMV per vector to the MFMA Z X kernel at this many
// right-hand sides; batches hold at most kMfmaMaxRhs columns (bounds the
// split-K workspace: 8 chunks x 256 x N complex128 at N = 30k).
constexpr i64 kMfmaMinRhs = 8;
constexpr i64 kMfmaMaxRhs = 256;

// ------------------------------------------------------------ device operators
// apply: y = A x (internal basis order, complex128 vectors of length n).
// apply_batch: Y = A X for R columns (leading dimension n); ws holds at least
// batch_ws(R) bytes.
struct DevOp {
  virtual ~DevOp() = default;
  // Relaxed products (inexact Krylov): later products may use a form whose
  // relative error bound is at most eta; returns the bound in use (operators
  // without looser forms ignore it and return 0).
  virtual double relax(double /*eta*/) { return 0.0; }
  virtual void apply(const double* x, double* y, hipStream_t s) = 0;
  virtual void apply_batch(const double* X, double* Y, int R, void* ws, hipStream_t s) = 0;
  virtual size_t batch_ws(int R) const = 0;
};
struct OpZ64 : DevOp {
  const dm_zmat* z;
  explicit OpZ64(const dm_zmat* z_) : z(z_) {}
  void apply(const double* x, double* y, hipStream_t s) override {
    ProfScope p(z->ctx, "gemv_z64", s);
    launch_gemv_z64(z->n, z->ld, z->z64.p, x, y, s);
  }
  void apply_batch(const double* X, double* Y, int R, void* ws, hipStream_t s) override {
    if (R < kMfmaMinRhs) {
      for (int j = 0; j < R; ++j) apply(X + 2 * z->n * j, Y + 2 * z->n * j, s);
      return;
    }
    ProfScope p(z->ctx, "zx_z64", s);
    dml::zx_mfma64(z->n, z->ld, z->z64.p, X, z->n, R, Y, z->n, ws, s);
  }
  size_t batch_ws(int R) const override { return dml::zx_workspace_bytes(z->n, R); }
};
// Error bound used for the three-part bf16 Z X (dropped part products below
// 2^-18 relative each, three of them; five times that for margin).
constexpr double kZx3TermBound = 5e-5;
struct OpZ32 : DevOp {
  const dm_zmat* z;
  DBuf<char> planes;  // bf16 x 3 planes of Z32, built on the first batched product
  int nterms = 6;     // bf16 part products of a batched product (3 once relaxed)
  explicit OpZ32(const dm_zmat* z_) : z(z_) {}
  double relax(double eta) override {
    nterms = eta >= kZx3TermBound ? 3 : 6;
    return nterms == 3 ? kZx3TermBound : 0.0;
  }
  void apply(const double* x, double* y, hipStream_t s) override {
    ProfScope p(z->ctx, "gemv_z32", s);
    launch_gemv_z32(z->n, z->ld, z->z32.p, x, y, s);
  }
  void apply_batch(const double* X, double* Y, int R, void* ws, hipStream_t s) override {
    if (R < kMfmaMinRhs) {
      for (int j = 0; j < R; ++j) apply(X + 2 * z->n * j, Y + 2 * z->n * j, s);
      return;
    }
    if (!planes.p) {
      ProfScope p(z->ctx, "zx_split", s);
      planes.alloc(dml::zx_bf16_planes_elems(z->ld, z->n) * 2);
      dml::zx_bf16_split(z->n, z->ld, z->z32.p, planes.p, s);
    }
    // FP32 accumulation chunk: key zx/c64/<n>/<R>
    const int kc = dmt::tune_choice(
        z->ctx, "zx/c64/" + std::to_string(z->n) + "/" + std::to_string(R), dml::zx_bf16_chunks(), 4096,
        [&](int c) { dml::zx_bf16(z->n, z->ld, planes.p, X, z->n, R, Y, z->n, ws, s, c); }, s);
    ProfScope p(z->ctx, nterms == 3 ? "zx_z32_relaxed" : "zx_z32", s);
    dml::zx_bf16(z->n, z->ld, planes.p, X, z->n, R, Y, z->n, ws, s, kc, nterms);
  }
  size_t batch_ws(int R) const override {
    return dml::zx_bf16_workspace_bytes(z->n, R);
  }
};

// Packed symmetric storage (DM_Z64P / DM_Z32P): one symmetric product per
// vector; batches loop over columns.
}  // namespace dms
namespace dmt {
// Tuned launch shape of the symmetric product (timed on first use per size).
int symv_shape(dm_ctx* ctx, const dm_zmat* z, int form, void* ws, hipStream_t s);
// The shape already chosen for this size, or the default; never times anything.
int symv_shape_known(const dm_ctx* ctx, const dm_zmat* z, int form);
// Tuned wave count of the compressed product, as a multiple of the resident
// wave slots (1..64).
int symv_hier_shape(dm_ctx* ctx, const dm_zmat* z, void* ws, hipStream_t s);
}  // namespace dmt
namespace dms {
struct OpHier : DevOp {
  const dm_zmat* z;
  DBuf<char> ws;
  // per level (0: the compression's eps, l: eps 10^l) the tiles' ranks and
  // the wave ranges, set up on the first product at that level
  DBuf<char> rbuf[1 + kHLevels];
  dml::HierRanges hr[1 + kHLevels];
  int level;
  explicit OpHier(const dm_zmat* z_) : z(z_), level(z_->hlevel) { ws.alloc(dml::symv_workspace_bytes(z->n)); }
  const int* ranks(int l) const { return l == 0 ? z->hrank.p : z->hrank_lv.p + (size_t)(l - 1) * packed_tiles(z->n); }
  double relax(double eta) override {
    int l = 0;
    while (l < kHLevels && z->heps * hlevel_factor(l + 1) <= eta) ++l;
    level = l;
    return z->heps * hlevel_factor(l);
  }
  void apply(const double* x, double* y, hipStream_t s) override {
    if (!hr[level].W) {
      // the wave ranges of every level in one batch of launches
      const int mult = dmt::symv_hier_shape(z->ctx, z, ws.p, s);
      const int W = mult * dml::hier_wave_slots();
      const int* rk[1 + kHLevels];
      void* bufs[1 + kHLevels];
      for (int l = 0; l <= kHLevels; ++l) {
        rbuf[l].alloc(dml::hier_ranges_bytes(z->n, W));
        rk[l] = ranks(l);
        bufs[l] = rbuf[l].p;
      }
      dml::hier_ranges_levels(z->n, 1 + kHLevels, rk, W, bufs, hr, s);
    }
    ProfScope p(z->ctx, level ? "symv_z32h_relaxed" : "symv_z32h", s);
    dml::symv_hier(z->n, z->z32p.p, ranks(level), z->hfac.p, z->hrmax, x, y, ws.p, hr[level], s, level >= 2);
  }
  void apply_batch(const double* X, double* Y, int R, void*, hipStream_t s) override {
    for (int j = 0; j < R; ++j) apply(X + 2 * z->n * j, Y + 2 * z->n * j, s);
  }
  size_t batch_ws(int) const override { return 0; }
};
struct OpPacked : DevOp {
  const dm_zmat* z;
  int form;
  const char* tag;
  int shape = 0;  // chosen on the first product (the matrix must be assembled by then)
  DBuf<char> ws;
  OpPacked(const dm_zmat* z_, int form_) : z(z_), form(form_) {
    tag = form == DM_Z32P ? "symv_z32" : "symv_z64";
    ws.alloc(dml::symv_workspace_bytes(z->n));
  }
  void apply(const double* x, double* y, hipStream_t s) override {
    if (!shape) shape = dmt::symv_shape(z->ctx, z, form, ws.p, s);
    ProfScope p(z->ctx, tag, s);
    dml::symv_packed(z->n, z->ref(form), x, y, ws.p, shape, s);
  }
  void apply_batch(const double* X, double* Y, int R, void*, hipStream_t s) override {
    for (int j = 0; j < R; ++j) apply(X + 2 * z->n * j, Y + 2 * z->n * j, s);
  }
  size_t batch_ws(int) const override { return 0; }
};

// The operator of one stored form.
std::unique_ptr<DevOp> make_op(const dm_zmat* z, int form);

// y = M x; apply_batch: R columns of length n (leading dimension n)
struct DevPrec {
  i64 n = 0;
  virtual ~DevPrec() = default;
  virtual void apply(const double* x, double* y, hipStream_t s) = 0;
  virtual void apply_batch(const double* X, double* Y, int R, hipStream_t s) {
    for (int j = 0; j < R; ++j) apply(X + 2 * n * j, Y + 2 * n * j, s);
  }
};
struct PrecDiag : DevPrec {
  DBuf<double> d;
  void apply(const double* x, double* y, hipStream_t s) override { dml::diag_apply(n, d.p, x, y, s); }
};
struct PrecSpai : DevPrec {
  std::shared_ptr<const SpaiPattern> pat;
  DBuf<double> val;
  void apply(const double* x, double* y, hipStream_t s) override {
    dmp::spmv(n, pat->d_rowptr.p, pat->d_colidx.p, val.p, x, y, s);
  }
  void apply_batch(const double* X, double* Y, int R, hipStream_t s) override {
    dmp::spmm(n, R, pat->d_rowptr.p, pat->d_colidx.p, val.p, X, Y, s);
  }
};
struct PrecBlock : DevPrec {
  int bs, nb;
  DBuf<double> inv;
  void apply(const double* x, double* y, hipStream_t s) override { dml::block_apply(n, bs, nb, inv.p, x, y, s); }
};

// Preconditioner of kind o.precond built from the stored form `form` of z
// (any of DM_Z64, DM_Z32, DM_Z64P, DM_Z32P).  Returns null for DM_PRECOND_NONE.
std::unique_ptr<DevPrec> build_precond(const dm_zmat* z, int form, const dm_solve_opts& o, hipStream_t s);
// DM_PRECOND_LEARNED: near-field SPAI S plus the trained network o.learned,
// M r = S r + network(S r, r) (nonlinear: right preconditioning only).
std::unique_ptr<DevPrec> make_learned_precond(const dm_zmat* z, int form, const dm_solve_opts& o, hipStream_t s);

// ------------------------------------------------------------------- GMRES
// One captured Krylov step: the graph executable for steps with k <= kb,
// valid for the operator, preconditioner and buffers it was captured with.
struct StepGraph {
  int kb = 0;
  hipGraphExec_t exec = nullptr;
};

struct Krylov {
  i64 n = 0;
  int m = 0;  // basis capacity
  DBuf<double> V, Zs, w, t, hbuf, ybuf, part, nrm;  // nrm: partials of ||w||^2
  // captured steps: the step index on the device, the step's input vector
  // (v_j) and preconditioned vector, the graphs and what they were captured for
  DBuf<int> jd;
  DBuf<double> vcur, zcur;
  int jhost = 0;
  std::vector<StepGraph> graphs;
  std::vector<const void*> graph_key;
  ~Krylov() { drop_graphs(); }
  void drop_graphs() {
    for (auto& g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    graphs.clear();
    graph_key.clear();
  }
  // device Givens state: H ((m + 1) x m complex), rotations, rotated rhs,
  // per-step (estimate, subdiagonal), 1 / subdiagonal
  DBuf<double> Hd, csd, snd, gd, stt, scd;
  double g0[2] = {0.0, 0.0};
  // Buffers only grow: m is the logical capacity (the kernels index with it),
  // so the refinement loop, whose remaining iteration budget shrinks at every
  // step, reuses the first step's buffers instead of reallocating them.
  void reserve(i64 n_, int m_, bool flexible) {
    if (n_ != n) drop_graphs();
    if (n_ != n) {
      V.release();
      Zs.release();
      w.release();
      t.release();
      part.release();
      nrm.release();
    }
    n = n_;
    m = m_;
    grow(Hd, (size_t)2 * (m + 1) * m);
    grow(csd, m);
    grow(snd, 2 * (size_t)m);
    grow(gd, 2 * (size_t)(m + 1));
    grow(stt, 2 * (size_t)m);
    grow(scd, 1);
    grow(part, 2 * dml::dots_part_count(n, m + 1));
    grow(nrm, dml::norm_part_count(n));
    grow(V, (size_t)2 * n * (m + 1));
    if (flexible) grow(Zs, (size_t)2 * n * m);
    grow(w, 2 * n);
    grow(t, 2 * n);
    grow(hbuf, 2 * (2 * (m + 1) + 1));
    grow(ybuf, 2 * (m + 1));
    if (!jd.p) jd.alloc(1);
    grow(vcur, 2 * n);
    grow(zcur, 2 * n);
  }
  static void grow(DBuf<double>& b, size_t count) {
    if (b.n < count) b.alloc(count);
  }
};

struct GmresOut {
  int iters = 0;
  bool solved = false;
  double est = 0;
};

// x_zero: x is known to be zero on entry (the first residual is then b).
// graph: steps after the first of each cycle replay a captured HIP graph.
// relax > 0: relaxed products (A.relax(relax rtol beta / est) at every check
// point, est the residual estimate and beta the cycle's initial residual).
GmresOut gmres(dm_ctx* ctx, i64 n, DevOp& A, DevPrec* M, int side, const double* b, double* x, double rtol,
               int maxiter, int restart, Krylov& K, hipStream_t s, bool x_zero = false, bool graph = false,
               double relax = 0.0);
double true_residual(DevOp& A, i64 n, const double* b, const double* x, double* tmp, double* hbuf, hipStream_t s);
double solve_one(dm_ctx* ctx, const dm_solve_opts& o, DevOp* A64, DevOp* A32, DevPrec* M, const double* b,
                 double* x, i64 n, Krylov& K, dm_solve_stats& st, bool& ok, hipEvent_t z64_ready);

// ------------------------------------------------------- lockstep batches
// Solve Z X = B for R right-hand sides at once (device, internal basis order,
// n x R with leading dimension n).  Every Krylov step applies the operator to
// all R columns with one product; each column keeps its own Arnoldi basis,
// Hessenberg matrix and stopping test, and drops out of the Arnoldi kernels
// once converged.  Method GMRES: FGMRES (right preconditioning) on A64 (or A32
// when A64 is null); MIXED: FGMRES on A32 with FP64 refinement against A64.
// Per-column iteration counts and true relative residuals are returned in
// iters / res; `ok` is set when every column met the tolerance.
void solve_batch(dm_ctx* ctx, const dm_solve_opts& o, DevOp* A64, DevOp* A32, DevPrec* M, const double* B,
                 double* X, i64 n, int R, std::vector<int>& iters, std::vector<double>& res, int& outer, bool& ok,
                 hipEvent_t z64_ready);

}  // namespace dms

// A preconditioner handed to the caller (dm_precond_create).
struct dm_precond {
  dm_ctx* ctx = nullptr;
  const dm_zmat* z = nullptr;
  std::unique_ptr<dms::DevPrec> p;
};
