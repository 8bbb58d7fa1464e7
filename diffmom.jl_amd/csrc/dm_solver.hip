// dm_solver.hip — preconditioner set-up and the single-right-hand-side Krylov
// solvers (GMRES / FGMRES with CGS2 Arnoldi, mixed-precision refinement).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>

#include "dm_solver.h"

namespace dms {

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::unique_ptr<DevOp> make_op(const dm_zmat* z, int form) {
  DM_REQUIRE(z->forms & form, DM_ERR_ARGUMENT, "requested form was not assembled");
  if (form == DM_Z64) return std::make_unique<OpZ64>(z);
  if (form == DM_Z32) return std::make_unique<OpZ32>(z);
  if (form == DM_Z32H) return std::make_unique<OpHier>(z);
  return std::make_unique<OpPacked>(z, form);
}

std::unique_ptr<DevPrec> build_precond(const dm_zmat* z, int form, const dm_solve_opts& o, hipStream_t s) {
  i64 n = z->n;
  DM_REQUIRE(z->forms & form, DM_ERR_ARGUMENT, "preconditioner source form was not assembled");
  const ZRef zr = z->ref(form);
  if (o.precond == DM_PRECOND_NONE) return nullptr;
  if (o.precond == DM_PRECOND_LEARNED) return make_learned_precond(z, form, o, s);
  if (o.precond == DM_PRECOND_DIAG) {
    auto p = std::make_unique<PrecDiag>();
    p->n = n;
    p->d.alloc(2 * n);
    dml::diag_inv(n, zr, p->d.p, s);
    return p;
  }
  if (o.precond == DM_PRECOND_NEARFIELD) {
    dm_rwg& r = const_cast<dm_rwg&>(*z->rwg);
    double cutJ = o.nf_cutoff > 0 ? o.nf_cutoff : 2.5 * r.mean_len;
    std::shared_ptr<const SpaiPattern> pat = dmh::spai_pattern(r, cutJ, dmp::spai_max_cols());
    const SpaiPattern& P = *pat;
    auto p = std::make_unique<PrecSpai>();
    p->n = n;
    p->pat = pat;
    DBuf<double> cscv;
    cscv.alloc(2 * P.nnz);
    p->val.alloc(2 * P.nnz);
    DBuf<int> st;
    st.alloc(1);
    HIP_OK(hipMemsetAsync(st.p, 0, sizeof(int), s));
    int maxJ = 1;
    for (i64 c = 0; c < n; ++c) maxJ = std::max(maxJ, P.jptr[c + 1] - P.jptr[c]);
    dmp::spai_build(n, zr, maxJ, P.d_jptr.p, P.d_jidx.p, P.d_iptr.p, P.d_iidx.p, cscv.p, st.p, s);
    dmp::csr_from_csc(P.nnz, P.d_map.p, cscv.p, p->val.p, s);
    int sv = 0;
    HIP_OK(hipMemcpyAsync(&sv, st.p, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    DM_REQUIRE(sv == 0, DM_ERR_ARGUMENT,
               "near-field preconditioner: normal equations of column " + std::to_string(sv) + " are not positive definite");
    return p;
  }
  if (o.precond == DM_PRECOND_BLOCK) {
    auto p = std::make_unique<PrecBlock>();
    p->n = n;
    p->bs = o.block_size > 0 ? o.block_size : 128;
    DM_REQUIRE(p->bs <= 1024, DM_ERR_ARGUMENT, "near-field block size must be <= 1024");
    p->bs = (int)std::min<i64>(p->bs, n);
    p->nb = (int)((n + p->bs - 1) / p->bs);
    p->inv.alloc((size_t)p->nb * p->bs * p->bs * 2);
    dml::extract_blocks(n, zr, p->bs, p->nb, p->inv.p, s);
    DBuf<int> st;
    st.alloc(p->nb);
    HIP_OK(hipMemsetAsync(st.p, 0, sizeof(int) * p->nb, s));
    dml::block_inverse(p->bs, p->nb, p->inv.p, st.p, s);
    std::vector<int> hs(p->nb);
    HIP_OK(hipMemcpyAsync(hs.data(), st.p, sizeof(int) * p->nb, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (int b = 0; b < p->nb; ++b)
      DM_REQUIRE(hs[b] == 0, DM_ERR_ARGUMENT, "near-field block " + std::to_string(b) + " is singular");
    return p;
  }
  dm_throw(DM_ERR_ARGUMENT, "preconditioner kind not available for this solve");
}

namespace {

// Ends a stream capture that an exception interrupts.
struct CaptureGuard {
  hipStream_t s;
  bool active = true;
  ~CaptureGuard() {
    if (!active) return;
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
  }
};

// The graph of one Krylov step for k = j + 1 <= kb, captured on first use:
// preconditioner and product on K.vcur (v_j), the preconditioned vector also
// copied into column j of K.Zs, two CGS passes and the Givens step (which
// writes v_{j+1} to V and K.vcur), then j += 1, all with j read from K.jd.
hipGraphExec_t step_graph(dm_ctx* ctx, i64 n, DevOp& A, DevPrec* M, int side, int mcap, int kb, Krylov& K,
                          hipStream_t s) {
  const std::vector<const void*> key = {&A, M, (const void*)(intptr_t)side, (const void*)(intptr_t)mcap,
                                        K.V.p, K.Zs.p, K.w.p, K.t.p, K.part.p, K.hbuf.p, K.Hd.p, K.gd.p};
  if (key != K.graph_key) {
    K.drop_graphs();
    K.graph_key = key;
  }
  for (auto& g : K.graphs)
    if (g.kb == kb) return g.exec;
  const bool right = M && side == 1, left = M && side == 0;
  const i64 ldv = n;  // complex columns
  const int prof = ctx->prof;
  ctx->prof = 0;  // no profiling events inside the graph
  hipGraph_t graph = nullptr;
  {
    HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    CaptureGuard guard{s};
    try {
      if (right) {
        M->apply(K.vcur.p, K.zcur.p, s);
        dml::copy_col_dev(n, K.zcur.p, K.Zs.p, ldv, K.jd.p, s);
        A.apply(K.zcur.p, K.w.p, s);
      } else if (left) {
        A.apply(K.vcur.p, K.t.p, s);
        M->apply(K.t.p, K.w.p, s);
      } else {
        A.apply(K.vcur.p, K.w.p, s);
      }
      dml::dots_dev(n, ldv, kb, K.jd.p, K.V.p, K.w.p, K.part.p, s);
      dml::update_dev(n, ldv, K.jd.p, 0, K.V.p, K.part.p, K.hbuf.p, K.w.p, nullptr, s);
      dml::dots_dev(n, ldv, kb, K.jd.p, K.V.p, K.w.p, K.part.p, s);
      dml::update_dev(n, ldv, K.jd.p, 1, K.V.p, K.part.p, K.hbuf.p, K.w.p, K.nrm.p, s);
      dml::givens_dev(kb - 1, K.jd.p, mcap + 1, K.hbuf.p, K.nrm.p, dml::norm_part_count(n), n, K.w.p, K.V.p, ldv,
                      K.vcur.p, K.Hd.p, K.csd.p, K.snd.p, K.gd.p, K.stt.p, K.scd.p, s);
      dml::step_inc(K.jd.p, s);
    } catch (...) {
      ctx->prof = prof;
      throw;
    }
    guard.active = false;
    HIP_OK(hipStreamEndCapture(s, &graph));
  }
  ctx->prof = prof;
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  HIP_OK(e);
  K.graphs.push_back({kb, exec});
  return exec;
}

}  // namespace

// Restarted (restart = 0: full) GMRES with classical Gram-Schmidt applied
// twice (CGS2) for the Arnoldi step, FP64 throughout; right preconditioning is
// flexible (FGMRES: the preconditioned directions are stored), left
// preconditioning follows Krylov.jl's semantics (stop on ||M r|| <= rtol ||M b||).
// x holds the initial guess on entry.  graph: the first step of each cycle
// runs its kernels directly (which also finishes any set-up the operator and
// preconditioner do on first use); the later steps replay a captured graph
// per step, the graph chosen by the bucket kb (a power of two >= 8, at most
// m + 1) that bounds k = j + 1, so the CGS grids grow with the basis.
GmresOut gmres(dm_ctx* ctx, i64 n, DevOp& A, DevPrec* M, int side, const double* b, double* x, double rtol,
               int maxiter, int restart, Krylov& K, hipStream_t s, bool x_zero, bool graph, double relax) {
  GmresOut out;
  int mcap = restart > 0 ? std::min(restart, maxiter) : maxiter;
  bool right = M && side == 1;
  bool left = M && side == 0;
  K.reserve(n, mcap, right);
  double* V = K.V.p;
  const i64 ldv = 2 * n;  // doubles per column
  auto col = [&](double* base, int j) { return base + (size_t)ldv * j; };
  // eps from the initial (preconditioned) residual with x0
  // (a zero initial guess needs no product: r0 = b)
  auto residual = [&](double* r) {
    HIP_OK(hipMemcpyAsync(r, b, 16 * n, hipMemcpyDeviceToDevice, s));
    if (!x_zero) {
      A.apply(x, K.t.p, s);
      dml::axpby(n, -1, 0, K.t.p, 1, 0, r, s);
    }
    x_zero = false;
    if (left) {
      M->apply(r, K.t.p, s);
      HIP_OK(hipMemcpyAsync(r, K.t.p, 16 * n, hipMemcpyDeviceToDevice, s));
    }
  };
  auto nrm = [&](const double* v) {
    double h;
    dml::norm2(n, v, K.hbuf.p, s);
    HIP_OK(hipMemcpyAsync(&h, K.hbuf.p, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    return std::sqrt(h);
  };
  double* r = col(V, 0);
  residual(r);
  double beta = nrm(r);
  double eps = rtol * beta;
  if (!(beta > 0)) {
    out.solved = true;
    return out;
  }
  // The Givens rotations run on the device (k_givens after each Arnoldi step),
  // so iterations are queued without a host round trip; the host reads the
  // residual estimates only at check points predicted from their decay rate
  // (every iteration near convergence).  Steps queued past the stopping point
  // are discarded: the leading columns of H and g do not change after their
  // own step.
  std::vector<double> sth(2 * (size_t)mcap);
  int total = 0;
  double est = beta;
  while (true) {
    dml::axpby(n, 1.0 / beta, 0, r, 0, 0, r, s);
    K.g0[0] = beta;
    K.g0[1] = 0.0;
    HIP_OK(hipMemsetAsync(K.gd.p, 0, K.gd.bytes(), s));
    HIP_OK(hipMemcpyAsync(K.gd.p, K.g0, sizeof(K.g0), hipMemcpyHostToDevice, s));
    int j = 0, jstop = -1, seen = 0, step = 1;
    int jprev = 0;
    double eprev = beta;
    est = beta;
    // relaxed products: the product error a step may carry grows as the
    // residual falls (inexact Krylov); a cycle starts at the tight form
    if (relax > 0) A.relax(relax * rtol);
    while (j < mcap && total < maxiter) {
      double* vj = col(V, j);
      if (graph && j >= 1) {
        if (j == 1) {  // the replayed steps of this cycle start at j = 1, v_1
          K.jhost = 1;
          HIP_OK(hipMemcpyAsync(K.jd.p, &K.jhost, sizeof(int), hipMemcpyHostToDevice, s));
          HIP_OK(hipMemcpyAsync(K.vcur.p, vj, 16 * n, hipMemcpyDeviceToDevice, s));
        }
        int kb = 8;
        while (kb < j + 1) kb *= 2;
        kb = std::min(kb, mcap + 1);
        hipGraphExec_t ex = step_graph(ctx, n, A, M, side, mcap, kb, K, s);
        ProfScope p(ctx, "step_graph", s);
        HIP_OK(hipGraphLaunch(ex, s));
        ++j;
        ++total;
        if (j - seen < step && j < mcap && total < maxiter) continue;
      } else {
        if (right) {
          double* zj = col(K.Zs.p, j);
          {
            ProfScope p(ctx, "precond", s);
            M->apply(vj, zj, s);
          }
          A.apply(zj, K.w.p, s);
        } else if (left) {
          A.apply(vj, K.t.p, s);
          ProfScope p(ctx, "precond", s);
          M->apply(K.t.p, K.w.p, s);
        } else {
          A.apply(vj, K.w.p, s);
        }
        const int k = j + 1;
        {
          ProfScope p(ctx, "arnoldi", s);
          double* h1 = K.hbuf.p;
          double* h2 = K.hbuf.p + 2 * k;
          dml::dots(n, ldv / 2, k, V, K.w.p, K.part.p, s);
          dml::update(n, ldv / 2, k, V, K.part.p, h1, K.w.p, nullptr, s);
          dml::dots(n, ldv / 2, k, V, K.w.p, K.part.p, s);
          dml::update(n, ldv / 2, k, V, K.part.p, h2, K.w.p, K.nrm.p, s);
          // rotations, and v_{j+1} = w / ||w|| in the same launch
          dml::givens(j, mcap + 1, K.hbuf.p, K.nrm.p, dml::norm_part_count(n), n, K.w.p, col(V, k), K.Hd.p, K.csd.p,
                      K.snd.p, K.gd.p, K.stt.p, K.scd.p, s);
        }
        ++j;
        ++total;
        if (j - seen < step && j < mcap && total < maxiter) continue;
      }
      HIP_OK(hipMemcpyAsync(sth.data() + 2 * seen, K.stt.p + 2 * seen, sizeof(double) * 2 * (j - seen),
                            hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      for (int i = seen; i < j; ++i) {
        est = sth[2 * i];
        if (est <= eps || sth[2 * i + 1] == 0) {
          jstop = i + 1;
          break;
        }
      }
      seen = j;
      if (jstop >= 0) break;
      if (relax > 0 && est > 0) A.relax(relax * rtol * beta / est);
      // next check: ~70 % of the steps the recent decay rate says are left
      const double rate = j > jprev && eprev > 0 ? std::pow(est / eprev, 1.0 / (j - jprev)) : 1.0;
      if (rate > 0 && rate < 0.999 && est > eps) {
        const double left_steps = std::log(eps / est) / std::log(rate);
        step = std::max(1, std::min(16, (int)(0.7 * left_steps)));
      } else {
        step = std::min(16, 2 * step);
      }
      jprev = j;
      eprev = est;
    }
    if (jstop < 0) jstop = j;
    total -= j - jstop;  // queued steps past the stopping point do not count
    if (relax > 0) A.relax(0.0);
    // y = H^-1 g, x += V y (or Zs y)
    if (jstop > 0) {
      dml::backsolve(jstop, mcap + 1, K.Hd.p, K.gd.p, K.ybuf.p, s);
      dml::combine(n, ldv / 2, jstop, right ? K.Zs.p : V, K.ybuf.p, x, s);
      HIP_OK(hipStreamSynchronize(s));
    }
    out.iters = total;
    out.est = est;
    if (est <= eps) {
      out.solved = true;
      break;
    }
    if (total >= maxiter) break;
    // restart from the true (preconditioned) residual
    residual(r);
    beta = nrm(r);
    if (beta <= eps) {
      out.solved = true;
      out.est = beta;
      break;
    }
  }
  return out;
}

double true_residual(DevOp& A, i64 n, const double* b, const double* x, double* tmp, double* hbuf, hipStream_t s) {
  A.apply(x, tmp, s);
  dml::axpby(n, 1, 0, b, -1, 0, tmp, s);
  double h[2];
  dml::norm2(n, tmp, hbuf, s);
  dml::norm2(n, b, hbuf + 1, s);
  HIP_OK(hipMemcpyAsync(h, hbuf, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (h[1] == 0) return h[0] == 0 ? 0.0 : INFINITY;
  return std::sqrt(h[0] / h[1]);
}

// One right-hand side on the device (internal basis order): FP64 GMRES on the
// matrix `A`, or (mixed) FGMRES on the complex64 matrix with FP64 Arnoldi and
// FP64 iterative refinement r = b - Z64 x.  `z64_ready`, when given, is waited
// on before the first complex128 product (the FP64 matrix may still be in
// assembly on the second stream).  Returns the true relative residual.
double solve_one(dm_ctx* ctx, const dm_solve_opts& o, DevOp* A64, DevOp* A32, DevPrec* M, const double* b,
                 double* x, i64 n, Krylov& K, dm_solve_stats& st, bool& ok, hipEvent_t z64_ready) {
  hipStream_t s = ctx->stream;
  DBuf<double> tmp, hb;
  tmp.alloc(2 * n);
  hb.alloc(4);
  HIP_OK(hipMemsetAsync(x, 0, 16 * n, s));
  if (o.method == DM_METHOD_GMRES) {
    DevOp& A = A64 ? *A64 : *A32;
    if (A64 && z64_ready) HIP_OK(hipStreamWaitEvent(s, z64_ready, 0));
    GmresOut g = gmres(ctx, n, A, M, o.precond_side, b, x, o.tol, o.maxiter, o.restart, K, s, true, o.graph == 1);
    st.iterations += g.iters;
    ok = g.solved;
    return true_residual(A, n, b, x, tmp.p, hb.p, s);
  }
  DBuf<double> rr, d;
  rr.alloc(2 * n);
  d.alloc(2 * n);
  HIP_OK(hipMemcpyAsync(rr.p, b, 16 * n, hipMemcpyDeviceToDevice, s));
  double rel = 1.0;
  double tau = o.inner_tol > 0 ? o.inner_tol : 0.5 * o.tol;
  int left = o.maxiter;
  bool waited = false;
  for (int outer = 0; outer < std::max(1, o.max_outer) && left > 0; ++outer) {
    HIP_OK(hipMemsetAsync(d.p, 0, 16 * n, s));
    GmresOut g = gmres(ctx, n, *A32, M, 1, rr.p, d.p, tau, left, o.restart, K, s, true, o.graph == 1, o.relax);
    st.iterations += g.iters;
    left -= g.iters;
    st.outer_iterations += 1;
    dml::axpby(n, 1, 0, d.p, 1, 0, x, s);
    if (z64_ready && !waited) {
      HIP_OK(hipStreamWaitEvent(s, z64_ready, 0));
      waited = true;
    }
    A64->apply(x, tmp.p, s);  // rr = b - Z64 x
    HIP_OK(hipMemcpyAsync(rr.p, b, 16 * n, hipMemcpyDeviceToDevice, s));
    dml::axpby(n, -1, 0, tmp.p, 1, 0, rr.p, s);
    double h[2];
    dml::norm2(n, rr.p, hb.p, s);
    dml::norm2(n, b, hb.p + 1, s);
    HIP_OK(hipMemcpyAsync(h, hb.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    rel = h[1] > 0 ? std::sqrt(h[0] / h[1]) : 0.0;
    if (rel <= o.tol) break;
    tau = std::min(0.5, std::max(1e-12, 0.5 * o.tol / rel));
  }
  ok = rel <= o.tol;
  return rel;
}

}  // namespace dms
