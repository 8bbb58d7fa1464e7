// dm_batch.hip — lockstep FGMRES over a batch of right-hand sides (many
// incidence angles of one scatterer).  One Z X product per Krylov step serves
// every column of the batch (MFMA kernel, Z streamed once per 64 columns);
// orthogonalisation is FP64 CGS2 per column, with each column's Hessenberg
// matrix, Givens rotations and stopping test on the device (one block per
// column).  The per-column arithmetic is that of the single-column solver in
// dm_solver.hip, so a batch of one reproduces it.
#include <algorithm>
#include <cmath>

#include "dm_solver.h"
#include "k_batch.h"

namespace dms {

namespace {

struct BatchKrylov {
  i64 n = 0;
  int R = 0, m = 0, kcap = 0;
  DBuf<double> V, Zs, W, T, part, h, nr, y, sc;
  DBuf<int> act, cnt;
  DBuf<char> ws;
  void reserve(i64 n_, int R_, int m_, bool flexible, size_t wsb) {
    n = n_;
    R = R_;
    m = m_;
    kcap = m + 1;
    const size_t blk = (size_t)2 * n * R;
    // buffers only grow, so the refinement loop (shrinking m) reuses them
    grow(V, blk * (m + 1));
    if (flexible) grow(Zs, blk * m);
    grow(W, blk);
    grow(T, blk);
    grow(part, (size_t)2 * R * kcap * dmb::dot_chunks(n));
    grow(h, (size_t)2 * 2 * R * kcap);
    grow(nr, R);
    grow(y, (size_t)2 * R * kcap);
    grow(sc, R);
    grow(act, R);
    grow(cnt, R);
    grow(ws, std::max<size_t>(wsb, 1));
  }
  template <class T_>
  static void grow(DBuf<T_>& b, size_t count) {
    if (b.n < count) b.alloc(count);
  }
};

std::vector<double> col_norms(BatchKrylov& K, const double* W, hipStream_t s) {
  std::vector<double> out(K.R);
  dmb::norm2(K.n, K.R, W, nullptr, K.nr.p, s);
  HIP_OK(hipMemcpyAsync(out.data(), K.nr.p, sizeof(double) * K.R, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (double& v : out) v = std::sqrt(v);
  return out;
}

// FGMRES (right preconditioning when M is given) on all columns r with
// live[r]; X holds the initial guess.  Column r stops when its residual
// estimate is <= rtol[r] times its initial residual norm.  Adds the Arnoldi
// steps of each column to iters[r]; solved[r] reports convergence.
void gmres_batch(dm_ctx* ctx, i64 n, int R, DevOp& A, DevPrec* M, const double* B, double* X,
                 const std::vector<double>& rtol, const std::vector<char>& live, int maxiter, int restart,
                 BatchKrylov& K, hipStream_t s, std::vector<int>& iters, std::vector<char>& solved,
                 bool x_zero = false, double relax = 0.0) {
  int mcap = restart > 0 ? std::min(restart, maxiter) : maxiter;
  mcap = std::min(mcap, dmb::max_basis() - 1);
  K.reserve(n, R, mcap, M != nullptr, A.batch_ws(R));
  const size_t blk = (size_t)2 * n * R;
  auto vblk = [&](double* base, int j) { return base + blk * j; };
  solved.assign(R, 0);
  // r0 = b - A x0
  // (a zero initial guess needs no product: r0 = b)
  auto residual = [&]() {
    HIP_OK(hipMemcpyAsync(K.V.p, B, 8 * blk, hipMemcpyDeviceToDevice, s));
    if (!x_zero) {
      A.apply_batch(X, K.T.p, R, K.ws.p, s);
      dml::axpby(n * R, -1, 0, K.T.p, 1, 0, K.V.p, s);
    }
    x_zero = false;
    return col_norms(K, K.V.p, s);
  };
  std::vector<double> beta = residual(), eps(R);
  std::vector<char> active(R, 0);
  for (int r = 0; r < R; ++r) {
    eps[r] = rtol[r] * beta[r];
    if (!live[r]) continue;
    if (beta[r] > 0)
      active[r] = 1;
    else
      solved[r] = 1;
  }
  // device Givens state of every column (see dmb::givens)
  DBuf<double> Hd, csd, snd, gd, epsd, betad;
  DBuf<int> jcnt, its, sol;
  Hd.alloc((size_t)2 * R * (mcap + 1) * mcap);
  csd.alloc((size_t)R * mcap);
  snd.alloc((size_t)2 * R * mcap);
  gd.alloc((size_t)2 * R * (mcap + 1));
  epsd.upload(eps, s);
  betad.alloc(R);
  jcnt.alloc(R);
  its.alloc(R);
  sol.alloc(R);
  std::vector<int> act_h(R), its_h(R), sol_h(R);
  // relaxed products (inexact Krylov): at every check the product may carry
  // a relative error of relax * eps_r / est_r for the most demanding active
  // column (eps_r = rtol_r beta_r, est_r its residual estimate)
  DBuf<double> estd;
  std::vector<double> est_h(R);
  if (relax > 0) estd.alloc(R);
  auto relax_to = [&](bool start) {
    if (relax <= 0) return;
    double eta = INFINITY;
    for (int r = 0; r < R; ++r) {
      if (!active[r] || (!start && !act_h[r])) continue;
      const double e = start ? rtol[r] : (est_h[r] > 0 ? eps[r] / est_h[r] : INFINITY);
      eta = std::min(eta, relax * e);
    }
    A.relax(std::isfinite(eta) ? eta : 0.0);
  };
  int total = 0;
  auto any_active = [&]() { return std::any_of(active.begin(), active.end(), [](char a) { return a != 0; }); };
  while (any_active() && total < maxiter) {
    for (int r = 0; r < R; ++r) act_h[r] = active[r];
    HIP_OK(hipMemcpyAsync(K.act.p, act_h.data(), sizeof(int) * R, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemcpyAsync(betad.p, beta.data(), sizeof(double) * R, hipMemcpyHostToDevice, s));
    HIP_OK(hipMemsetAsync(gd.p, 0, gd.bytes(), s));
    HIP_OK(hipMemsetAsync(sol.p, 0, sizeof(int) * R, s));
    dmb::givens_init(R, mcap, betad.p, K.act.p, gd.p, K.sc.p, jcnt.p, its.p, s);
    dmb::scale(n, R, K.V.p, K.sc.p, K.V.p, s);
    relax_to(true);
    // steps are queued without a host round trip; the host reads the active
    // flags every 4 steps (columns that converge in between are masked on
    // the device, so only the Z X of a fully converged batch is wasted)
    for (int j = 0; j < mcap && total < maxiter; ++j) {
      double* vj = vblk(K.V.p, j);
      if (M) {
        double* zj = vblk(K.Zs.p, j);
        {
          ProfScope p(ctx, "precond", s);
          M->apply_batch(vj, zj, R, s);
        }
        A.apply_batch(zj, K.W.p, R, K.ws.p, s);
      } else {
        A.apply_batch(vj, K.W.p, R, K.ws.p, s);
      }
      const int k = j + 1;
      double* h1 = K.h.p;
      double* h2 = K.h.p + 2 * (size_t)R * K.kcap;
      {
        ProfScope p(ctx, "arnoldi", s);
        dmb::dots(n, R, k, K.kcap, K.V.p, K.W.p, K.act.p, K.part.p, h1, s);
        dmb::update(n, R, k, K.kcap, K.V.p, h1, K.act.p, K.W.p, s);
        dmb::dots(n, R, k, K.kcap, K.V.p, K.W.p, K.act.p, K.part.p, h2, s);
        dmb::update(n, R, k, K.kcap, K.V.p, h2, K.act.p, K.W.p, s);
        dmb::norm2(n, R, K.W.p, K.act.p, K.nr.p, s);
        dmb::givens(R, j, mcap, K.kcap, h1, h2, K.nr.p, Hd.p, csd.p, snd.p, gd.p, K.sc.p, K.act.p, epsd.p, jcnt.p,
                    its.p, sol.p, s, relax > 0 ? estd.p : nullptr);
        // v_{j+1} = w / ||w|| for the columns still iterating (sc = 0 otherwise)
        if (k <= mcap) dmb::scale(n, R, K.W.p, K.sc.p, vblk(K.V.p, k), s);
      }
      ++total;
      if (k % 4 == 0 || k == mcap || total >= maxiter) {
        HIP_OK(hipMemcpyAsync(act_h.data(), K.act.p, sizeof(int) * R, hipMemcpyDeviceToHost, s));
        if (relax > 0) HIP_OK(hipMemcpyAsync(est_h.data(), estd.p, sizeof(double) * R, hipMemcpyDeviceToHost, s));
        HIP_OK(hipStreamSynchronize(s));
        bool any = false;
        for (int r = 0; r < R; ++r) any = any || act_h[r] != 0;
        if (!any) break;
        relax_to(false);
      }
    }
    HIP_OK(hipMemcpyAsync(its_h.data(), its.p, sizeof(int) * R, hipMemcpyDeviceToHost, s));
    HIP_OK(hipMemcpyAsync(sol_h.data(), sol.p, sizeof(int) * R, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (int r = 0; r < R; ++r) {
      iters[r] += its_h[r];
      if (sol_h[r]) solved[r] = 1;
      active[r] = 0;
    }
    // x_r += (Zs or V) y_r with y_r = H_r^-1 g_r over the column's own steps
    dmb::backsolve(R, mcap, K.kcap, Hd.p, gd.p, jcnt.p, K.y.p, s);
    dmb::combine(n, R, K.kcap, M ? K.Zs.p : K.V.p, K.y.p, jcnt.p, X, s);
    // a column that filled its basis (or broke down) unconverged restarts
    // from its true residual, as the single-column solver does
    bool pending = false;
    for (int r = 0; r < R; ++r) pending = pending || (live[r] && !solved[r]);
    if (!pending || total >= maxiter) break;
    beta = residual();
    for (int r = 0; r < R; ++r) {
      active[r] = 0;
      if (!live[r] || solved[r]) continue;
      if (beta[r] <= eps[r])
        solved[r] = 1;
      else
        active[r] = 1;
    }
  }
  if (relax > 0) A.relax(0.0);
  HIP_OK(hipStreamSynchronize(s));
}

}  // namespace

void solve_batch(dm_ctx* ctx, const dm_solve_opts& o, DevOp* A64, DevOp* A32, DevPrec* M, const double* B,
                 double* X, i64 n, int R, std::vector<int>& iters, std::vector<double>& res, int& outer, bool& ok,
                 hipEvent_t z64_ready) {
  hipStream_t s = ctx->stream;
  DM_REQUIRE(R >= 1 && R <= kMfmaMaxRhs, DM_ERR_ARGUMENT, "batch width out of range");
  DM_REQUIRE(o.precond_side == 1 || M == nullptr, DM_ERR_ARGUMENT,
             "batched solves use right (flexible) preconditioning");
  iters.assign(R, 0);
  res.assign(R, 0.0);
  outer = 0;
  const size_t blk = (size_t)2 * n * R;
  BatchKrylov K;
  HIP_OK(hipMemsetAsync(X, 0, 8 * blk, s));
  DBuf<double> bn_d;
  bn_d.alloc(R);
  std::vector<double> bn(R);
  dmb::norm2(n, R, B, nullptr, bn_d.p, s);
  HIP_OK(hipMemcpyAsync(bn.data(), bn_d.p, sizeof(double) * R, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  for (double& v : bn) v = std::sqrt(v);
  std::vector<char> solved(R, 0), live(R, 1);
  auto true_res = [&](DevOp& A, double* Rr, void* ws) {
    // Rr = B - A X, per-column relative norms
    A.apply_batch(X, Rr, R, ws, s);
    dml::axpby(n * R, 1, 0, B, -1, 0, Rr, s);
    dmb::norm2(n, R, Rr, nullptr, bn_d.p, s);
    std::vector<double> rr(R);
    HIP_OK(hipMemcpyAsync(rr.data(), bn_d.p, sizeof(double) * R, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
    for (int r = 0; r < R; ++r) rr[r] = bn[r] > 0 ? std::sqrt(rr[r]) / bn[r] : (rr[r] == 0 ? 0.0 : INFINITY);
    return rr;
  };
  if (o.method == DM_METHOD_GMRES) {
    DevOp& A = A64 ? *A64 : *A32;
    if (A64 && z64_ready) HIP_OK(hipStreamWaitEvent(s, z64_ready, 0));
    std::vector<double> rtol(R, o.tol);
    gmres_batch(ctx, n, R, A, M, B, X, rtol, live, o.maxiter, o.restart, K, s, iters, solved, true);
    DBuf<double> Rr;
    DBuf<char> ws;
    Rr.alloc(blk);
    ws.alloc(std::max<size_t>(A.batch_ws(R), 1));
    res = true_res(A, Rr.p, ws.p);
    ok = std::all_of(solved.begin(), solved.end(), [](char c) { return c != 0; });
    return;
  }
  // mixed precision: FGMRES on the complex64 matrix, FP64 refinement
  DBuf<double> RR, D, ws;
  RR.alloc(blk);
  D.alloc(blk);
  ws.alloc(std::max<size_t>(A64->batch_ws(R), 1) / sizeof(double) + 1);
  HIP_OK(hipMemcpyAsync(RR.p, B, 8 * blk, hipMemcpyDeviceToDevice, s));
  std::vector<double> rel(R), tau(R, o.inner_tol > 0 ? o.inner_tol : 0.5 * o.tol);
  for (int r = 0; r < R; ++r) rel[r] = bn[r] > 0 ? 1.0 : 0.0;
  bool waited = false;
  for (int it = 0; it < std::max(1, o.max_outer); ++it) {
    int used = *std::max_element(iters.begin(), iters.end());
    int left = o.maxiter - used;
    for (int r = 0; r < R; ++r) live[r] = rel[r] > o.tol;
    if (left <= 0 || !std::any_of(live.begin(), live.end(), [](char c) { return c != 0; })) break;
    HIP_OK(hipMemsetAsync(D.p, 0, 8 * blk, s));
    gmres_batch(ctx, n, R, *A32, M, RR.p, D.p, tau, live, left, o.restart, K, s, iters, solved, true, o.relax);
    outer += 1;
    dml::axpby(n * R, 1, 0, D.p, 1, 0, X, s);
    if (z64_ready && !waited) {
      HIP_OK(hipStreamWaitEvent(s, z64_ready, 0));
      waited = true;
    }
    rel = true_res(*A64, RR.p, ws.p);
    for (int r = 0; r < R; ++r) tau[r] = std::min(0.5, std::max(1e-12, 0.5 * o.tol / std::max(rel[r], 1e-300)));
  }
  res = rel;
  ok = std::all_of(rel.begin(), rel.end(), [&](double v) { return v <= o.tol; });
}

}  // namespace dms
