// dm_mpnn.hip — the learned preconditioner: network state, training data
// resident on the device, the training loop (forward, objective, backward and
// Adam on the GPU), and the flexible preconditioner M r = S (r + d(S r, r))
// used inside FGMRES, with S the near-field sparse approximate inverse and d
// the network's correction.
//
// The objective is the relative residual that one application of M leaves,
// ||Z M r - r||^2 / ||r||^2, the quantity right-preconditioned FGMRES
// consumes.  Each training mesh contributes its operator A = Z S (a
// complex64 copy on the device); the correction passes through S, so the
// objective's curvature in d is that of A^H A (clustered near the identity
// by S) rather than that of Z^H Z.
#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <random>

#include "dm_solver.h"
#include "k_mpnn.h"

struct dm_mpnn {
  dm_ctx* ctx = nullptr;
  DBuf<float> p, m, v, g;  // FP32 master parameters, Adam moments, gradient
  DBuf<uint16_t> pb;       // bfloat16 copy used by the kernels
  int step = 0;
};

struct dm_mpnn_data {
  dm_ctx* ctx = nullptr;
  i64 nnode = 0;
  int ns = 0;
  std::vector<int64_t> ptr;  // host copy of the sample offsets
  DBuf<int> nbr, smp;
  DBuf<float> sgn, dinv, geom;
  DBuf<double> z, r;
  DBuf<int64_t> d_ptr;
  // node-order complex64 operators A = Z S, the operator of each sample (-1:
  // none yet) and p = r - A r per node
  std::vector<std::unique_ptr<DBuf<float>>> mats;
  std::vector<int> mat_of;
  DBuf<double> p;
};

namespace dms {

// Graph and geometry features of one mesh in the given node order
// (internal = true: the library's internal basis order, else caller order):
// bases sharing a triangle, neighbours in increasing node index, sign =
// product of the two bases' orientations on the shared triangle.
void mpnn_graph(const dm_rwg& r, double k, bool internal, std::vector<int>& nbr, std::vector<float>& sgn,
                std::vector<float>& dinv, std::vector<float>& geom) {
  const i64 n = r.n;
  const i64 nt = r.mesh->nt;
  auto node = [&](i64 e) { return internal ? (i64)r.int_of_ext[e] : e; };
  std::vector<std::vector<std::pair<i64, float>>> on_tri(nt);
  for (i64 e = 0; e < n; ++e) {
    on_tri[r.tplus[e] - 1].push_back({e, 1.0f});
    on_tri[r.tminus[e] - 1].push_back({e, -1.0f});
  }
  nbr.assign(n * dmn::DEG, -1);
  sgn.assign(n * dmn::DEG, 0.0f);
  dinv.assign(n, 1.0f);
  geom.assign(2 * n, 0.0f);
  for (i64 e = 0; e < n; ++e) {
    std::vector<std::pair<i64, float>> lst;
    const i64 tris[2] = {r.tplus[e] - 1, r.tminus[e] - 1};
    const float se[2] = {1.0f, -1.0f};
    for (int a = 0; a < 2; ++a)
      for (auto& q : on_tri[tris[a]])
        if (q.first != e) lst.push_back({node(q.first), se[a] * q.second});
    std::sort(lst.begin(), lst.end());
    DM_REQUIRE(lst.size() <= (size_t)dmn::DEG, DM_ERR_ARGUMENT, "a basis shares its triangles with more than 4 bases");
    const i64 i = node(e);
    for (size_t q = 0; q < lst.size(); ++q) {
      nbr[i * dmn::DEG + q] = (int)lst[q].first;
      sgn[i * dmn::DEG + q] = lst[q].second;
    }
    dinv[i] = 1.0f / (float)std::max<size_t>(lst.size(), 1);
    const double len = r.len[e];
    geom[2 * i] = (float)(k * len / (2.0 * 3.14159265358979323846 / 10.0));
    geom[2 * i + 1] = (float)((r.area_p[e] + r.area_m[e]) / (len * len));
  }
}

namespace {

// Activation buffers for batches of up to `cap` nodes.
struct Work {
  i64 cap = 0;
  DBuf<float> u, h, t, m, y, gy, bwork, part, q;
  DBuf<double> cden, parts, loss, zwork;
  void reserve(i64 nodes, int samples) {
    if (nodes > cap) {
      cap = nodes;
      q.alloc(2 * cap);
      zwork.alloc(2 * cap);
      u.alloc(cap * dmn::FIN);
      h.alloc((size_t)(dmn::L + 1) * cap * dmn::H);
      t.alloc((size_t)dmn::L * cap * dmn::H);
      m.alloc((size_t)dmn::L * cap * dmn::H);
      y.alloc(2 * cap);
      gy.alloc(2 * cap);
      bwork.alloc(dmn::backward_floats(cap));
      parts.alloc(cap);
      part.alloc(dmn::outer_floats());
    }
    if (cden.n < (size_t)4 * samples) cden.alloc((size_t)4 * samples);
  }
};

// forward + objective (+ backward into net.g when grad) over samples [s0, s1)
void run_batch(dm_mpnn& net, const dm_mpnn_data& d, int s0, int s1, Work& w, double* loss_dst, bool grad,
               hipStream_t s) {
  const int64_t node0 = d.ptr[s0];
  const i64 nn = d.ptr[s1] - node0;
  const int ns = s1 - s0;
  w.reserve(nn, ns);
  dmn::stats(ns, d.d_ptr.p + s0, node0, d.z.p + 2 * node0, d.r.p + 2 * node0, w.cden.p, s);
  dmn::features(nn, d.smp.p + node0, s0, w.cden.p, d.z.p + 2 * node0, d.r.p + 2 * node0, d.geom.p + 2 * node0,
                w.u.p, s);
  dmn::Graph g{nn, d.nbr.p + dmn::DEG * node0, d.sgn.p + dmn::DEG * node0, d.dinv.p + node0};
  dmn::Cache c{w.u.p, w.h.p, w.t.p, w.m.p, w.y.p};
  dmn::forward(g, node0, net.pb.p, c, nullptr, s);
  // consecutive samples that share an operator form one group of columns
  for (int a = s0; a < s1;) {
    const int mi = d.mat_of[a];
    int b = a + 1;
    while (b < s1 && d.mat_of[b] == mi) ++b;
    const i64 n = d.ptr[a + 1] - d.ptr[a];
    const i64 lo = d.ptr[a] - node0;
    dmn::res_loss(n, b - a, d.mats[mi]->p, w.y.p + 2 * lo, d.p.p + 2 * d.ptr[a], w.cden.p + 4 * (a - s0), ns,
                  w.q.p + 2 * lo, w.parts.p + lo, grad ? w.gy.p + 2 * lo : nullptr, w.zwork.p, s);
    a = b;
  }
  dmn::sum(nn, w.parts.p, loss_dst, s);
  if (grad) dmn::backward(g, node0, net.pb.p, c, w.gy.p, net.g.p, w.bwork.p, w.part.p, s);
}

void require_operators(const dm_mpnn_data& d, int64_t s0, int64_t s1) {
  bool ok = !d.mat_of.empty();
  for (int64_t a = s0; ok && a < s1; ++a) ok = d.mat_of[a] >= 0;
  DM_REQUIRE(ok, DM_ERR_ARGUMENT, "every training sample needs its operator A = Z S (dm_mpnn_data_set_matrix)");
}

// M r = S (r + d(S r, r)) = z + S d: flexible, for right-preconditioned FGMRES
struct PrecLearned : DevPrec {
  std::unique_ptr<DevPrec> spai;
  const dm_mpnn* net = nullptr;
  DBuf<int> nbr, smp;
  DBuf<float> sgn, dinv, geom, u, h, hb, y;
  DBuf<double> zb, db, sd, cden;
  DBuf<int64_t> ptr;
  void apply(const double* x, double* out, hipStream_t s) override {
    spai->apply(x, zb.p, s);
    ProfScope p(net->ctx, "learned", s);
    dmn::stats(1, ptr.p, 0, zb.p, x, cden.p, s);
    dmn::features(n, smp.p, 0, cden.p, zb.p, x, geom.p, u.p, s);
    dmn::Graph g{n, nbr.p, sgn.p, dinv.p};
    dmn::Cache c{u.p, h.p, nullptr, nullptr, y.p};
    dmn::forward(g, 0, net->pb.p, c, hb.p, s);
    dmn::correction(n, y.p, cden.p, db.p, s);
    spai->apply(db.p, sd.p, s);
    dmn::add(n, zb.p, sd.p, out, s);
  }
};

}  // namespace

std::unique_ptr<DevPrec> make_learned_precond(const dm_zmat* z, int form, const dm_solve_opts& o, hipStream_t s) {
  DM_REQUIRE(o.learned != nullptr, DM_ERR_ARGUMENT, "learned preconditioner requested without a trained model");
  DM_REQUIRE(o.learned->ctx == z->ctx, DM_ERR_ARGUMENT, "model belongs to another context");
  DM_REQUIRE(o.precond_side == 1, DM_ERR_ARGUMENT,
             "the learned preconditioner is nonlinear: use right (flexible) preconditioning");
  dm_solve_opts so = o;
  so.precond = DM_PRECOND_NEARFIELD;
  // a network whose decoder is zero adds nothing: the SPAI alone, without
  // running the network (the gate of a trained network may be 0)
  std::vector<float> wd(2 * dmn::H);
  HIP_OK(hipMemcpyAsync(wd.data(), o.learned->p.p + dmn::OFF_WD, sizeof(float) * wd.size(), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  if (std::all_of(wd.begin(), wd.end(), [](float v) { return v == 0.0f; })) return build_precond(z, form, so, s);
  auto p = std::make_unique<PrecLearned>();
  p->spai = build_precond(z, form, so, s);
  p->net = o.learned;
  const i64 n = z->n;
  p->n = n;
  std::vector<int> nbr;
  std::vector<float> sgn, dinv, geom;
  mpnn_graph(*z->rwg, z->k_re, true, nbr, sgn, dinv, geom);
  p->nbr.upload(nbr, s);
  p->sgn.upload(sgn, s);
  p->dinv.upload(dinv, s);
  p->geom.upload(geom, s);
  std::vector<int> zeros(n, 0);
  p->smp.upload(zeros, s);
  std::vector<int64_t> pr = {0, n};
  p->ptr.upload(pr, s);
  p->u.alloc(n * dmn::FIN);
  p->h.alloc(n * dmn::H);
  p->hb.alloc(n * dmn::H);
  p->y.alloc(2 * n);
  p->zb.alloc(2 * n);
  p->db.alloc(2 * n);
  p->sd.alloc(2 * n);
  p->cden.alloc(4);
  HIP_OK(hipStreamSynchronize(s));
  return p;
}

}  // namespace dms

using namespace dms;

namespace {

template <class F>
int mp_guard(F&& f) {
  try {
    f();
    return DM_OK;
  } catch (const DmError& e) {
    dm_set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    dm_set_error("host allocation failed");
    return DM_ERR_MEMORY;
  } catch (const std::exception& e) {
    dm_set_error(e.what());
    return DM_ERR_DEVICE;
  }
}

struct Scope {
  std::lock_guard<std::mutex> lk;
  explicit Scope(dm_ctx* c) : lk(c->mu) { HIP_OK(hipSetDevice(c->device)); }
};
}  // namespace

extern "C" {

int64_t dm_mpnn_param_count(void) { return dmn::NPARAM; }

int dm_mpnn_create(dm_ctx* ctx, const float* params, dm_mpnn** out) {
  return mp_guard([&] {
    DM_REQUIRE(ctx && params && out, DM_ERR_ARGUMENT, "NULL argument");
    for (int i = 0; i < dmn::NPARAM; ++i) DM_REQUIRE(std::isfinite(params[i]), DM_ERR_ARGUMENT, "parameters must be finite");
    Scope sc(ctx);
    hipStream_t s = ctx->stream;
    auto net = std::make_unique<dm_mpnn>();
    net->ctx = ctx;
    net->p.upload(params, dmn::NPARAM, s);
    net->m.alloc(dmn::NPARAM);
    net->v.alloc(dmn::NPARAM);
    net->g.alloc(dmn::NPARAM);
    net->pb.alloc(dmn::NPARAM);
    HIP_OK(hipMemsetAsync(net->m.p, 0, net->m.bytes(), s));
    HIP_OK(hipMemsetAsync(net->v.p, 0, net->v.bytes(), s));
    HIP_OK(hipMemsetAsync(net->g.p, 0, net->g.bytes(), s));
    dmn::to_bf16(net->p.p, net->pb.p, s);
    HIP_OK(hipStreamSynchronize(s));
    *out = net.release();
  });
}

int dm_mpnn_get_params(const dm_mpnn* net, float* out) {
  return mp_guard([&] {
    DM_REQUIRE(net && out, DM_ERR_ARGUMENT, "NULL argument");
    Scope sc(net->ctx);
    HIP_OK(hipMemcpy(out, net->p.p, net->p.bytes(), hipMemcpyDeviceToHost));
  });
}

int dm_mpnn_free(dm_mpnn* net) {
  return mp_guard([&] {
    if (!net) return;
    Scope sc(net->ctx);
    delete net;
  });
}

int dm_rwg_graph(const dm_rwg* rwg, double k, int32_t* nbr, float* sgn, float* geom) {
  return mp_guard([&] {
    DM_REQUIRE(rwg && nbr && sgn && geom, DM_ERR_ARGUMENT, "NULL argument");
    DM_REQUIRE(std::isfinite(k) && k > 0, DM_ERR_ARGUMENT, "wavenumber must be finite and positive");
    std::vector<int> nb;
    std::vector<float> sg, di, ge;
    mpnn_graph(*rwg, k, false, nb, sg, di, ge);
    std::copy(nb.begin(), nb.end(), nbr);
    std::copy(sg.begin(), sg.end(), sgn);
    std::copy(ge.begin(), ge.end(), geom);
  });
}

int dm_mpnn_data_create(dm_ctx* ctx, int64_t nnode, const int32_t* nbr, const float* sgn, const float* geom,
                        const double* z, const double* r, int64_t nsample, const int64_t* sample_ptr,
                        dm_mpnn_data** out) {
  return mp_guard([&] {
    DM_REQUIRE(ctx && nbr && sgn && geom && z && r && sample_ptr && out, DM_ERR_ARGUMENT, "NULL argument");
    DM_REQUIRE(nsample >= 1 && nsample < (1LL << 31), DM_ERR_ARGUMENT, "sample count out of range");
    DM_REQUIRE(sample_ptr[0] == 0 && sample_ptr[nsample] == nnode, DM_ERR_ARGUMENT,
               "sample offsets must start at 0 and end at the node count");
    std::vector<int> smp(nnode);
    std::vector<float> dinv(nnode);
    for (int64_t s = 0; s < nsample; ++s) {
      DM_REQUIRE(sample_ptr[s + 1] > sample_ptr[s], DM_ERR_ARGUMENT, "empty or decreasing sample");
      for (int64_t i = sample_ptr[s]; i < sample_ptr[s + 1]; ++i) {
        smp[i] = (int)s;
        int deg = 0;
        for (int q = 0; q < dmn::DEG; ++q) {
          int j = nbr[i * dmn::DEG + q];
          if (j < 0) continue;
          DM_REQUIRE(j >= sample_ptr[s] && j < sample_ptr[s + 1], DM_ERR_ARGUMENT,
                     "a neighbour lies outside its sample's nodes");
          ++deg;
        }
        dinv[i] = 1.0f / (float)std::max(deg, 1);
      }
    }
    Scope sc(ctx);
    hipStream_t s = ctx->stream;
    auto d = std::make_unique<dm_mpnn_data>();
    d->ctx = ctx;
    d->nnode = nnode;
    d->ns = (int)nsample;
    d->ptr.assign(sample_ptr, sample_ptr + nsample + 1);
    d->nbr.upload(nbr, nnode * dmn::DEG, s);
    d->sgn.upload(sgn, nnode * dmn::DEG, s);
    d->geom.upload(geom, 2 * nnode, s);
    d->z.upload(z, 2 * nnode, s);
    d->r.upload(r, 2 * nnode, s);
    d->smp.upload(smp, s);
    d->dinv.upload(dinv, s);
    d->d_ptr.upload(d->ptr, s);
    HIP_OK(hipStreamSynchronize(s));
    *out = d.release();
  });
}

int dm_mpnn_residual_operator(dm_ctx* ctx, const dm_zmat* z, const dm_precond* p, float* out) {
  return mp_guard([&] {
    DM_REQUIRE(ctx && z && p && out, DM_ERR_ARGUMENT, "NULL argument");
    DM_REQUIRE(z->ctx == ctx && p->ctx == ctx, DM_ERR_ARGUMENT, "objects belong to another context");
    DM_REQUIRE(p->p && p->p->n == z->n && p->z && p->z->rwg == z->rwg, DM_ERR_ARGUMENT,
               "the preconditioner was built for another mesh");
    DM_REQUIRE(z->forms & DM_Z64, DM_ERR_ARGUMENT, "the operator needs the full complex128 matrix (DM_Z64)");
    const i64 n = z->n;
    DM_REQUIRE(n <= 16384, DM_ERR_ARGUMENT, "training operators are limited to 16384 bases");
    Scope sc(ctx);
    hipStream_t s = ctx->stream;
    // S as a dense internal-order matrix (S applied to the identity), then
    // A = Z S on the FP64 matrix cores, then the node-order complex64 copy
    DBuf<double> eye, sd, a;
    eye.alloc((size_t)2 * n * n);
    sd.alloc((size_t)2 * n * n);
    a.alloc((size_t)2 * n * n);
    dmn::identity(n, eye.p, s);
    p->p->apply_batch(eye.p, sd.p, (int)n, s);
    DBuf<char> ws;
    ws.alloc(dml::gemm_c128_workspace_bytes(n, (int)n));
    dml::gemm_c128(n, n, z->z64.p, z->ld, sd.p, n, (int)n, a.p, n, ws.p, s);
    DBuf<float> ac;
    ac.alloc((size_t)2 * n * n);
    dmn::gather_c64(n, a.p, z->rwg->d_iofe.p, ac.p, s);
    HIP_OK(hipMemcpyAsync(out, ac.p, ac.bytes(), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_mpnn_data_set_matrix(dm_mpnn_data* d, int64_t s0, int64_t s1, const float* a, int64_t n) {
  return mp_guard([&] {
    DM_REQUIRE(d && a, DM_ERR_ARGUMENT, "NULL argument");
    DM_REQUIRE(s0 >= 0 && s0 < s1 && s1 <= d->ns, DM_ERR_ARGUMENT, "sample range out of bounds");
    DM_REQUIRE(n >= 1 && n <= 16384, DM_ERR_ARGUMENT, "operator order out of range (1 .. 16384)");
    for (int64_t k = s0; k < s1; ++k)
      DM_REQUIRE(d->ptr[k + 1] - d->ptr[k] == n, DM_ERR_ARGUMENT, "a sample's node count differs from the operator order");
    for (i64 i = 0; i < 2 * n * n; ++i) DM_REQUIRE(std::isfinite(a[i]), DM_ERR_ARGUMENT, "operator entries must be finite");
    Scope sc(d->ctx);
    hipStream_t s = d->ctx->stream;
    if (d->mat_of.empty()) d->mat_of.assign(d->ns, -1);
    if (d->p.n == 0) d->p.alloc(2 * d->nnode);
    d->mats.push_back(std::make_unique<DBuf<float>>());
    DBuf<float>& ac = *d->mats.back();
    ac.upload(a, (size_t)2 * n * n, s);
    const int64_t o = d->ptr[s0];
    dmn::res_rhs(n, (int)(s1 - s0), ac.p, d->r.p + 2 * o, d->p.p + 2 * o, s);
    for (int64_t k = s0; k < s1; ++k) d->mat_of[k] = (int)d->mats.size() - 1;
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_mpnn_data_free(dm_mpnn_data* d) {
  return mp_guard([&] {
    if (!d) return;
    Scope sc(d->ctx);
    delete d;
  });
}

int dm_mpnn_loss_grad(dm_ctx* ctx, dm_mpnn* net, const dm_mpnn_data* d, int64_t s0, int64_t s1, double* loss,
                      float* grad) {
  return mp_guard([&] {
    DM_REQUIRE(ctx && net && d && loss, DM_ERR_ARGUMENT, "NULL argument");
    DM_REQUIRE(net->ctx == ctx && d->ctx == ctx, DM_ERR_ARGUMENT, "objects belong to another context");
    DM_REQUIRE(s0 >= 0 && s0 < s1 && s1 <= d->ns, DM_ERR_ARGUMENT, "sample range out of bounds");
    require_operators(*d, s0, s1);
    Scope sc(ctx);
    hipStream_t s = ctx->stream;
    Work w;
    w.loss.alloc(1);
    run_batch(*net, *d, (int)s0, (int)s1, w, w.loss.p, grad != nullptr, s);
    HIP_OK(hipMemcpyAsync(loss, w.loss.p, sizeof(double), hipMemcpyDeviceToHost, s));
    if (grad) HIP_OK(hipMemcpyAsync(grad, net->g.p, net->g.bytes(), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_mpnn_train(dm_ctx* ctx, dm_mpnn* net, const dm_mpnn_data* d, int epochs, int64_t batch_samples, double lr,
                  uint64_t seed, double* epoch_loss) {
  return mp_guard([&] {
    DM_REQUIRE(ctx && net && d, DM_ERR_ARGUMENT, "NULL argument");
    DM_REQUIRE(net->ctx == ctx && d->ctx == ctx, DM_ERR_ARGUMENT, "objects belong to another context");
    DM_REQUIRE(epochs >= 1 && batch_samples >= 1, DM_ERR_ARGUMENT, "epochs and batch size must be positive");
    DM_REQUIRE(std::isfinite(lr) && lr > 0, DM_ERR_ARGUMENT, "learning rate must be finite and positive");
    require_operators(*d, 0, d->ns);
    Scope sc(ctx);
    hipStream_t s = ctx->stream;
    std::vector<std::pair<int, int>> batches;
    for (int a = 0; a < d->ns; a += (int)batch_samples) batches.push_back({a, std::min<int>(d->ns, a + (int)batch_samples)});
    Work w;
    i64 maxn = 0;
    for (auto& b : batches) maxn = std::max<i64>(maxn, d->ptr[b.second] - d->ptr[b.first]);
    w.reserve(maxn, (int)batch_samples);
    w.loss.alloc(batches.size());
    std::mt19937_64 rng(seed);
    std::vector<double> bl(batches.size());
    for (int ep = 0; ep < epochs; ++ep) {
      std::shuffle(batches.begin(), batches.end(), rng);
      for (size_t b = 0; b < batches.size(); ++b) {
        run_batch(*net, *d, batches[b].first, batches[b].second, w, w.loss.p + b, true, s);
        dmn::adam(net->p.p, net->g.p, net->m.p, net->v.p, ++net->step, (float)lr, net->pb.p, s);
      }
      HIP_OK(hipMemcpyAsync(bl.data(), w.loss.p, sizeof(double) * bl.size(), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      double tot = 0;
      size_t cnt = 0;
      for (size_t b = 0; b < batches.size(); ++b) {
        tot += bl[b] * (batches[b].second - batches[b].first);
        cnt += batches[b].second - batches[b].first;
      }
      if (epoch_loss) epoch_loss[ep] = tot / (double)cnt;
      DM_REQUIRE(std::isfinite(tot), DM_ERR_CONVERGENCE, "training diverged (non-finite loss)");
    }
  });
}

int dm_mpnn_apply(dm_ctx* ctx, const dm_mpnn* net, const dm_rwg* rwg, double k, const double* z, const double* r,
                  double* out) {
  return mp_guard([&] {
    DM_REQUIRE(ctx && net && rwg && z && r && out, DM_ERR_ARGUMENT, "NULL argument");
    DM_REQUIRE(net->ctx == ctx && rwg->ctx == ctx, DM_ERR_ARGUMENT, "objects belong to another context");
    Scope sc(ctx);
    hipStream_t s = ctx->stream;
    const i64 n = rwg->n;
    std::vector<int> nbr;
    std::vector<float> sgn, dinv, geom;
    mpnn_graph(*rwg, k, false, nbr, sgn, dinv, geom);
    DBuf<int> dnbr, smp;
    DBuf<float> dsgn, ddinv, dgeom, u, h, hb, y;
    DBuf<double> dz, dr, dout, cden;
    DBuf<int64_t> ptr;
    dnbr.upload(nbr, s);
    dsgn.upload(sgn, s);
    ddinv.upload(dinv, s);
    dgeom.upload(geom, s);
    std::vector<int> zeros(n, 0);
    smp.upload(zeros, s);
    std::vector<int64_t> pr = {0, n};
    ptr.upload(pr, s);
    dz.upload(z, 2 * n, s);
    dr.upload(r, 2 * n, s);
    u.alloc(n * dmn::FIN);
    h.alloc(n * dmn::H);
    hb.alloc(n * dmn::H);
    y.alloc(2 * n);
    dout.alloc(2 * n);
    cden.alloc(4);
    dmn::stats(1, ptr.p, 0, dz.p, dr.p, cden.p, s);
    dmn::features(n, smp.p, 0, cden.p, dz.p, dr.p, dgeom.p, u.p, s);
    dmn::Graph g{n, dnbr.p, dsgn.p, ddinv.p};
    dmn::Cache c{u.p, h.p, nullptr, nullptr, y.p};
    dmn::forward(g, 0, net->pb.p, c, hb.p, s);
    dmn::correction(n, y.p, cden.p, dout.p, s);
    HIP_OK(hipMemcpyAsync(out, dout.p, 16 * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  });
}

}  // extern "C"
