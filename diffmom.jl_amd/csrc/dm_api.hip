// dm_api.hip — the C ABI of libdiffmom (include/diffmom.h) and the host-side
// orchestration of assembly, products and Krylov solves.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>

#include "dm_common.h"
#include "k_linalg.h"
#include "dm_solver.h"
#include "k_batch.h"

using namespace dms;

thread_local DevCache* tl_cache = nullptr;

// Scratch of the allocation-free product (dm_apply_noalloc): a file-scope
// device array, one copy per GPU for the life of the process.  Host code never
// uses the array's name as an address (that is a host-side stand-in): dm_create
// takes each device's address of it once with hipGetSymbolAddress, with that
// device current, and kernels get that pointer.
static constexpr size_t kScratchBytes = size_t(16) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};

// with `device` current
void scratch_init(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  DM_REQUIRE(device >= 0 && device < kMaxDevices, DM_ERR_ARGUMENT, "device index out of range for the scratch table");
  if (g_scratch[device]) return;
  void* ptr = nullptr;
  HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  DM_REQUIRE(ptr != nullptr, DM_ERR_DEVICE, "no device address for the scratch array");
  g_scratch[device] = ptr;
}

void* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch[device] : nullptr;
}
}  // namespace

void* DevCache::take(size_t b) {
  auto it = free_blocks.find(b);
  if (it == free_blocks.end()) return nullptr;
  void* p = it->second;
  free_blocks.erase(it);
  bytes -= b;
  return p;
}
bool DevCache::give(void* p, size_t b) {
  if (bytes + b > cap) return false;
  pending.emplace_back(b, p);
  bytes += b;
  return true;
}
void DevCache::promote(hipStream_t a, hipStream_t b) {
  if (pending.empty()) return;
  const hipError_t ea = hipStreamQuery(a);
  const hipError_t eb = ea == hipSuccess ? hipStreamQuery(b) : hipSuccess;
  if (ea != hipSuccess || eb != hipSuccess) {
    // work still queued: drop only that status, so that a later check of the
    // last error does not see it (any other error is left for its owner)
    if ((ea == hipSuccess || ea == hipErrorNotReady) && (eb == hipSuccess || eb == hipErrorNotReady))
      (void)hipGetLastError();
    return;
  }
  for (auto& e : pending) free_blocks.emplace(e.first, e.second);
  pending.clear();
}
void DevCache::trim() {
  for (auto& e : free_blocks) {
    (void)hipFree(e.second);
    bytes -= e.first;
  }
  free_blocks.clear();
}
void DevCache::clear() {
  trim();
  for (auto& e : pending) (void)hipFree(e.second);
  pending.clear();
  bytes = 0;
}

namespace {
thread_local std::string g_err;
}  // namespace

void dm_set_error(const std::string& msg) { g_err = msg; }

namespace {

template <class F>
int guard(F&& f) {
  try {
    f();
    return DM_OK;
  } catch (const DmError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return DM_ERR_MEMORY;
  } catch (const std::exception& e) {
    g_err = e.what();
    return DM_ERR_DEVICE;
  }
}

// A public call on a context: its lock, its device, and its memory cache for
// the buffers the call allocates and releases.
struct CtxScope {
  std::lock_guard<std::mutex> lk;
  dm_ctx* c;
  DevCache* prev;
  explicit CtxScope(dm_ctx* c_) : lk(c_->mu), c(c_), prev(tl_cache) {
    HIP_OK(hipSetDevice(c->device));
    c->cache.promote(c->stream, c->stream2);
    tl_cache = &c->cache;
  }
  ~CtxScope() {
    c->cache.promote(c->stream, c->stream2);
    tl_cache = prev;
  }
};


// _validated_efie_prefactors: src/assembly/EFIE.jl:297-330
void validate_prefactors(double kr, double ki, double eta0) {
  DM_REQUIRE(std::isfinite(kr) && std::isfinite(ki) && (kr != 0 || ki != 0), DM_ERR_ARGUMENT,
             "EFIE wavenumber must be finite and nonzero");
  DM_REQUIRE(std::isfinite(eta0) && eta0 != 0, DM_ERR_ARGUMENT, "EFIE eta0 must be numeric, finite, and nonzero");
  cd k(kr, ki);
  cd k2 = k * k;
  DM_REQUIRE(std::isfinite(k2.real()) && std::isfinite(k2.imag()) && std::abs(k2) != 0, DM_ERR_OVERFLOW,
             "EFIE wavenumber squared is not representable");
  cd inv = 1.0 / k2;
  DM_REQUIRE(std::isfinite(inv.real()) && std::isfinite(inv.imag()) && std::abs(inv) != 0, DM_ERR_OVERFLOW,
             "EFIE inverse wavenumber squared is not representable");
  cd om = k * eta0;
  DM_REQUIRE(std::isfinite(om.real()) && std::isfinite(om.imag()) && std::abs(om) != 0, DM_ERR_OVERFLOW,
             "EFIE k*eta0 prefactor is not representable");
}

struct NearTables {
  DBuf<double> jself, jadj;
  DBuf<int> flag;
};

// Assembly arguments with the near-field tables written to caller-provided
// device buffers (jself: Nt x 18, jadj: Nt x 54 doubles; flag: one int that the
// caller cleared, shared by several meshes if it likes).
AsmArgs make_args_into(dm_rwg& r, const TriTables& T, double kr, double ki, double eta0, i64 ld, double* jself,
                       double* jadj, int* flag, bool launch_near = true) {
  AsmArgs a;
  a.n = r.n;
  a.ld = ld;
  a.nt = r.mesh->nt;
  a.nblk = r.nblk;
  a.TB = r.TB;
  a.blk_off = r.d_blk_off.p;
  a.blk_tri = r.d_blk_tri.p;
  a.hrwg = r.d_hrwg.p;
  a.hsig = r.d_hsig.p;
  a.nb = r.d_nb.p;
  a.ext = r.d_ext.p;
  a.len = r.d_len.p;
  a.tplus_i = r.d_tplus_i.p;
  a.cen = T.cen.p;
  a.uq = T.uq.p;
  a.av = T.av.p;
  double xi[14], w[7];
  int nq = dmh::quad_rule(T.quad_order, xi, w);
  for (int i = 0; i < 7; ++i) a.w[i] = i < nq ? w[i] : 0.0;
  a.nq = nq;
  a.k_re = kr;
  a.k_im = ki;
  a.eta0 = eta0;
  hipStream_t s = r.ctx->stream;
  a.jself = jself;
  a.jadj = jadj;
  a.flag = flag;
  if (launch_near) launch_near_tables(a, T, jself, jadj, s);
  // tile-side images (depend on the RWG and quadrature order only; kept on the RWG)
  a.side_cap = r.max_blk_tris;
  {
    std::lock_guard<std::mutex> lk(r.tab_mu);
    for (int f = 0; f < 2; ++f) {
      DBuf<char>& img = r.side_img[T.quad_order][f];
      const size_t stride = side_image_bytes(nq, f == 1, r.max_blk_tris);
      if (!img.p) {
        img.alloc(stride * r.nblk);
        launch_build_sides(a, r.max_blk_tris, f == 1, img.p, s);
      }
      if (f == 1) {
        a.side_img64 = img.p;
        a.side_stride64 = (i64)stride;
      } else {
        a.side_img32 = img.p;
        a.side_stride32 = (i64)stride;
      }
    }
  }
  return a;
}

AsmArgs make_args(dm_rwg& r, const TriTables& T, double kr, double ki, double eta0, i64 ld, NearTables& nt) {
  const i64 ntri = r.mesh->nt;
  nt.jself.alloc(ntri * 18);
  nt.jadj.alloc(ntri * 54);
  nt.flag.alloc(1);
  HIP_OK(hipMemsetAsync(nt.flag.p, 0, sizeof(int), r.ctx->stream));
  return make_args_into(r, T, kr, ki, eta0, ld, nt.jself.p, nt.jadj.p, nt.flag.p);
}

void check_flag(NearTables& nt, hipStream_t s) {
  int f = 0;
  HIP_OK(hipMemcpyAsync(&f, nt.flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  DM_REQUIRE(f == 0, DM_ERR_OVERFLOW, "EFIE entry is outside the representable ComplexF64 range.");
}

i64 round_ld(i64 n) { return (n + 63) / 64 * 64; }

constexpr int kAllForms = DM_Z64 | DM_Z32 | DM_Z64P | DM_Z32P;

// FP64 tile launches with the tuned threads per workgroup: key
// asm/<form>/q<quadrature points>/<n>, form z64 / z64p, "+z32" when the same
// launch also writes the complex64 rounding.
void asm_tuned(dm_ctx* ctx, const AsmArgs& a, double* z, int cap, i64 n, bool packed, hipStream_t s) {
  const std::string key = std::string("asm/") + (packed ? "z64p" : "z64") + (a.z32_dual ? "+z32" : "") + "/q" +
                          std::to_string(a.nq) + "/" + std::to_string(n);
  auto run = [&](int t) {
    if (packed)
      launch_assemble_z64p(a, z, cap, s, t);
    else
      launch_assemble_z64(a, z, cap, s, t);
  };
  run(dmt::tune_choice(ctx, key, {1024, 512}, 1024, run, s));
}

// Packed symmetric forms requested in `forms` (storage zeroed first: entries
// past n stay zero).
void assemble_packed(dm_zmat* z, const AsmArgs& a, const dm_rwg& r, int forms, hipStream_t s) {
  if (forms & DM_Z64P) {
    z->z64p.alloc(packed_scalars(z->n));
    zero_packed_padding(z->n, z->z64p.p, sizeof(double), s);
    ProfScope p(z->ctx, "assemble_z64", s);
    asm_tuned(z->ctx, a, z->z64p.p, r.max_blk_tris, z->n, true, s);
  }
  if (forms & DM_Z32P) {
    z->z32p.alloc(packed_scalars(z->n));
    zero_packed_padding(z->n, z->z32p.p, sizeof(float), s);
    ProfScope p(z->ctx, "assemble_z32", s);
    launch_assemble_z32p(a, z->z32p.p, r.max_blk_tris, s);
  }
}


}  // namespace

namespace {

// Widest lockstep batch that fits in half of the free device memory: per
// column the Krylov basis and its preconditioned copy (2 m + 1 vectors), five
// work vectors and the split-K product workspace.
int batch_width(i64 n, i64 nrhs, int maxiter, int restart) {
  i64 m = restart > 0 ? std::min(restart, maxiter) : maxiter;
  m = std::min<i64>(m, dmb::max_basis() - 1);
  size_t free_b = 0, total_b = 0;
  HIP_OK(hipMemGetInfo(&free_b, &total_b));
  const double per_col = 16.0 * n * (2.0 * m + 7.0) + (double)dml::zx_workspace_bytes(n, 1);
  i64 w = (i64)(0.5 * (double)free_b / per_col);
  w = std::min<i64>(std::min<i64>(w, kMfmaMaxRhs), nrhs);
  DM_REQUIRE(w >= 1, DM_ERR_MEMORY, "not enough device memory for the Krylov basis of one right-hand side");
  return (int)w;
}

// Lockstep batches serve several right-hand sides with right (flexible) or no
// preconditioning.
bool batched_solve(const dm_solve_opts& o, i64 nrhs) {
  if (o.method == DM_METHOD_LU || o.batch < 0 || nrhs < 1) return false;
  if (o.precond != DM_PRECOND_NONE && o.precond_side != 1) return false;
  return o.batch > 0 || nrhs >= kMfmaMinRhs;
}

bool use_batch(const dm_solve_opts& o, i64 nrhs, const DevPrec* M) {
  return batched_solve(o, nrhs) && (M == nullptr || o.precond_side == 1);
}

// Operator form of one precision held by z: the packed symmetric form for
// one-column products (half the bytes), the full form for lockstep batches
// (the MFMA product reads full storage).  0 when neither is stored.
int pick_form(const dm_zmat* z, bool f32, bool batched) {
  const int full = f32 ? DM_Z32 : DM_Z64, packed = f32 ? DM_Z32P : DM_Z64P;
  const bool hf = (z->forms & full) != 0, hp = (z->forms & packed) != 0;
  if (!hf && !hp) return 0;
  if (batched) return hf ? full : packed;
  return hp ? packed : full;
}

// Operator form of the complex64 side of a mixed solve: the compressed form
// when it exists and columns are solved one at a time.
int op_form(const dm_zmat* z, bool batched) {
  if (!batched && (z->forms & DM_Z32H)) return DM_Z32H;
  return pick_form(z, true, batched);
}

constexpr int kAcaRankCap = 48;

// Cross approximation of the packed complex64 form (see dm_zmat_compress):
// queued on stream s without a host round trip; compression_stats reads the
// ranks back.
void compress_z32p(dm_zmat* z, double eps, int rmax, hipStream_t s) {
  DM_REQUIRE(z->forms & DM_Z32P, DM_ERR_ARGUMENT, "compression needs the packed complex64 form (DM_Z32P)");
  DM_REQUIRE(eps > 0 && eps < 1, DM_ERR_ARGUMENT, "compression tolerance must be in (0, 1)");
  DM_REQUIRE(rmax >= 1 && rmax <= 64, DM_ERR_ARGUMENT, "rank cap must be in 1..64");
  const i64 n = z->n, T = packed_tiles(n);
  z->hwork.alloc(T + 1);
  z->hrank.alloc(T);
  z->hrank_lv.alloc((size_t)kHLevels * T);
  z->hfac.alloc((size_t)T * 2 * kPT * rmax * 2);
  z->hrmax = rmax;
  z->heps = eps;
  z->hlevel = 0;
  dml::aca_compress(n, z->z32p.p, eps, rmax, z->hwork.p, z->hrank.p, z->hrank_lv.p, z->hfac.p, s);
  z->forms |= DM_Z32H;
}

// [0] stored tiles, [1] low-rank tiles, [2] their mean rank, [3] bytes a
// product reads relative to the dense packed form.
void compression_stats(dm_zmat* z, double* stats, hipStream_t s, int level = 0) {
  const i64 T = packed_tiles(z->n);
  std::vector<int> rk(T);
  const int* src = level == 0 ? z->hrank.p : z->hrank_lv.p + (size_t)(level - 1) * T;
  HIP_OK(hipMemcpyAsync(rk.data(), src, sizeof(int) * T, hipMemcpyDeviceToHost, s));
  HIP_OK(hipStreamSynchronize(s));
  i64 low = 0;
  double rsum = 0, bytes = 0;
  for (i64 t = 0; t < T; ++t) {
    if (rk[t] >= 0) {
      ++low;
      rsum += rk[t];
      bytes += 2.0 * kPT * rk[t] * 8;
    } else {
      bytes += (double)kPT * kPT * 8;
    }
  }
  if (level == 0) z->hbytes = bytes;
  stats[0] = (double)T;
  stats[1] = (double)low;
  stats[2] = low ? rsum / low : 0.0;
  stats[3] = bytes / ((double)T * kPT * kPT * 8);
}

// Solve the nrhs columns of Bd (device, internal basis order, n x nrhs) and
// write the currents in caller order to host X (leading dimension ldx).
void solve_columns(dm_ctx* ctx, const dm_solve_opts& o, const dm_rwg& r, DevOp* A64, DevOp* A32, DevPrec* M,
                   const double* Bd, i64 n, i64 nrhs, double* X, i64 ldx, dm_solve_stats& st, double& worst,
                   bool& all_ok, hipEvent_t z64_ready, bool conj_out = false) {
  hipStream_t s = ctx->stream;
  if (nrhs <= 0) return;
  if (!use_batch(o, nrhs, M)) {
    DBuf<double> x, xe;
    x.alloc(2 * n);
    xe.alloc(2 * n);
    Krylov K;
    for (i64 j = 0; j < nrhs; ++j) {
      bool ok = false;
      double res = solve_one(ctx, o, A64, A32, M, Bd + 2 * n * j, x.p, n, K, st, ok, z64_ready);
      worst = std::max(worst, res);
      all_ok = all_ok && ok;
      launch_permute(n, r.d_ext.p, x.p, xe.p, 0, s);
      if (conj_out) dml::conj(n, xe.p, s);
      HIP_OK(hipMemcpyAsync(X + 2 * ldx * j, xe.p, 16 * n, hipMemcpyDeviceToHost, s));
    }
    HIP_OK(hipStreamSynchronize(s));
    return;
  }
  const int W = batch_width(n, nrhs, o.maxiter, o.restart);
  DBuf<double> Xb, xe;
  Xb.alloc((size_t)2 * n * W);
  xe.alloc((size_t)2 * n * W);
  for (i64 j0 = 0; j0 < nrhs; j0 += W) {
    const int R = (int)std::min<i64>(W, nrhs - j0);
    std::vector<int> iters;
    std::vector<double> res;
    int outer = 0;
    bool ok = false;
    solve_batch(ctx, o, A64, A32, M, Bd + 2 * n * j0, Xb.p, n, R, iters, res, outer, ok, z64_ready);
    for (int c = 0; c < R; ++c) {
      st.iterations += iters[c];
      worst = std::max(worst, res[c]);
      launch_permute(n, r.d_ext.p, Xb.p + 2 * n * c, xe.p + 2 * n * c, 0, s);
    }
    st.outer_iterations = std::max(st.outer_iterations, outer);
    all_ok = all_ok && ok;
    if (conj_out) dml::conj(n * R, xe.p, s);
    HIP_OK(hipMemcpy2DAsync(X + 2 * ldx * j0, 16 * ldx, xe.p, 16 * n, 16 * n, R, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  }
}

void validate_opts(const dm_solve_opts& o) {
  DM_REQUIRE(std::isfinite(o.tol) && o.tol > 0, DM_ERR_ARGUMENT, "GMRES tol must be finite and positive");
  DM_REQUIRE(o.maxiter >= 1 && o.maxiter <= 1000, DM_ERR_ARGUMENT, "maxiter must be in [1, 1000]");
  DM_REQUIRE(o.precond_side == 0 || o.precond_side == 1, DM_ERR_ARGUMENT, "precond_side must be 0 (left) or 1 (right)");
  DM_REQUIRE(o.method >= DM_METHOD_LU && o.method <= DM_METHOD_MIXED, DM_ERR_ARGUMENT, "unknown solve method");
  DM_REQUIRE(std::isfinite(o.compress_tol) && o.compress_tol >= 0 && o.compress_tol < 1, DM_ERR_ARGUMENT,
             "compress_tol must be in [0, 1)");
  DM_REQUIRE(o.storage == 0 || o.storage == 1, DM_ERR_ARGUMENT, "storage must be 0 (packed) or 1 (full)");
  DM_REQUIRE(o.graph == 0 || o.graph == 1, DM_ERR_ARGUMENT, "graph must be 0 or 1");
  DM_REQUIRE(std::isfinite(o.relax) && o.relax >= 0, DM_ERR_ARGUMENT, "relax must be finite and >= 0");
}

void check_ctx(const void* a, const char* what) { DM_REQUIRE(a != nullptr, DM_ERR_ARGUMENT, std::string(what) + " is NULL"); }

}  // namespace

extern "C" {

const char* dm_last_error(void) { return g_err.c_str(); }
int dm_version(void) { return 100; }

int dm_create(int device, dm_ctx** out) {
  return guard([&] {
    check_ctx(out, "output handle");
    int count = 0;
    HIP_OK(hipGetDeviceCount(&count));
    DM_REQUIRE(device >= 0 && device < count, DM_ERR_ARGUMENT, "device index out of range");
    HIP_OK(hipSetDevice(device));
    auto c = std::make_unique<dm_ctx>();
    c->device = device;
    HIP_OK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_OK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    c->cache.device = device;
    c->cache.cap = std::min<size_t>(size_t(64) << 30, prop.totalGlobalMem / 4);
    scratch_init(device);
    *out = c.release();
  });
}

int dm_destroy(dm_ctx* ctx) {
  return guard([&] {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipStreamSynchronize(ctx->stream2);
    ctx->cache.clear();
    for (auto& r : ctx->recs) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : ctx->ev_free) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(ctx->stream);
    (void)hipStreamDestroy(ctx->stream2);
    delete ctx;
  });
}

int dm_sync(dm_ctx* ctx) {
  return guard([&] {
    check_ctx(ctx, "context");
    CtxScope sc(ctx);
    HIP_OK(hipStreamSynchronize(ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream2));
  });
}

int dm_profile_enable(dm_ctx* ctx, int on) {
  return guard([&] {
    check_ctx(ctx, "context");
    CtxScope sc(ctx);
    DM_REQUIRE(on >= 0 && on <= 2, DM_ERR_ARGUMENT, "profile mode must be 0 (off), 1 (all) or 2 (products)");
    ctx->prof = on;
  });
}

int dm_profile_reset(dm_ctx* ctx) {
  return guard([&] {
    check_ctx(ctx, "context");
    CtxScope sc(ctx);
    HIP_OK(hipStreamSynchronize(ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream2));
    for (auto& r : ctx->recs) {
      ctx->ev_free.push_back(r.a);
      ctx->ev_free.push_back(r.b);
    }
    ctx->recs.clear();
  });
}

int dm_profile_read(dm_ctx* ctx, const char* name, int64_t* count, double* total_ms) {
  return guard([&] {
    check_ctx(ctx, "context");
    CtxScope sc(ctx);
    i64 c = 0;
    double t = 0;
    for (auto& r : ctx->recs) {
      if (r.name != name) continue;
      HIP_OK(hipEventSynchronize(r.b));
      float ms = 0;
      HIP_OK(hipEventElapsedTime(&ms, r.a, r.b));
      t += ms;
      ++c;
    }
    *count = c;
    *total_ms = t;
  });
}

int dm_mesh_create(dm_ctx* ctx, const double* xyz, int64_t nv, const int64_t* tri, int64_t nt, dm_mesh** out) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(out, "output handle");
    DM_REQUIRE(nv >= 0 && nt >= 0, DM_ERR_ARGUMENT, "negative mesh size");
    DM_REQUIRE((nv == 0 || xyz) && (nt == 0 || tri), DM_ERR_ARGUMENT, "mesh arrays are NULL");
    auto m = std::make_unique<dm_mesh>();
    m->ctx = ctx;
    m->nv = nv;
    m->nt = nt;
    m->xyz.assign(xyz, xyz + 3 * nv);
    m->tri.assign(tri, tri + 3 * nt);
    for (i64 i = 0; i < 3 * nt; ++i)
      DM_REQUIRE(m->tri[i] >= 1 && m->tri[i] <= nv, DM_ERR_ARGUMENT, "triangle vertex index out of range");
    *out = m.release();
  });
}

int dm_mesh_free(dm_mesh* mesh) {
  return guard([&] { delete mesh; });
}

int dm_mesh_check(const dm_mesh* mesh, int allow_boundary, int require_closed, double area_tol_rel) {
  return guard([&] {
    check_ctx(mesh, "mesh");
    dmh::mesh_check(*mesh, allow_boundary, require_closed, area_tol_rel);
  });
}

int dm_mesh_areas(const dm_mesh* mesh, double* areas) {
  return guard([&] {
    check_ctx(mesh, "mesh");
    for (i64 t = 0; t < mesh->nt; ++t) areas[t] = dmh::triangle_area(*mesh, t);
  });
}

int dm_build_rwg(const dm_mesh* mesh, int precheck, int allow_boundary, int require_closed, double area_tol_rel,
                 dm_rwg** out) {
  return guard([&] {
    check_ctx(mesh, "mesh");
    check_ctx(out, "output handle");
    CtxScope sc(mesh->ctx);
    if (precheck) dmh::mesh_check(*mesh, allow_boundary, require_closed, area_tol_rel);
    auto r = std::make_unique<dm_rwg>();
    r->mesh = mesh;
    r->ctx = mesh->ctx;
    dmh::build_rwg_canonical(*mesh, *r);
    DM_REQUIRE(r->n >= 1, DM_ERR_ARGUMENT,
               "mesh produces no RWG unknowns; at least two triangles sharing an edge are required");
    dmh::finalize_rwg(*r);
    *out = r.release();
  });
}

int dm_rwg_create(const dm_mesh* mesh, int64_t n, const int64_t* tplus, const int64_t* tminus, const int64_t* evert,
                  const int64_t* vplus_opp, const int64_t* vminus_opp, const double* len, const double* area_plus,
                  const double* area_minus, const double* coeff_plus, const double* coeff_minus, dm_rwg** out) {
  return guard([&] {
    check_ctx(mesh, "mesh");
    check_ctx(out, "output handle");
    DM_REQUIRE(n >= 1, DM_ERR_ARGUMENT, "RWG data has no basis functions");
    DM_REQUIRE(tplus && tminus && vplus_opp && vminus_opp && len && area_plus && area_minus, DM_ERR_ARGUMENT,
               "RWG arrays are NULL");
    CtxScope sc(mesh->ctx);
    auto r = std::make_unique<dm_rwg>();
    r->mesh = mesh;
    r->ctx = mesh->ctx;
    r->n = n;
    r->tplus.assign(tplus, tplus + n);
    r->tminus.assign(tminus, tminus + n);
    if (evert) r->evert.assign(evert, evert + 2 * n);
    r->vplus_opp.assign(vplus_opp, vplus_opp + n);
    r->vminus_opp.assign(vminus_opp, vminus_opp + n);
    r->len.assign(len, len + n);
    r->area_p.assign(area_plus, area_plus + n);
    r->area_m.assign(area_minus, area_minus + n);
    if (coeff_plus) r->coeff_p.assign(coeff_plus, coeff_plus + n); else r->coeff_p.assign(n, 1.0);
    if (coeff_minus) r->coeff_m.assign(coeff_minus, coeff_minus + n); else r->coeff_m.assign(n, 1.0);
    dmh::finalize_rwg(*r);
    *out = r.release();
  });
}

int dm_rwg_size(const dm_rwg* rwg, int64_t* n) {
  return guard([&] {
    check_ctx(rwg, "rwg");
    *n = rwg->n;
  });
}

int dm_rwg_export(const dm_rwg* r, int64_t* tplus, int64_t* tminus, int64_t* evert, int64_t* vplus_opp,
                  int64_t* vminus_opp, double* len, double* area_plus, double* area_minus) {
  return guard([&] {
    check_ctx(r, "rwg");
    i64 n = r->n;
    if (tplus) std::copy(r->tplus.begin(), r->tplus.end(), tplus);
    if (tminus) std::copy(r->tminus.begin(), r->tminus.end(), tminus);
    if (evert) {
      if ((i64)r->evert.size() == 2 * n) std::copy(r->evert.begin(), r->evert.end(), evert);
      else std::fill(evert, evert + 2 * n, 0);
    }
    if (vplus_opp) std::copy(r->vplus_opp.begin(), r->vplus_opp.end(), vplus_opp);
    if (vminus_opp) std::copy(r->vminus_opp.begin(), r->vminus_opp.end(), vminus_opp);
    if (len) std::copy(r->len.begin(), r->len.end(), len);
    if (area_plus) std::copy(r->area_p.begin(), r->area_p.end(), area_plus);
    if (area_minus) std::copy(r->area_m.begin(), r->area_m.end(), area_minus);
  });
}

int dm_rwg_free(dm_rwg* rwg) {
  return guard([&] {
    if (!rwg) return;
    std::lock_guard<std::mutex> lk(rwg->ctx->mu);
    (void)hipSetDevice(rwg->ctx->device);
    delete rwg;
  });
}

int dm_rwg_adjacent_pairs(const dm_rwg* rwg, int64_t* count) {
  return guard([&] {
    check_ctx(rwg, "rwg");
    *count = rwg->adj_pairs;
  });
}

int dm_assemble_z(dm_ctx* ctx, const dm_rwg* rwg_c, double k_re, double k_im, double eta0, int quad_order, int forms,
                  int64_t max_output_bytes, dm_zmat** out) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(rwg_c, "rwg");
    check_ctx(out, "output handle");
    DM_REQUIRE(rwg_c->ctx == ctx, DM_ERR_ARGUMENT, "RWG data belongs to another context");
    const bool rounded = (forms & DM_Z32_ROUNDED) != 0;
    forms &= ~DM_Z32_ROUNDED;
    DM_REQUIRE(forms > 0 && (forms & ~kAllForms) == 0, DM_ERR_ARGUMENT,
               "forms must be a mask of DM_Z64, DM_Z32, DM_Z64P, DM_Z32P (and DM_Z32_ROUNDED)");
    if (rounded)
      DM_REQUIRE(((forms & DM_Z32) == 0 || (forms & DM_Z64)) && ((forms & DM_Z32P) == 0 || (forms & DM_Z64P)),
                 DM_ERR_ARGUMENT, "DM_Z32_ROUNDED needs the complex128 form of the same storage");
    CtxScope sc(ctx);
    dm_rwg& r = const_cast<dm_rwg&>(*rwg_c);
    i64 n = r.n;
    // byte gate on the dense complex128 payload (src/assembly/EFIE.jl:711-716)
    if (max_output_bytes > 0) {
      long double bytes = (long double)n * (long double)n * 16.0L;
      DM_REQUIRE(bytes <= (long double)max_output_bytes, DM_ERR_ARGUMENT,
                 "dense EFIE matrix requires " + std::to_string((unsigned long long)bytes) +
                     " bytes, exceeding max_output_bytes=" + std::to_string(max_output_bytes));
    }
    validate_prefactors(k_re, k_im, eta0);
    const TriTables& T = dmh::tables(r, quad_order);
    auto z = std::make_unique<dm_zmat>();
    z->ctx = ctx;
    z->rwg = rwg_c;
    z->n = n;
    z->ld = round_ld(n);
    z->forms = forms;
    z->k_re = k_re;
    z->k_im = k_im;
    z->eta0 = eta0;
    z->quad_order = quad_order;
    hipStream_t s = ctx->stream;
    NearTables nt;
    AsmArgs a;
    {
      ProfScope p(ctx, "near_tables", s);
      a = make_args(r, T, k_re, k_im, eta0, z->ld, nt);
    }
    if (forms & DM_Z32) z->z32.alloc((size_t)z->ld * n * 2);
    if (forms & DM_Z64) {
      z->z64.alloc((size_t)z->ld * n * 2);
      ProfScope p(ctx, "assemble_z64", s);
      a.z32_dual = rounded && (forms & DM_Z32) ? z->z32.p : nullptr;
      asm_tuned(ctx, a, z->z64.p, r.max_blk_tris, n, false, s);
      a.z32_dual = nullptr;
    }
    if ((forms & DM_Z32) && !rounded) {
      ProfScope p(ctx, "assemble_z32", s);
      launch_assemble_z32(a, z->z32.p, r.max_blk_tris, s);
    }
    if (rounded && (forms & DM_Z32P)) {
      z->z64p.alloc(packed_scalars(n));
      z->z32p.alloc(packed_scalars(n));
      zero_packed_padding(n, z->z64p.p, sizeof(double), s);
      zero_packed_padding(n, z->z32p.p, sizeof(float), s);
      ProfScope p(ctx, "assemble_z64_z32", s);
      a.z32_dual = z->z32p.p;
      asm_tuned(ctx, a, z->z64p.p, r.max_blk_tris, n, true, s);
      a.z32_dual = nullptr;
    } else {
      assemble_packed(z.get(), a, r, forms, s);
    }
    check_flag(nt, s);
    *out = z.release();
  });
}

int dm_assemble_z_multi(dm_ctx* ctx, int64_t nmesh, const dm_rwg* const* rwgs, double k_re, double k_im,
                        double eta0, int quad_order, int forms, dm_zmat** outs) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(rwgs, "rwg list");
    check_ctx(outs, "output handles");
    DM_REQUIRE(nmesh >= 1 && nmesh <= 65536, DM_ERR_ARGUMENT, "mesh count out of range");
    DM_REQUIRE(forms > 0 && (forms & ~kAllForms) == 0, DM_ERR_ARGUMENT,
               "forms must be a mask of DM_Z64, DM_Z32, DM_Z64P, DM_Z32P");
    for (i64 m = 0; m < nmesh; ++m) {
      check_ctx(rwgs[m], "rwg");
      DM_REQUIRE(rwgs[m]->ctx == ctx, DM_ERR_ARGUMENT, "RWG data belongs to another context");
    }
    validate_prefactors(k_re, k_im, eta0);
    CtxScope sc(ctx);
    hipStream_t s = ctx->stream;
    std::vector<std::unique_ptr<dm_zmat>> zs(nmesh);
    std::vector<AsmArgs> args(nmesh);
    int cap = 0;
    // near-field tables of all meshes in one pair of buffers, one overflow flag
    std::vector<i64> toff(nmesh + 1, 0);
    for (i64 m = 0; m < nmesh; ++m) toff[m + 1] = toff[m] + rwgs[m]->mesh->nt;
    NearTables nt;
    nt.jself.alloc((size_t)toff[nmesh] * 18);
    nt.jadj.alloc((size_t)toff[nmesh] * 54);
    nt.flag.alloc(1);
    HIP_OK(hipMemsetAsync(nt.flag.p, 0, sizeof(int), s));
    std::vector<const TriTables*> tabs(nmesh);
    DBuf<char> near_staging;
    {
      ProfScope p(ctx, "near_tables", s);
      for (i64 m = 0; m < nmesh; ++m) {
        dm_rwg& r = const_cast<dm_rwg&>(*rwgs[m]);
        const TriTables& T = dmh::tables(r, quad_order);
        auto z = std::make_unique<dm_zmat>();
        z->ctx = ctx;
        z->rwg = rwgs[m];
        z->n = r.n;
        z->ld = round_ld(r.n);
        z->forms = forms;
        z->k_re = k_re;
        z->k_im = k_im;
        z->eta0 = eta0;
        z->quad_order = quad_order;
        args[m] = make_args_into(r, T, k_re, k_im, eta0, z->ld, nt.jself.p + 18 * toff[m], nt.jadj.p + 54 * toff[m],
                                 nt.flag.p, false);
        tabs[m] = &T;
        cap = std::max(cap, r.max_blk_tris);
        zs[m] = std::move(z);
      }
      // every mesh's near-field tables in one launch per kind
      launch_near_tables_multi(args, tabs, nt.jself.p, nt.jadj.p, near_staging, s);
    }
    // work list: largest meshes first
    std::vector<i64> order(nmesh);
    for (i64 m = 0; m < nmesh; ++m) order[m] = m;
    std::stable_sort(order.begin(), order.end(), [&](i64 a, i64 b) { return args[a].nblk > args[b].nblk; });
    std::vector<AsmArgs> oargs(nmesh);
    for (i64 q = 0; q < nmesh; ++q) oargs[q] = args[order[q]];
    // the persistent kernel lays out LDS for the largest block; meshes with
    // smaller blocks get side images built for that layout (kept on the RWG
    // set per quadrature order and layout, like its own-layout images)
    for (i64 q = 0; q < nmesh; ++q) {
      AsmArgs& a = oargs[q];
      if (a.side_cap == cap) continue;
      dm_rwg& r = const_cast<dm_rwg&>(*rwgs[order[q]]);
      std::lock_guard<std::mutex> lk(r.tab_mu);
      for (int f = 0; f < 2; ++f) {
        const size_t stride = side_image_bytes(a.nq, f == 1, cap);
        DBuf<char>& img = r.side_img_cap[{quad_order * 2 + f, cap}];
        if (!img.p) {
          img.alloc(stride * a.nblk);
          launch_build_sides(a, cap, f == 1, img.p, s);
        }
        if (f == 1) {
          a.side_img64 = img.p;
          a.side_stride64 = (i64)stride;
        } else {
          a.side_img32 = img.p;
          a.side_stride32 = (i64)stride;
        }
      }
      a.side_cap = cap;
    }
    for (int form : {DM_Z64, DM_Z32, DM_Z64P, DM_Z32P}) {
      if (!(forms & form)) continue;
      std::vector<void*> ptrs(nmesh);
      for (i64 q = 0; q < nmesh; ++q) {
        dm_zmat& z = *zs[order[q]];
        const i64 n = z.n;
        if (form == DM_Z64) {
          z.z64.alloc((size_t)z.ld * n * 2);
          ptrs[q] = z.z64.p;
        } else if (form == DM_Z32) {
          z.z32.alloc((size_t)z.ld * n * 2);
          ptrs[q] = z.z32.p;
        } else if (form == DM_Z64P) {
          z.z64p.alloc(packed_scalars(n));
          zero_packed_padding(z.n, z.z64p.p, sizeof(double), s);
          ptrs[q] = z.z64p.p;
        } else {
          z.z32p.alloc(packed_scalars(n));
          zero_packed_padding(z.n, z.z32p.p, sizeof(float), s);
          ptrs[q] = z.z32p.p;
        }
      }
      ProfScope p(ctx, "assemble_multi", s);
      launch_assemble_multi(oargs, ptrs, form, cap, s);
    }
    check_flag(nt, s);
    for (i64 m = 0; m < nmesh; ++m) outs[m] = zs[m].release();
  });
}

int dm_zmat_size(const dm_zmat* z, int64_t* n, int* forms) {
  return guard([&] {
    check_ctx(z, "zmat");
    if (n) *n = z->n;
    if (forms) *forms = z->forms;
  });
}

static void need_form(const dm_zmat* z, int form) {
  DM_REQUIRE(form == DM_Z64 || form == DM_Z32 || form == DM_Z64P || form == DM_Z32P, DM_ERR_ARGUMENT,
             "form must be one of DM_Z64, DM_Z32, DM_Z64P, DM_Z32P");
  DM_REQUIRE(z->forms & form, DM_ERR_ARGUMENT, "requested form was not assembled");
}

int dm_zmat_compress(dm_zmat* z, double eps, int rmax, double* stats) {
  return guard([&] {
    check_ctx(z, "zmat");
    CtxScope sc(z->ctx);
    compress_z32p(z, eps, rmax, z->ctx->stream);
    double st[4];
    compression_stats(z, st, z->ctx->stream);
    if (stats) std::copy(st, st + 4, stats);
  });
}

int dm_zmat_hlevel(dm_zmat* z, int level, double* stats) {
  return guard([&] {
    check_ctx(z, "zmat");
    DM_REQUIRE(z->forms & DM_Z32H, DM_ERR_ARGUMENT, "the matrix has no compressed form");
    DM_REQUIRE(level >= 0 && level <= kHLevels, DM_ERR_ARGUMENT, "compression level out of range");
    CtxScope sc(z->ctx);
    z->hlevel = level;
    if (stats) {
      compression_stats(z, stats, z->ctx->stream, level);
      stats[4] = z->heps * hlevel_factor(level);
    }
  });
}

int dm_zmat_get(const dm_zmat* z, int form, double* out) {
  return guard([&] {
    check_ctx(z, "zmat");
    check_ctx(out, "output buffer");
    need_form(z, form);
    CtxScope sc(z->ctx);
    hipStream_t s = z->ctx->stream;
    i64 n = z->n;
    // column chunks of at most ~1 GiB of staging
    i64 cols = std::max<i64>(1, std::min<i64>(n, (1LL << 26) / std::max<i64>(n, 1)));
    DBuf<double> stage;
    stage.alloc((size_t)2 * n * cols);
    for (i64 c0 = 0; c0 < n; c0 += cols) {
      i64 nc = std::min(cols, n - c0);
      dml::unpermute(n, z->ref(form), c0, nc, z->rwg->d_iofe.p, stage.p, s);
      HIP_OK(hipMemcpyAsync(out + 2 * n * c0, stage.p, 16 * n * nc, hipMemcpyDeviceToHost, s));
    }
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_zmat_entries(const dm_zmat* z, int form, const int64_t* m, const int64_t* nidx, int64_t count, double* out) {
  return guard([&] {
    check_ctx(z, "zmat");
    need_form(z, form);
    if (count <= 0) return;
    for (i64 i = 0; i < count; ++i)
      DM_REQUIRE(m[i] >= 0 && m[i] < z->n && nidx[i] >= 0 && nidx[i] < z->n, DM_ERR_ARGUMENT, "entry index out of range");
    CtxScope sc(z->ctx);
    hipStream_t s = z->ctx->stream;
    DBuf<i64> dmi, dni;
    DBuf<double> o;
    dmi.upload(m, count, s);
    dni.upload(nidx, count, s);
    o.alloc(2 * count);
    dml::entries(z->ref(form), count, z->rwg->d_iofe.p, dmi.p, dni.p, o.p, s);
    HIP_OK(hipMemcpyAsync(out, o.p, 16 * count, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_zmat_free(dm_zmat* z) {
  return guard([&] {
    if (!z) return;
    std::lock_guard<std::mutex> lk(z->ctx->mu);
    (void)hipSetDevice(z->ctx->device);
    delete z;
  });
}

int dm_apply(dm_ctx* ctx, const dm_zmat* z, int form, int adjoint, const double* X, int64_t ldx, int64_t nrhs,
             double* Y, int64_t ldy) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(z, "zmat");
    if (form == DM_Z32H)
      DM_REQUIRE(z->forms & DM_Z32H, DM_ERR_ARGUMENT, "requested form was not assembled");
    else
      need_form(z, form);
    i64 n = z->n;
    DM_REQUIRE(nrhs >= 0 && ldx >= n && ldy >= n, DM_ERR_DIMENSION, "leading dimension smaller than N");
    CtxScope sc(ctx);
    hipStream_t s = ctx->stream;
    // >= 8 right-hand sides on a full form: one Z X product on the matrix
    // cores per block of columns (Z streamed once per 64 columns); otherwise
    // one GEMV (full forms) or symmetric product (packed forms) per column
    auto op = make_op(z, form);
    for (i64 j0 = 0; j0 < nrhs; j0 += kMfmaMaxRhs) {
      int R = (int)std::min<i64>(kMfmaMaxRhs, nrhs - j0);
      DBuf<double> xe, xi, yi;
      DBuf<char> ws;
      xe.alloc(2 * n * R);
      xi.alloc(2 * n * R);
      yi.alloc(2 * n * R);
      ws.alloc(std::max<size_t>(op->batch_ws(R), 1));
      HIP_OK(hipMemcpy2DAsync(xe.p, 16 * n, X + 2 * ldx * j0, 16 * ldx, 16 * n, R, hipMemcpyHostToDevice, s));
      if (adjoint) dml::conj(n * R, xe.p, s);
      for (int j = 0; j < R; ++j) launch_permute(n, z->rwg->d_ext.p, xe.p + 2 * n * j, xi.p + 2 * n * j, 1, s);
      op->apply_batch(xi.p, yi.p, R, ws.p, s);
      for (int j = 0; j < R; ++j) launch_permute(n, z->rwg->d_ext.p, yi.p + 2 * n * j, xe.p + 2 * n * j, 0, s);
      if (adjoint) dml::conj(n * R, xe.p, s);
      HIP_OK(hipMemcpy2DAsync(Y + 2 * ldy * j0, 16 * ldy, xe.p, 16 * n, 16 * n, R, hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
    }
  });
}

int dm_apply_noalloc(dm_ctx* ctx, const dm_zmat* z, int form, int adjoint, const double* x, double* y) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(z, "zmat");
    DM_REQUIRE(x != nullptr && y != nullptr, DM_ERR_ARGUMENT, "x and y must be device pointers");
    DM_REQUIRE(z->ctx == ctx, DM_ERR_ARGUMENT, "matrix belongs to another context");
    DM_REQUIRE(form == DM_Z64 || form == DM_Z32 || form == DM_Z64P || form == DM_Z32P, DM_ERR_ARGUMENT,
               "dm_apply_noalloc takes one stored form: DM_Z64, DM_Z32, DM_Z64P or DM_Z32P");
    need_form(z, form);
    const i64 n = z->n;
    const bool packed = form == DM_Z64P || form == DM_Z32P;
    // scratch layout: x and y in internal order, then the packed product's partials
    const size_t vec = (size_t)16 * n;
    const size_t vec_al = (vec + 255) & ~(size_t)255;
    const size_t need = 2 * vec_al + (packed ? dml::symv_workspace_bytes(n) : 0);
    if (need > kScratchBytes)
      throw DmError(DM_ERR_ARGUMENT, "dm_apply_noalloc: N = " + std::to_string(n) + " needs " + std::to_string(need) +
                                         " bytes of scratch for this form, more than its " +
                                         std::to_string(kScratchBytes) + " bytes");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_OK(hipSetDevice(ctx->device));
    unsigned char* base = static_cast<unsigned char*>(scratch_of(ctx->device));
    DM_REQUIRE(base != nullptr, DM_ERR_DEVICE, "scratch of this device was not set up (dm_create)");
    double* xi = reinterpret_cast<double*>(base);
    double* yi = reinterpret_cast<double*>(base + vec_al);
    void* ws = base + 2 * vec_al;
    hipStream_t s = ctx->stream;
    const int* ext = z->rwg->d_ext.p;
    launch_permute(n, ext, x, xi, 1, s);
    if (adjoint) dml::conj(n, xi, s);
    const ZRef zr = z->ref(form);
    if (form == DM_Z64)
      launch_gemv_z64(n, z->ld, z->z64.p, xi, yi, s);
    else if (form == DM_Z32)
      launch_gemv_z32(n, z->ld, z->z32.p, xi, yi, s);
    else
      dml::symv_packed(n, zr, xi, yi, ws, dmt::symv_shape_known(ctx, z, form), s);
    launch_permute(n, ext, yi, y, 0, s);
    if (adjoint) dml::conj(n, y, s);
    HIP_OK(hipGetLastError());
  });
}

int dm_mf_create(dm_ctx* ctx, const dm_rwg* rwg, double k_re, double k_im, double eta0, int quad_order, dm_mfop** out) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(rwg, "rwg");
    check_ctx(out, "output handle");
    validate_prefactors(k_re, k_im, eta0);
    CtxScope sc(ctx);
    dmh::tables(const_cast<dm_rwg&>(*rwg), quad_order);
    auto m = std::make_unique<dm_mfop>();
    m->ctx = ctx;
    m->rwg = rwg;
    m->k_re = k_re;
    m->k_im = k_im;
    m->eta0 = eta0;
    m->quad_order = quad_order;
    *out = m.release();
  });
}

int dm_mf_apply(dm_ctx* ctx, const dm_mfop* op, int adjoint, const double* x, double* y) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(op, "operator");
    CtxScope sc(ctx);
    dm_rwg& r = const_cast<dm_rwg&>(*op->rwg);
    const TriTables& T = dmh::tables(r, op->quad_order);
    hipStream_t s = ctx->stream;
    i64 n = r.n;
    NearTables nt;
    AsmArgs a = make_args(r, T, op->k_re, op->k_im, op->eta0, 0, nt);
    DBuf<double> xe, xi, yi;
    xe.alloc(2 * n);
    xi.alloc(2 * n);
    yi.alloc(2 * n);
    HIP_OK(hipMemcpyAsync(xe.p, x, 16 * n, hipMemcpyHostToDevice, s));
    if (adjoint) dml::conj(n, xe.p, s);
    launch_permute(n, r.d_ext.p, xe.p, xi.p, 1, s);
    HIP_OK(hipMemsetAsync(yi.p, 0, 16 * n, s));
    launch_mf_apply(a, xi.p, yi.p, r.max_blk_tris, s);
    launch_permute(n, r.d_ext.p, yi.p, xe.p, 0, s);
    if (adjoint) dml::conj(n, xe.p, s);
    HIP_OK(hipMemcpyAsync(y, xe.p, 16 * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_mf_free(dm_mfop* op) {
  return guard([&] { delete op; });
}

// Plane-wave parameter checks: src/assembly/Excitation.jl:1549-1598
static void validate_planewave(const double* k, const double* p, double E0) {
  for (int c = 0; c < 3; ++c)
    DM_REQUIRE(std::isfinite(k[c]) && std::isfinite(p[c]), DM_ERR_ARGUMENT, "plane-wave k_vec and polarization must be finite.");
  double kn = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
  DM_REQUIRE(kn > 0 && std::isfinite(kn), DM_ERR_ARGUMENT, "plane-wave k_vec must have a finite, nonzero norm.");
  DM_REQUIRE(std::isfinite(E0), DM_ERR_ARGUMENT, "plane-wave amplitude E0 must be finite.");
  double pn = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
  DM_REQUIRE(pn > 0, DM_ERR_ARGUMENT, "plane-wave polarization must have a finite, nonzero norm.");
  DM_REQUIRE(std::fabs(pn - 1.0) <= 1e-12 + 1e-10 * std::max(pn, 1.0), DM_ERR_ARGUMENT,
             "plane-wave polarization must be a unit vector; norm=" + std::to_string(pn) + ".");
  double tr = std::fabs((k[0] * p[0] + k[1] * p[1] + k[2] * p[2]) / (kn * pn));
  DM_REQUIRE(tr <= 1e-10, DM_ERR_ARGUMENT, "plane-wave polarization must be transverse to k_vec.");
}

int dm_excitation_planewave(dm_ctx* ctx, const dm_rwg* rwg, const double* kvec, const double* pol, const double* E0,
                            int64_t M, int quad_order, double* V) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(rwg, "rwg");
    DM_REQUIRE(M >= 0, DM_ERR_ARGUMENT, "negative excitation count");
    for (i64 m = 0; m < M; ++m) validate_planewave(kvec + 3 * m, pol + 3 * m, E0[m]);
    if (M == 0) return;
    CtxScope sc(ctx);
    dm_rwg& r = const_cast<dm_rwg&>(*rwg);
    const TriTables& T = dmh::tables(r, quad_order);
    hipStream_t s = ctx->stream;
    DBuf<double> dk, dp, de, dv;
    dk.upload(kvec, 3 * M, s);
    dp.upload(pol, 3 * M, s);
    de.upload(E0, M, s);
    dv.alloc((size_t)2 * r.n * M);
    dmf::planewave(r, T, dk.p, dp.p, de.p, (int)M, r.d_ext.p, dv.p, r.n, s);
    HIP_OK(hipMemcpyAsync(V, dv.p, 16 * r.n * M, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_radiation_vectors(dm_ctx* ctx, const dm_rwg* rwg, const double* rhat, int64_t nomega, double k, double eta0,
                         int quad_order, double* G) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(rwg, "rwg");
    DM_REQUIRE(nomega > 0, DM_ERR_ARGUMENT, "NΩ must be positive");
    DM_REQUIRE(std::isfinite(k) && std::isfinite(eta0), DM_ERR_ARGUMENT, "radiation_vectors wavenumber and eta0 must be finite");
    for (i64 q = 0; q < nomega; ++q) {
      double nn = rhat[3 * q] * rhat[3 * q] + rhat[3 * q + 1] * rhat[3 * q + 1] + rhat[3 * q + 2] * rhat[3 * q + 2];
      DM_REQUIRE(std::isfinite(nn) && nn > 0, DM_ERR_ARGUMENT, "far-field direction must be finite and nonzero");
    }
    CtxScope sc(ctx);
    dm_rwg& r = const_cast<dm_rwg&>(*rwg);
    const TriTables& T = dmh::tables(r, quad_order);
    hipStream_t s = ctx->stream;
    DBuf<double> dr, dg;
    dr.upload(rhat, 3 * nomega, s);
    dg.alloc((size_t)2 * 3 * nomega * r.n);
    dmf::radiation_vectors(r, T, dr.p, nomega, k, eta0, dg.p, s);
    HIP_OK(hipMemcpyAsync(G, dg.p, dg.bytes(), hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_farfield(dm_ctx* ctx, const dm_rwg* rwg, const double* rhat, int64_t nomega, double k, double eta0,
                int quad_order, const double* I, int64_t ldi, int64_t ncur, double* E, int64_t lde) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(rwg, "rwg");
    check_ctx(I, "currents");
    check_ctx(E, "output");
    DM_REQUIRE(nomega > 0 && ncur >= 1, DM_ERR_ARGUMENT, "NΩ and the current count must be positive");
    DM_REQUIRE(ldi >= rwg->n && lde >= 3 * nomega, DM_ERR_DIMENSION, "leading dimension too small");
    DM_REQUIRE(std::isfinite(k) && std::isfinite(eta0), DM_ERR_ARGUMENT, "wavenumber and eta0 must be finite");
    for (i64 q = 0; q < nomega; ++q) {
      double nn = rhat[3 * q] * rhat[3 * q] + rhat[3 * q + 1] * rhat[3 * q + 1] + rhat[3 * q + 2] * rhat[3 * q + 2];
      DM_REQUIRE(std::isfinite(nn) && nn > 0, DM_ERR_ARGUMENT, "far-field direction must be finite and nonzero");
    }
    CtxScope sc(ctx);
    dm_rwg& r = const_cast<dm_rwg&>(*rwg);
    const TriTables& T = dmh::tables(r, quad_order);
    hipStream_t s = ctx->stream;
    const i64 n = r.n;
    // radiation vectors of a chunk of directions (<= ~1 GiB), then E = G I on
    // the matrix cores, column by column of currents in blocks of 1024
    const i64 chunk = std::max<i64>(1, std::min<i64>(nomega, (1LL << 30) / (48 * std::max<i64>(n, 1))));
    const int RB = (int)std::min<i64>(ncur, 1024);
    DBuf<double> dr, dg, di, de, ws;
    dr.upload(rhat, 3 * nomega, s);
    dg.alloc((size_t)2 * 3 * chunk * n);
    di.alloc((size_t)2 * n * RB);
    de.alloc((size_t)2 * 3 * chunk * RB);
    ws.alloc(dml::gemm_c128_workspace_bytes(n, RB) / sizeof(double));
    for (i64 j0 = 0; j0 < ncur; j0 += RB) {
      const int R = (int)std::min<i64>(RB, ncur - j0);
      HIP_OK(hipMemcpy2DAsync(di.p, 16 * n, I + 2 * ldi * j0, 16 * ldi, 16 * n, R, hipMemcpyHostToDevice, s));
      for (i64 q0 = 0; q0 < nomega; q0 += chunk) {
        const i64 nq = std::min(chunk, nomega - q0);
        {
          ProfScope p(ctx, "radiation", s);
          dmf::radiation_vectors(r, T, dr.p + 3 * q0, nq, k, eta0, dg.p, s);
        }
        {
          ProfScope p(ctx, "farfield_gemm", s);
          dml::gemm_c128(3 * nq, n, dg.p, 3 * nq, di.p, n, R, de.p, 3 * nq, ws.p, s);
        }
        HIP_OK(hipMemcpy2DAsync(E + 2 * (lde * j0 + 3 * q0), 16 * lde, de.p, 16 * 3 * nq, 16 * 3 * nq, R,
                                hipMemcpyDeviceToHost, s));
      }
    }
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_precond_create(dm_ctx* ctx, const dm_zmat* z, int kind, double nf_cutoff, int block_size, int form,
                      dm_precond** out) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(z, "zmat");
    check_ctx(out, "output handle");
    need_form(z, form);
    DM_REQUIRE(kind == DM_PRECOND_DIAG || kind == DM_PRECOND_NEARFIELD || kind == DM_PRECOND_BLOCK, DM_ERR_ARGUMENT,
               "preconditioner kind must be DM_PRECOND_DIAG, DM_PRECOND_NEARFIELD or DM_PRECOND_BLOCK");
    CtxScope sc(ctx);
    dm_solve_opts o;
    dm_solve_opts_default(&o);
    o.precond = kind;
    o.nf_cutoff = nf_cutoff;
    o.block_size = block_size;
    auto h = std::make_unique<dm_precond>();
    h->ctx = ctx;
    h->z = z;
    h->p = build_precond(z, form, o, ctx->stream);
    *out = h.release();
  });
}

int dm_nearfield_pattern_info(const dm_rwg* rwg, double nf_cutoff, int64_t* nnz, double* mean_radius,
                              int* max_cols, int64_t* capped_cols) {
  return guard([&] {
    check_ctx(rwg, "rwg");
    DM_REQUIRE(std::isfinite(nf_cutoff), DM_ERR_ARGUMENT, "near-field cutoff must be finite");
    CtxScope sc(rwg->ctx);
    dm_rwg& r = const_cast<dm_rwg&>(*rwg);
    const double cutJ = nf_cutoff > 0 ? nf_cutoff : 2.5 * r.mean_len;
    auto pat = dmh::spai_pattern(r, cutJ, dmp::spai_max_cols());
    if (nnz) *nnz = pat->nnz;
    if (mean_radius) *mean_radius = pat->mean_radius;
    if (max_cols) *max_cols = dmp::spai_max_cols();
    if (capped_cols) *capped_cols = pat->capped;
  });
}

int dm_precond_apply(dm_ctx* ctx, const dm_precond* P, const double* x, double* y) {
  return guard([&] {
    check_ctx(ctx, "context");
    check_ctx(P, "preconditioner");
    CtxScope sc(ctx);
    hipStream_t s = ctx->stream;
    const dm_rwg& r = *P->z->rwg;
    i64 n = r.n;
    DBuf<double> xe, xi, yi;
    xe.alloc(2 * n);
    xi.alloc(2 * n);
    yi.alloc(2 * n);
    HIP_OK(hipMemcpyAsync(xe.p, x, 16 * n, hipMemcpyHostToDevice, s));
    launch_permute(n, r.d_ext.p, xe.p, xi.p, 1, s);
    P->p->apply(xi.p, yi.p, s);
    launch_permute(n, r.d_ext.p, yi.p, xe.p, 0, s);
    HIP_OK(hipMemcpyAsync(y, xe.p, 16 * n, hipMemcpyDeviceToHost, s));
    HIP_OK(hipStreamSynchronize(s));
  });
}

int dm_precond_free(dm_precond* P) {
  return guard([&] {
    if (!P) return;
    std::lock_guard<std::mutex> lk(P->ctx->mu);
    (void)hipSetDevice(P->ctx->device);
    delete P;
  });
}

void dm_solve_opts_default(dm_solve_opts* o) {
  if (!o) return;
  o->method = DM_METHOD_GMRES;
  o->compress_tol = 0.0;
  o->precond = DM_PRECOND_NEARFIELD;
  o->precond_side = 1;
  o->maxiter = 300;
  o->tol = 1e-6;
  o->nf_cutoff = 0;
  o->restart = 0;
  o->check_convergence = 1;
  o->inner_tol = 0;
  o->max_outer = 10;
  o->block_size = 128;
  o->batch = 0;
  o->learned = nullptr;
  o->adjoint = 0;
  o->storage = 0;
  o->graph = 0;
  o->relax = 0.0;
}

int dm_solve_scattering_pw(dm_ctx* ctx, const dm_rwg* rwg_c, double k_re, double k_im, double eta0, int quad_order,
                           const double* kvec, const double* pol, const double* E0, int64_t nrhs,
                           const dm_solve_opts* opts, double* X, dm_solve_stats* stats, double* assembly_time_s,
                           dm_zmat** zout) {
  bool unconverged = false;
  int rc = guard([&] {
    check_ctx(ctx, "context");
    check_ctx(rwg_c, "rwg");
    DM_REQUIRE(rwg_c->ctx == ctx, DM_ERR_ARGUMENT, "RWG data belongs to another context");
    dm_solve_opts o;
    dm_solve_opts_default(&o);
    if (opts) o = *opts;
    validate_opts(o);
    DM_REQUIRE(o.method != DM_METHOD_LU || nrhs >= 0, DM_ERR_ARGUMENT, "bad RHS count");
    DM_REQUIRE(nrhs >= 1, DM_ERR_ARGUMENT, "at least one plane wave is required");
    validate_prefactors(k_re, k_im, eta0);
    for (i64 m = 0; m < nrhs; ++m) validate_planewave(kvec + 3 * m, pol + 3 * m, E0[m]);
    CtxScope sc(ctx);
    dm_rwg& r = const_cast<dm_rwg&>(*rwg_c);
    i64 n = r.n;
    const TriTables& T = dmh::tables(r, quad_order);
    hipStream_t s = ctx->stream;
    double t0 = now();
    // the matrix buffers of this handle's previous solve, when it kept them
    // (the previous call ended with the stream synchronised)
    std::unique_ptr<dm_zmat> z = std::move(r.spare);
    if (!z || z->n != n) z = std::make_unique<dm_zmat>();
    z->ctx = ctx;
    z->rwg = rwg_c;
    z->n = n;
    z->ld = round_ld(n);
    // packed symmetric forms for one-column solves (half the bytes per
    // product), full forms for lockstep batches (the MFMA product reads full
    // storage) and dense LU; the FP64 GMRES path packs unless asked for full
    // storage
    const bool batched = batched_solve(o, nrhs);
    if (o.method == DM_METHOD_MIXED)
      z->forms = batched ? (DM_Z64 | DM_Z32) : (DM_Z64P | DM_Z32P);
    else
      z->forms = (o.method == DM_METHOD_LU || batched || o.storage == 1) ? DM_Z64 : DM_Z64P;
    // buffers of a kept matrix that this solve does not use
    if (!(z->forms & DM_Z64)) z->z64.release();
    if (!(z->forms & DM_Z32)) z->z32.release();
    if (!(z->forms & DM_Z64P)) z->z64p.release();
    if (!(z->forms & DM_Z32P)) z->z32p.release();
    z->hrmax = 0;
    z->hbytes = 0;
    z->k_re = k_re;
    z->k_im = k_im;
    z->eta0 = eta0;
    z->quad_order = quad_order;
    NearTables nt;
    AsmArgs a;
    {
      ProfScope p(ctx, "near_tables", s);
      a = make_args(r, T, k_re, k_im, eta0, z->ld, nt);
    }
    if (o.method == DM_METHOD_MIXED) {
      // one FP64 assembly writes both forms: the complex128 matrix for the
      // refinement residual and its complex64 rounding for the inner solve
      ProfScope p(ctx, "assemble_z64_z32", s);
      if (batched) {
        z->z64.alloc((size_t)z->ld * n * 2);
        z->z32.alloc((size_t)z->ld * n * 2);
        a.z32_dual = z->z32.p;
        asm_tuned(ctx, a, z->z64.p, r.max_blk_tris, n, false, s);
      } else {
        z->z64p.alloc(packed_scalars(n));
        z->z32p.alloc(packed_scalars(n));
        zero_packed_padding(n, z->z64p.p, sizeof(double), s);
        zero_packed_padding(n, z->z32p.p, sizeof(float), s);
        a.z32_dual = z->z32p.p;
        asm_tuned(ctx, a, z->z64p.p, r.max_blk_tris, n, true, s);
      }
      a.z32_dual = nullptr;
    } else if (z->forms & DM_Z64P) {
      z->z64p.alloc(packed_scalars(n));
      zero_packed_padding(n, z->z64p.p, sizeof(double), s);
      ProfScope p(ctx, "assemble_z64", s);
      asm_tuned(ctx, a, z->z64p.p, r.max_blk_tris, n, true, s);
    } else {
      z->z64.alloc((size_t)z->ld * n * 2);
      ProfScope p(ctx, "assemble_z64", s);
      asm_tuned(ctx, a, z->z64.p, r.max_blk_tris, n, false, s);
    }
    // compression of the complex64 matrix on the second stream, overlapped
    // with the preconditioner build (both only read the packed matrix)
    hipEvent_t e_asm = nullptr, e_h = nullptr;
    struct EvGuard {
      hipEvent_t& a;
      hipEvent_t& b;
      ~EvGuard() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
      }
    } evg{e_asm, e_h};
    const bool compress = o.method == DM_METHOD_MIXED && !batched && o.compress_tol > 0;
    if (!compress) {
      z->hrank.release();
      z->hfac.release();
      z->hwork.release();
    }
    if (compress) {
      HIP_OK(hipEventCreateWithFlags(&e_asm, hipEventDisableTiming));
      HIP_OK(hipEventCreateWithFlags(&e_h, hipEventDisableTiming));
      HIP_OK(hipEventRecord(e_asm, s));
      HIP_OK(hipStreamWaitEvent(ctx->stream2, e_asm, 0));
      {
        ProfScope pc(ctx, "compress", ctx->stream2);
        compress_z32p(z.get(), o.compress_tol, kAcaRankCap, ctx->stream2);
      }
      HIP_OK(hipEventRecord(e_h, ctx->stream2));
    }
    DBuf<double> dk, dp, de, V, x, xe;
    dk.upload(kvec, 3 * nrhs, s);
    dp.upload(pol, 3 * nrhs, s);
    de.upload(E0, nrhs, s);
    V.alloc((size_t)2 * n * nrhs);
    x.alloc(2 * n);
    xe.alloc(2 * n);
    dmf::planewave(r, T, dk.p, dp.p, de.p, (int)nrhs, nullptr, V.p, n, s);
    HIP_OK(hipStreamSynchronize(s));
    double t_asm = now() - t0;
    dm_solve_stats st{};
    std::unique_ptr<DevOp> A64 = make_op(z.get(), pick_form(z.get(), false, batched)), A32;
    if (o.method == DM_METHOD_MIXED) A32 = make_op(z.get(), op_form(z.get(), batched));
    double t1 = now();
    double worst = 0;
    bool all_ok = true;
    if (o.method == DM_METHOD_LU) {
      DBuf<double> A;
      DBuf<int> ipiv, status;
      A.alloc((size_t)2 * z->ld * n);
      HIP_OK(hipMemcpyAsync(A.p, z->z64.p, A.bytes(), hipMemcpyDeviceToDevice, s));
      ipiv.alloc(n);
      status.alloc(1);
      HIP_OK(hipMemsetAsync(status.p, 0, sizeof(int), s));
      dml::lu_factor(n, z->ld, A.p, ipiv.p, status.p, s);
      int sv = 0;
      HIP_OK(hipMemcpyAsync(&sv, status.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      DM_REQUIRE(sv == 0, DM_ERR_ARGUMENT, "dense LU: matrix is singular");
      st.precond_time_s = now() - t1;
      DBuf<double> tmp, hb;
      tmp.alloc(2 * n);
      hb.alloc(4);
      for (i64 j = 0; j < nrhs; ++j) {
        double* bj = V.p + 2 * n * j;
        HIP_OK(hipMemcpyAsync(x.p, bj, 16 * n, hipMemcpyDeviceToDevice, s));
        dml::lu_solve(n, z->ld, A.p, ipiv.p, x.p, n, 1, s);
        worst = std::max(worst, true_residual(*A64, n, bj, x.p, tmp.p, hb.p, s));
        launch_permute(n, r.d_ext.p, x.p, xe.p, 0, s);
        HIP_OK(hipMemcpyAsync(X + 2 * n * j, xe.p, 16 * n, hipMemcpyDeviceToHost, s));
      }
    } else {
      auto M = build_precond(z.get(), pick_form(z.get(), o.method == DM_METHOD_MIXED, batched), o, s);
      HIP_OK(hipStreamSynchronize(s));
      st.precond_time_s = now() - t1;
      if (compress) HIP_OK(hipStreamWaitEvent(s, e_h, 0));
      solve_columns(ctx, o, r, A64.get(), A32.get(), M.get(), V.p, n, nrhs, X, n, st, worst, all_ok, nullptr);
    }
    HIP_OK(hipStreamSynchronize(s));
    check_flag(nt, s);
    st.solved = all_ok ? 1 : 0;
    st.residual = worst;
    st.solve_time_s = now() - t1 - st.precond_time_s;
    if (stats) *stats = st;
    if (assembly_time_s) *assembly_time_s = t_asm;
    if (zout)
      *zout = z.release();
    else
      r.spare = std::move(z);
    if (!all_ok && o.check_convergence) unconverged = true;
  });
  if (rc == DM_OK && unconverged) {
    g_err = "GMRES did not converge within maxiter";
    return DM_ERR_CONVERGENCE;
  }
  return rc;
}

int dm_solve(dm_ctx* ctx, const dm_zmat* z, const double* B, int64_t ldb, int64_t nrhs, const dm_solve_opts* opts,
             double* X, int64_t ldx, dm_solve_stats* stats) {
  bool unconverged = false;
  int rc = guard([&] {
    check_ctx(ctx, "context");
    check_ctx(z, "zmat");
    dm_solve_opts o;
    dm_solve_opts_default(&o);
    if (opts) o = *opts;
    i64 n = z->n;
    DM_REQUIRE(nrhs >= 0 && ldb >= n && ldx >= n, DM_ERR_DIMENSION, "leading dimension smaller than N");
    validate_opts(o);
    for (i64 j = 0; j < nrhs; ++j)
      for (i64 i = 0; i < 2 * n; ++i)
        DM_REQUIRE(std::isfinite(B[2 * ldb * j + i]), DM_ERR_ARGUMENT, "right-hand side must be finite");
    CtxScope sc(ctx);
    hipStream_t s = ctx->stream;
    const dm_rwg& r = *z->rwg;
    dm_solve_stats st{};
    DBuf<double> be, b, x, tmp, hb;
    be.alloc(2 * n);
    b.alloc(2 * n);
    x.alloc(2 * n);
    tmp.alloc(2 * n);
    hb.alloc(4);
    double t0 = now();
    if (o.method == DM_METHOD_LU) {
      DM_REQUIRE(z->forms & DM_Z64, DM_ERR_ARGUMENT, "dense LU needs the complex128 matrix");
      DBuf<double> A, bb;
      DBuf<int> ipiv, status;
      A.alloc((size_t)2 * z->ld * n);
      HIP_OK(hipMemcpyAsync(A.p, z->z64.p, A.bytes(), hipMemcpyDeviceToDevice, s));
      ipiv.alloc(n);
      status.alloc(1);
      HIP_OK(hipMemsetAsync(status.p, 0, sizeof(int), s));
      dml::lu_factor(n, z->ld, A.p, ipiv.p, status.p, s);
      int sv = 0;
      HIP_OK(hipMemcpyAsync(&sv, status.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_OK(hipStreamSynchronize(s));
      DM_REQUIRE(sv == 0, DM_ERR_ARGUMENT, "dense LU: matrix is singular (zero pivot " + std::to_string(sv) + ")");
      st.precond_time_s = now() - t0;
      OpZ64 A64(z);
      double worst = 0;
      for (i64 j = 0; j < nrhs; ++j) {
        HIP_OK(hipMemcpyAsync(be.p, B + 2 * ldb * j, 16 * n, hipMemcpyHostToDevice, s));
        if (o.adjoint) dml::conj(n, be.p, s);
        launch_permute(n, r.d_ext.p, be.p, b.p, 1, s);
        HIP_OK(hipMemcpyAsync(x.p, b.p, 16 * n, hipMemcpyDeviceToDevice, s));
        dml::lu_solve(n, z->ld, A.p, ipiv.p, x.p, n, 1, s);
        worst = std::max(worst, true_residual(A64, n, b.p, x.p, tmp.p, hb.p, s));
        launch_permute(n, r.d_ext.p, x.p, be.p, 0, s);
        if (o.adjoint) dml::conj(n, be.p, s);
        HIP_OK(hipMemcpyAsync(X + 2 * ldx * j, be.p, 16 * n, hipMemcpyDeviceToHost, s));
      }
      HIP_OK(hipStreamSynchronize(s));
      st.solved = 1;
      st.residual = worst;
      st.solve_time_s = now() - t0 - st.precond_time_s;
      if (stats) *stats = st;
      return;
    }
    const bool batched = batched_solve(o, nrhs);
    const int f64 = pick_form(z, false, batched), f32 = pick_form(z, true, batched);
    if (o.method == DM_METHOD_MIXED)
      DM_REQUIRE(f64 && f32, DM_ERR_ARGUMENT, "mixed-precision solve needs a complex128 and a complex64 form");
    const int pform = (o.method == DM_METHOD_MIXED || !f64) ? f32 : f64;
    auto M = build_precond(z, pform, o, s);
    HIP_OK(hipStreamSynchronize(s));
    st.precond_time_s = now() - t0;
    double t1 = now();
    std::unique_ptr<DevOp> A64, A32;
    if (f64) A64 = make_op(z, f64);
    if (f32) A32 = make_op(z, o.method == DM_METHOD_MIXED ? op_form(z, batched) : f32);
    bool all_ok = true;
    double worst = 0;
    DBuf<double> Bd;
    Bd.alloc((size_t)2 * n * std::max<i64>(nrhs, 1));
    for (i64 j = 0; j < nrhs; ++j) {
      HIP_OK(hipMemcpyAsync(be.p, B + 2 * ldb * j, 16 * n, hipMemcpyHostToDevice, s));
      if (o.adjoint) dml::conj(n, be.p, s);
      launch_permute(n, r.d_ext.p, be.p, Bd.p + 2 * n * j, 1, s);
    }
    solve_columns(ctx, o, r, A64.get(), A32.get(), M.get(), Bd.p, n, nrhs, X, ldx, st, worst, all_ok, nullptr,
                  o.adjoint != 0);
    st.solved = all_ok ? 1 : 0;
    st.residual = worst;
    st.solve_time_s = now() - t1;
    if (stats) *stats = st;
    if (!all_ok && o.check_convergence) unconverged = true;
  });
  if (rc == DM_OK && unconverged) {
    g_err = "GMRES did not converge within maxiter";
    return DM_ERR_CONVERGENCE;
  }
  return rc;
}

}  // extern "C"
