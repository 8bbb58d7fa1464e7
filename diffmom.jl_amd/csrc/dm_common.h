// dm_common.h — shared internals of libdiffmom (host + device).
#pragma once
#include <cmath>
#include <hip/hip_runtime.h>

#include <cstring>

#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/diffmom.h"

typedef int64_t i64;

// ------------------------------------------------------------------ errors
struct DmError : std::runtime_error {
  int code;
  DmError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] inline void dm_throw(int code, const std::string& msg) { throw DmError(code, msg); }

#define HIP_OK(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      dm_throw(e_ == hipErrorOutOfMemory ? DM_ERR_MEMORY : DM_ERR_DEVICE,                \
               std::string(#expr) + ": " + hipGetErrorString(e_));                       \
  } while (0)

#define DM_REQUIRE(cond, code, msg) \
  do {                              \
    if (!(cond)) dm_throw(code, msg); \
  } while (0)

// --------------------------------------------------------------- device buffers
// Device memory a context keeps for reuse.  hipFree waits for the whole
// device and hipMalloc of a large block costs ~0.3 ms, so blocks released
// during a public call of a context are kept (up to `cap` bytes) and handed
// out again for requests of exactly the same size.  A released block first
// waits in `pending`: it becomes reusable only once both of the context's
// streams have been seen idle, so no queued kernel can still touch it.
// dm_destroy frees everything kept; a failed hipMalloc frees the reusable
// blocks and retries.
struct DevCache {
  int device = -1;
  size_t cap = 0, bytes = 0;
  std::multimap<size_t, void*> free_blocks;
  std::vector<std::pair<size_t, void*>> pending;
  void* take(size_t b);
  bool give(void* p, size_t b);
  void promote(hipStream_t a, hipStream_t b);  // pending -> reusable when both streams are idle
  void trim();                                 // free the reusable blocks
  void clear();                                // free everything
};
// the cache of the context whose public call runs on this thread (or null)
extern thread_local DevCache* tl_cache;

template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  int dev = -1;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      const size_t b = n * sizeof(T);
      if (!(tl_cache && tl_cache->device == dev && tl_cache->give(p, b))) (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    const size_t b = count * sizeof(T);
    void* q = tl_cache ? tl_cache->take(b) : nullptr;
    if (q) {
      p = static_cast<T*>(q);
      dev = tl_cache->device;
    } else {
      hipError_t e = hipMalloc(&p, b);
      if (e == hipErrorOutOfMemory && tl_cache) {
        (void)hipGetLastError();
        tl_cache->trim();
        e = hipMalloc(&p, b);
      }
      HIP_OK(e);
      HIP_OK(hipGetDevice(&dev));
    }
    n = count;
  }
  void upload(const T* host, size_t count, hipStream_t s) {
    alloc(count);
    if (count) HIP_OK(hipMemcpyAsync(p, host, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  size_t bytes() const { return n * sizeof(T); }
};

// ------------------------------------------------------------------ handles
struct ProfRec {
  std::string name;
  hipEvent_t a, b;
};

struct dm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;   // main stream: every public call runs here
  hipStream_t stream2 = nullptr;  // second stream for overlapped work
  std::mutex mu;
  int cu_count = 256;
  int prof = 0;  // HIP events around tracked launches: 0 none, 1 all, 2 operator products only
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> ev_free;  // recycled profiling events
  std::map<std::string, int> tune;  // launch parameters chosen by the tuner, by shape key
  DevCache cache;                   // device memory kept for reuse by this context's calls
  hipEvent_t ev_get() {
    if (ev_free.empty()) {
      hipEvent_t e;
      HIP_OK(hipEventCreate(&e));
      return e;
    }
    hipEvent_t e = ev_free.back();
    ev_free.pop_back();
    return e;
  }
};

// Brackets one tracked launch with HIP events on its stream when profiling is on.
struct ProfScope {
  dm_ctx* c;
  hipStream_t s;
  ProfRec rec;
  bool on;
  ProfScope(dm_ctx* c_, const char* name, hipStream_t s_)
      : c(c_), s(s_),
        on(c_ && (c_->prof == 1 ||
                  (c_->prof == 2 && (std::strncmp(name, "symv", 4) == 0 || std::strncmp(name, "gemv", 4) == 0)))) {
    if (!on) return;
    rec.name = name;
    rec.a = c->ev_get();
    rec.b = c->ev_get();
    HIP_OK(hipEventRecord(rec.a, s));
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(rec.b, s);
    c->recs.push_back(rec);
  }
};

// Message returned by dm_last_error() for the calling thread.
void dm_set_error(const std::string& msg);

struct dm_mesh {
  dm_ctx* ctx = nullptr;
  i64 nv = 0, nt = 0;
  std::vector<double> xyz;  // 3 x nv
  std::vector<i64> tri;     // 3 x nt, 1-based
};

// Per-quadrature-order triangle tables (device), built lazily per RWG handle.
struct TriTables {
  int quad_order = 0, nq = 0;
  DBuf<double> cen;   // nt*3 centroid
  DBuf<double> uq;    // nt*nq*3 quad points relative to the centroid
  DBuf<double> av;    // nt*9 vertex offsets from the centroid
  DBuf<double> rq;    // nt*nq*3 absolute quad points
  DBuf<double> rq7;   // nt*7*3 absolute 7-point rule points
  DBuf<double> vtx;   // nt*9 absolute vertices
  DBuf<double> area;  // nt
  std::vector<double> h_rq, h_area;
};

// Near-field sparse-approximate-inverse pattern (internal basis order).
struct SpaiPattern {
  double cutJ = 0, cutI = 0;
  double mean_radius = 0;  // mean over columns of the radius J_j covers
  i64 capped = 0;          // columns whose cap left bases within the cutoff out
  i64 nnz = 0;
  std::vector<int> jptr, jidx, iptr, iidx;   // per column: support J_j, rows I_j
  std::vector<int> rowptr, colidx, map;      // CSR of M and CSR position -> CSC position
  DBuf<int> d_jptr, d_jidx, d_iptr, d_iidx, d_rowptr, d_colidx, d_map;
};

struct dm_zmat;
struct dm_rwg {
  const dm_mesh* mesh = nullptr;
  dm_ctx* ctx = nullptr;
  i64 n = 0;
  // caller-order arrays (1-based triangle / vertex ids)
  std::vector<i64> tplus, tminus, evert, vplus_opp, vminus_opp;
  std::vector<double> len, area_p, area_m, coeff_p, coeff_m;
  // derived
  std::vector<int> eplus, eminus;     // local vertex index (0..2) opposite the edge
  std::vector<int> ext_of_int;        // internal index -> caller index (0-based)
  std::vector<int> int_of_ext;        // caller index -> internal index
  std::vector<int> nb;                // nt*3 edge neighbours (0-based, -1 none)
  i64 adj_pairs = 0;
  // tiles (internal order): blocks of TB consecutive internal RWGs
  int TB = 64;
  int nblk = 0;
  int max_blk_tris = 0;
  std::vector<int> blk_off, blk_tri;  // CSR: triangles touched by each block
  // device copies
  DBuf<int> d_ext;          // n
  DBuf<int> d_iofe;         // n: caller index -> internal index
  DBuf<double> d_len;       // n (internal order)
  DBuf<int> d_hrwg;         // nt*3 internal RWG of half (t, e), -1 if none
  DBuf<double> d_hsig;      // nt*3 sign/coefficient of half (t, e)
  DBuf<int> d_nb;           // nt*3
  DBuf<int> d_blk_off, d_blk_tri;
  DBuf<int> d_tplus_i, d_tminus_i, d_eplus_i, d_eminus_i;  // internal order, 0-based
  TriTables tabs[8];        // indexed by quad order (1,3,4,7)
  std::mutex tab_mu;
  std::vector<double> cen_i;  // internal-order basis centres (3 per basis)
  double mean_len = 0;
  std::shared_ptr<SpaiPattern> spai;  // latest pattern; preconditioners keep their own reference
  // tile-side images per quadrature order and precision (built lazily)
  DBuf<char> side_img[8][2];
  // side images for another mesh's block layout (many-mesh launches), keyed by
  // (2 x quadrature order + FP64 flag, triangles per block of the layout)
  std::map<std::pair<int, int>, DBuf<char>> side_img_cap;
  // the matrix buffers of the last dm_solve_scattering_pw that did not hand
  // its matrix to the caller, kept for the next solve on this handle (same
  // sizes: no allocation, no free at the end of the solve); freed with the
  // handle
  std::unique_ptr<dm_zmat> spare;
};

// ------------------------------------------------------------ Z storage
// Full storage: column-major ld x n, interleaved (re, im).  Packed symmetric
// storage: tiles of kPT x kPT bases (I <= J) at tile index J (J + 1) / 2 + I,
// each tile column-major; entries past n are zero.
constexpr int kPT = 128;
// Looser levels of the compressed form: level l (1 .. kHLevels) bounds the
// error by eps 10^(l/2) with the prefix of the factors the cross
// approximation had at that tolerance (half-decade steps up to 10^5 eps).
constexpr int kHLevels = 10;
inline double hlevel_factor(int l) { return std::pow(10.0, 0.5 * l); }
inline i64 packed_tiles(i64 n) {
  i64 nb = (n + kPT - 1) / kPT;
  return nb * (nb + 1) / 2;
}
inline size_t packed_scalars(i64 n) { return (size_t)packed_tiles(n) * kPT * kPT * 2; }

struct ZRef {
  const void* p = nullptr;
  i64 ld = 0;
  int f32 = 0;     // complex64 storage
  int packed = 0;  // packed symmetric tiles
};

__device__ __host__ inline i64 packed_index(i64 r, i64 c) {
  if (r > c) {
    i64 t = r;
    r = c;
    c = t;
  }
  i64 I = r / kPT, J = c / kPT;
  i64 l = J * (J + 1) / 2 + I;
  return (l * kPT + (c - J * kPT)) * kPT + (r - I * kPT);
}

__device__ inline double2 zat(const ZRef& z, i64 r, i64 c) {
  i64 idx = z.packed ? packed_index(r, c) : r + c * z.ld;
  if (z.f32) {
    float2 v = reinterpret_cast<const float2*>(z.p)[idx];
    return make_double2((double)v.x, (double)v.y);
  }
  return reinterpret_cast<const double2*>(z.p)[idx];
}

struct dm_zmat {
  dm_ctx* ctx = nullptr;
  const dm_rwg* rwg = nullptr;
  i64 n = 0, ld = 0;
  int forms = 0;
  double k_re = 0, k_im = 0, eta0 = 0;
  int quad_order = 3;
  DBuf<double> z64;   // ld x n complex128 (interleaved), internal order
  DBuf<float> z32;    // ld x n complex64, internal order
  DBuf<double> z64p;  // packed symmetric complex128
  DBuf<float> z32p;   // packed symmetric complex64
  // compressed complex64 form over z32p: per stored tile its rank (-1: dense)
  // and a slot of 2 x hrmax columns of 128 complex64 (U then V)
  DBuf<int> hrank;
  DBuf<float> hfac;
  DBuf<double> hwork;  // tile norms and the tolerance
  DBuf<int> hrank_lv;  // kHLevels x T: per level the rank of each tile (-1: dense)
  double heps = 0;     // eps of the compression
  int hlevel = 0;      // level of the products dm_apply runs on the compressed form
  int hrmax = 0;
  double hbytes = 0;  // bytes one product reads
  ZRef ref(int form) const {
    ZRef r;
    r.ld = ld;
    r.f32 = (form == DM_Z32 || form == DM_Z32P) ? 1 : 0;
    r.packed = (form == DM_Z64P || form == DM_Z32P) ? 1 : 0;
    r.p = form == DM_Z64 ? (const void*)z64.p
          : form == DM_Z32 ? (const void*)z32.p
          : form == DM_Z64P ? (const void*)z64p.p
                            : (const void*)z32p.p;
    return r;
  }
};

struct dm_mfop {
  dm_ctx* ctx = nullptr;
  const dm_rwg* rwg = nullptr;
  double k_re = 0, k_im = 0, eta0 = 0;
  int quad_order = 3;
};

// ------------------------------------------------------------ host geometry
namespace dmh {
void build_rwg_canonical(const dm_mesh& m, dm_rwg& r);
void finalize_rwg(dm_rwg& r);  // validates, orders, tiles, uploads
void mesh_check(const dm_mesh& m, int allow_boundary, int require_closed, double area_tol_rel);
double triangle_area(const dm_mesh& m, i64 t0);
const TriTables& tables(dm_rwg& r, int quad_order);
int quad_rule(int order, double* xi, double* w);
std::shared_ptr<const SpaiPattern> spai_pattern(dm_rwg& r, double cutJ, int maxJ);
}  // namespace dmh

namespace dmp {
int spai_max_cols();
// maxJ: widest column support of the pattern (<= spai_max_cols())
void spai_build(i64 n, const ZRef& z, int maxJ, const int* jptr, const int* jidx, const int* iptr, const int* iidx,
                double* vals, int* status, hipStream_t s);
void csr_from_csc(i64 nnz, const int* map, const double* cscval, double* csrval, hipStream_t s);
void spmv(i64 n, const int* rowptr, const int* colidx, const double* val, const double* x, double* y, hipStream_t s);
// Y = S X for R columns of length n (leading dimension n)
void spmm(i64 n, int R, const int* rowptr, const int* colidx, const double* val, const double* X, double* Y,
          hipStream_t s);
}  // namespace dmp

// ------------------------------------------------------------ kernel launchers
struct AsmArgs {
  i64 n, ld, nt;
  int nblk, TB;
  const int* blk_off;
  const int* blk_tri;
  const int* hrwg;
  const double* hsig;
  const int* nb;
  const int* ext;
  const double* len;
  const int* tplus_i;  // internal basis -> T+ triangle (0-based)
  const double* cen;
  const double* uq;
  const double* av;
  const double* jself;  // nt*9 complex
  const double* jadj;   // nt*3*9 complex
  double w[7];
  double k_re, k_im, eta0;
  int nq;
  int* flag;  // set to 1 when an assembled entry is non-finite
  // per-block LDS images of the tile sides (built once per RWG, quadrature
  // order and precision): block b at img + b * stride
  const char* side_img32;
  const char* side_img64;
  i64 side_stride32, side_stride64;
  int side_cap;
  // FP64 assembly only: when set, every entry is also stored rounded to
  // complex64 at the same position of this second matrix (same form)
  float* z32_dual = nullptr;
};

void launch_near_tables(const AsmArgs& a, const TriTables& t, double* jself, double* jadj,
                        hipStream_t s);
// near-field tables of several meshes in one launch per kind: mesh m's tables
// start at jself + 18 x (triangles of meshes 0..m-1), jadj + 54 x the same;
// staging keeps the per-mesh arguments on the device until the work is done
void launch_near_tables_multi(const std::vector<AsmArgs>& a, const std::vector<const TriTables*>& t, double* jself,
                              double* jadj, DBuf<char>& staging, hipStream_t s);
// Bytes of one block's tile-side image and the kernel that fills nblk of them.
size_t side_image_bytes(int nq, bool f64, int cap);
void launch_build_sides(const AsmArgs& a, int cap, bool f64, char* img, hipStream_t s);
// nthr: threads per workgroup of the FP64 tile kernel (0 = 1024, or 512)
void launch_assemble_z64(const AsmArgs& a, double* z, int max_tris, hipStream_t s, int nthr = 0);
void launch_assemble_z32(const AsmArgs& a, float* z, int max_tris, hipStream_t s);
// packed symmetric storage (buffer of packed_scalars(n); entries past n must be
// zeroed beforehand: zero_packed_padding)
void launch_assemble_z64p(const AsmArgs& a, double* z, int max_tris, hipStream_t s, int nthr = 0);
void launch_assemble_z32p(const AsmArgs& a, float* z, int max_tris, hipStream_t s);
// Zero the entries of packed storage that lie past n (only the last tile
// column holds any); every other entry is written by the assembly.
void zero_packed_padding(i64 n, void* z, size_t scalar_bytes, hipStream_t s);
// Several meshes' matrices (one stored form) in one persistent launch;
// packed forms need zeroed buffers.  Synchronises the stream on return.
void launch_assemble_multi(const std::vector<AsmArgs>& args, const std::vector<void*>& z, int form, int max_tris,
                           hipStream_t s);
// matrix-free y += Z x in internal order (FP64), x and y complex128
void launch_mf_apply(const AsmArgs& a, const double* x, double* y, int max_tris, hipStream_t s);

// dense products (internal order, column-major, complex interleaved)
void launch_gemv_z64(i64 n, i64 ld, const double* z, const double* x, double* y, hipStream_t s);
void launch_gemv_z32(i64 n, i64 ld, const float* z, const double* x, double* y, hipStream_t s);

// small vector kernels
void launch_permute(i64 n, const int* idx, const double* src, double* dst, int gather, hipStream_t s);
void launch_cast_c128_to_c64_matrix(i64 n, i64 ld, const double* src, float* dst, hipStream_t s);
