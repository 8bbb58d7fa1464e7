// dm_tune.hip — the library's own launch-shape tuner.
//
// A tunable kernel asks for its parameter with a key naming the shape (kernel,
// precision, problem size).  On the first request in a context the candidates
// are timed with HIP events on the calling stream (one warm-up and three timed
// launches each) and the fastest is kept in the context; later requests reuse
// it.  Results can be exported to / imported from a flat JSON object so a
// caller can persist them across processes (dm_tune_get / dm_tune_set).
// DM_TUNE=0 in the environment disables timing and uses the defaults.
#include <algorithm>
#include <cstdlib>
#include <functional>
#include <vector>
#include <cstring>
#include <sstream>
#include <string>

#include "dm_solver.h"

namespace dmt {

namespace {

bool tuning_enabled() {
  const char* e = std::getenv("DM_TUNE");
  return !(e && std::strcmp(e, "0") == 0);
}

template <class Launch>
double time_ms(Launch&& run, hipStream_t s) {
  hipEvent_t a, b;
  HIP_OK(hipEventCreate(&a));
  HIP_OK(hipEventCreate(&b));
  run();  // warm-up
  HIP_OK(hipEventRecord(a, s));
  for (int r = 0; r < 3; ++r) run();
  HIP_OK(hipEventRecord(b, s));
  HIP_OK(hipEventSynchronize(b));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms / 3.0;
}

}  // namespace

int tune_choice(dm_ctx* ctx, const std::string& key, const std::vector<int>& cands, int fallback,
                const std::function<void(int)>& run, hipStream_t s) {
  auto it = ctx->tune.find(key);
  if (it != ctx->tune.end()) {
    DM_REQUIRE(std::find(cands.begin(), cands.end(), it->second) != cands.end(), DM_ERR_ARGUMENT,
               "tuned value out of range for " + key);
    return it->second;
  }
  if (!tuning_enabled()) return fallback;
  int best = fallback;
  double best_ms = 1e300;
  for (int c : cands) {
    const double ms = time_ms([&] { run(c); }, s);
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  ctx->tune[key] = best;
  return best;
}

namespace {
std::string symv_key(const dm_zmat* z, int form) {
  return std::string("symv/") + (form == DM_Z32P ? "c64/" : "c128/") + std::to_string(z->n);
}
constexpr int kSymvFallback = 8;
}  // namespace

int symv_shape_known(const dm_ctx* ctx, const dm_zmat* z, int form) {
  auto it = ctx->tune.find(symv_key(z, form));
  return it != ctx->tune.end() ? it->second : kSymvFallback;
}

int symv_shape(dm_ctx* ctx, const dm_zmat* z, int form, void* ws, hipStream_t s) {
  const int fallback = kSymvFallback;
  const std::string key = symv_key(z, form);
  auto it = ctx->tune.find(key);
  if (it != ctx->tune.end()) return it->second;
  if (!tuning_enabled()) return fallback;
  DBuf<double> x, y;
  x.alloc(2 * z->n);
  y.alloc(2 * z->n);
  HIP_OK(hipMemsetAsync(x.p, 0, x.bytes(), s));
  const ZRef zr = z->ref(form);
  int best = fallback;
  double best_ms = 1e300;
  for (int shape : dml::symv_shapes(z->n, form == DM_Z32P)) {
    double ms = time_ms([&] { dml::symv_packed(z->n, zr, x.p, y.p, ws, shape, s); }, s);
    if (ms < best_ms) {
      best_ms = ms;
      best = shape;
    }
  }
  ctx->tune[key] = best;
  return best;
}

int symv_hier_shape(dm_ctx* ctx, const dm_zmat* z, void* ws, hipStream_t s) {
  const int fallback = 2;
  const std::string key = "symvh/c64/" + std::to_string(z->n);
  auto it = ctx->tune.find(key);
  if (it != ctx->tune.end()) {
    DM_REQUIRE(it->second >= 1 && it->second <= 64, DM_ERR_ARGUMENT, "compressed product wave multiple out of range");
    return it->second;
  }
  if (!tuning_enabled()) return fallback;
  DBuf<double> x, y;
  x.alloc(2 * z->n);
  y.alloc(2 * z->n);
  HIP_OK(hipMemsetAsync(x.p, 0, x.bytes(), s));
  const int slots = dml::hier_wave_slots();
  int best = fallback;
  double best_ms = 1e300;
  for (int mult : {1, 2, 4, 8}) {
    DBuf<char> rb;
    rb.alloc(dml::hier_ranges_bytes(z->n, mult * slots));
    dml::HierRanges hr;
    dml::hier_ranges(z->n, z->hrank.p, mult * slots, rb.p, hr, s);
    double ms = time_ms(
        [&] { dml::symv_hier(z->n, z->z32p.p, z->hrank.p, z->hfac.p, z->hrmax, x.p, y.p, ws, hr, s); }, s);
    if (ms < best_ms) {
      best_ms = ms;
      best = mult;
    }
  }
  ctx->tune[key] = best;
  return best;
}

}  // namespace dmt

extern "C" {

int dm_tune_get(dm_ctx* ctx, char* buf, int64_t cap, int64_t* len) {
  if (!ctx || !len) {
    dm_set_error("NULL argument");
    return DM_ERR_ARGUMENT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  std::ostringstream o;
  o << "{";
  bool first = true;
  for (auto& kv : ctx->tune) {
    o << (first ? "" : ", ") << "\"" << kv.first << "\": " << kv.second;
    first = false;
  }
  o << "}";
  const std::string js = o.str();
  *len = (int64_t)js.size();
  if (buf && cap > 0) {
    const size_t m = std::min<size_t>((size_t)cap - 1, js.size());
    std::memcpy(buf, js.data(), m);
    buf[m] = 0;
  }
  return DM_OK;
}

// Flat {"key": int, ...} objects only (what dm_tune_get writes).
int dm_tune_set(dm_ctx* ctx, const char* json) {
  if (!ctx || !json) {
    dm_set_error("NULL argument");
    return DM_ERR_ARGUMENT;
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  const std::string s(json);
  size_t p = 0;
  std::map<std::string, int> got;
  while ((p = s.find('"', p)) != std::string::npos) {
    size_t q = s.find('"', p + 1);
    size_t c = q == std::string::npos ? q : s.find(':', q);
    if (c == std::string::npos) {
      dm_set_error("malformed tuning record");
      return DM_ERR_ARGUMENT;
    }
    const std::string key = s.substr(p + 1, q - p - 1);
    char* end = nullptr;
    long v = std::strtol(s.c_str() + c + 1, &end, 10);
    if (end == s.c_str() + c + 1) {
      dm_set_error("malformed tuning record");
      return DM_ERR_ARGUMENT;
    }
    got[key] = (int)v;
    p = (size_t)(end - s.c_str());
  }
  for (auto& kv : got) ctx->tune[kv.first] = kv.second;
  return DM_OK;
}

}  // extern "C"
