// dm_host.cpp — host-side mesh / RWG preparation for libdiffmom.
//
// Everything here is O(Nt log Nt) set-up work that runs once per RWG handle:
// mesh-quality checks, canonical RWG construction, validation of caller RWG
// arrays, the edge-adjacency table, the spatial (bisection) basis order used by the
// tiled kernels, and the per-quadrature-order triangle tables.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <sstream>
#include <tuple>

#include "dm_common.h"

namespace dmh {

namespace {
struct P3 {
  double x, y, z;
};
inline P3 vert(const dm_mesh& m, i64 v1) {
  const double* p = &m.xyz[3 * (v1 - 1)];
  return {p[0], p[1], p[2]};
}
inline double comp_det(double a, double b, double c, double d) {
  double p1 = a * b, p2 = c * d;
  return (p1 - p2) + (std::fma(a, b, -p1) - std::fma(c, d, -p2));
}
}  // namespace

// Triangle quadrature rules, src/basis/Quadrature.jl:17-54.  The point and
// weight tables must equal the reference's for parity.
int quad_rule(int order, double* xi, double* w) {
  if (order == 1) {
    xi[0] = 1.0 / 3; xi[1] = 1.0 / 3; w[0] = 0.5;
    return 1;
  }
  if (order == 3) {
    const double p[3][2] = {{1.0 / 6, 1.0 / 6}, {2.0 / 3, 1.0 / 6}, {1.0 / 6, 2.0 / 3}};
    for (int q = 0; q < 3; ++q) { xi[2 * q] = p[q][0]; xi[2 * q + 1] = p[q][1]; w[q] = 1.0 / 6; }
    return 3;
  }
  if (order == 4) {
    const double p[4][2] = {{1.0 / 3, 1.0 / 3}, {0.6, 0.2}, {0.2, 0.6}, {0.2, 0.2}};
    const double ww[4] = {-27.0 / 96, 25.0 / 96, 25.0 / 96, 25.0 / 96};
    for (int q = 0; q < 4; ++q) { xi[2 * q] = p[q][0]; xi[2 * q + 1] = p[q][1]; w[q] = ww[q]; }
    return 4;
  }
  if (order == 7) {
    const double a1 = 0.059715871789770, b1 = 0.470142064105115;
    const double a2 = 0.797426985353087, b2 = 0.101286507323456;
    const double p[7][2] = {{1.0 / 3, 1.0 / 3}, {b1, b1}, {a1, b1}, {b1, a1},
                            {b2, b2}, {a2, b2}, {b2, a2}};
    const double w0 = 0.1125, w1 = 0.0661970763942530, w2 = 0.0629695902724135;
    const double ww[7] = {w0, w1, w1, w1, w2, w2, w2};
    for (int q = 0; q < 7; ++q) { xi[2 * q] = p[q][0]; xi[2 * q + 1] = p[q][1]; w[q] = ww[q]; }
    return 7;
  }
  return 0;
}

// triangle_area with exponent scaling and a compensated cross product:
// src/geometry/Mesh.jl:2940-3080,3205-3232 (fast path).
double triangle_area(const dm_mesh& m, i64 t0) {
  const i64* tr = &m.tri[3 * t0];
  P3 v1 = vert(m, tr[0]), v2 = vert(m, tr[1]), v3 = vert(m, tr[2]);
  double e1x = v2.x - v1.x, e1y = v2.y - v1.y, e1z = v2.z - v1.z;
  double e2x = v3.x - v1.x, e2y = v3.y - v1.y, e2z = v3.z - v1.z;
  double s1 = std::max(std::fabs(e1x), std::max(std::fabs(e1y), std::fabs(e1z)));
  double s2 = std::max(std::fabs(e2x), std::max(std::fabs(e2y), std::fabs(e2z)));
  if (!(s1 > 0 && s2 > 0) || !std::isfinite(s1) || !std::isfinite(s2)) return 0.0;
  int x1, x2;
  std::frexp(s1, &x1);
  std::frexp(s2, &x2);
  double a1 = std::ldexp(e1x, -x1), a2 = std::ldexp(e1y, -x1), a3 = std::ldexp(e1z, -x1);
  double b1 = std::ldexp(e2x, -x2), b2 = std::ldexp(e2y, -x2), b3 = std::ldexp(e2z, -x2);
  double cx = comp_det(a2, b3, a3, b2), cy = comp_det(a3, b1, a1, b3), cz = comp_det(a1, b2, a2, b1);
  double cn = std::hypot(std::hypot(cx, cy), cz);
  if (cn == 0.0) return 0.0;
  int xn;
  double fr = std::frexp(cn, &xn);
  return std::ldexp(0.5 * fr, xn + x1 + x2);
}

// assert_mesh_quality: src/geometry/Mesh.jl:1093-1255.
void mesh_check(const dm_mesh& m, int allow_boundary, int require_closed, double area_tol_rel) {
  DM_REQUIRE(std::isfinite(area_tol_rel) && area_tol_rel >= 0, DM_ERR_ARGUMENT,
             "area_tol_rel must be finite and nonnegative");
  i64 nv = m.nv, nt = m.nt;
  std::vector<char> finite(nv, 1);
  i64 invalid_vertices = 0;
  for (i64 i = 0; i < nv; ++i)
    if (!(std::isfinite(m.xyz[3 * i]) && std::isfinite(m.xyz[3 * i + 1]) && std::isfinite(m.xyz[3 * i + 2]))) {
      finite[i] = 0;
      ++invalid_vertices;
    }
  // mesh scale: max bounding-box extent of valid positive-area faces
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  auto valid_face = [&](i64 t) {
    const i64* tr = &m.tri[3 * t];
    for (int j = 0; j < 3; ++j)
      if (tr[j] < 1 || tr[j] > nv) return false;
    if (tr[0] == tr[1] || tr[1] == tr[2] || tr[2] == tr[0]) return false;
    return finite[tr[0] - 1] && finite[tr[1] - 1] && finite[tr[2] - 1];
  };
  for (i64 t = 0; t < nt; ++t) {
    if (!valid_face(t)) continue;
    double a = triangle_area(m, t);
    if (!(std::isfinite(a) && a > 0)) continue;
    for (int j = 0; j < 3; ++j)
      for (int c = 0; c < 3; ++c) {
        double v = m.xyz[3 * (m.tri[3 * t + j] - 1) + c];
        lo[c] = std::min(lo[c], v);
        hi[c] = std::max(hi[c], v);
      }
  }
  double scale = 0;
  for (int c = 0; c < 3; ++c)
    if (hi[c] >= lo[c]) scale = std::max(scale, hi[c] - lo[c]);
  double tol_abs = 0;
  if (area_tol_rel > 0) {
    double s = std::sqrt(area_tol_rel) * scale;
    tol_abs = s * s;
  }
  i64 invalid_tris = 0, degenerate = 0, duplicate = 0, nonmanifold = 0, conflicts = 0, boundary = 0;
  std::map<std::array<i64, 3>, int> faces;
  std::map<std::pair<i64, i64>, std::vector<int>> edges;
  for (i64 t = 0; t < nt; ++t) {
    if (!valid_face(t)) { ++invalid_tris; continue; }
    const i64* tr = &m.tri[3 * t];
    std::array<i64, 3> key = {tr[0], tr[1], tr[2]};
    std::sort(key.begin(), key.end());
    bool dup = faces.count(key) > 0;
    if (dup) ++duplicate; else faces[key] = 1;
    double a = triangle_area(m, t);
    if (!std::isfinite(a)) { ++invalid_tris; continue; }
    if (a <= tol_abs) ++degenerate;
    if (dup) continue;
    for (int j = 0; j < 3; ++j) {
      i64 a0 = tr[j], b0 = tr[(j + 1) % 3];
      edges[{std::min(a0, b0), std::max(a0, b0)}].push_back(a0 < b0 ? 1 : -1);
    }
  }
  for (auto& kv : edges) {
    size_t d = kv.second.size();
    if (d == 1) ++boundary;
    else if (d == 2) { if (kv.second[0] == kv.second[1]) ++conflicts; }
    else ++nonmanifold;
  }
  std::ostringstream os;
  bool bad = false;
  auto prob = [&](const std::string& s) { os << (bad ? "; " : "") << s; bad = true; };
  if (nv < 3 || nt < 1) prob("mesh needs at least three vertices and one triangle");
  if (invalid_vertices) prob(std::to_string(invalid_vertices) + " non-finite vertices");
  if (invalid_tris) prob(std::to_string(invalid_tris) + " invalid triangles");
  if (degenerate) prob(std::to_string(degenerate) + " degenerate triangles");
  if (duplicate) prob(std::to_string(duplicate) + " duplicate triangles");
  if (nonmanifold) prob(std::to_string(nonmanifold) + " non-manifold edges");
  if (conflicts) prob(std::to_string(conflicts) + " orientation conflicts");
  if ((require_closed || !allow_boundary) && boundary)
    prob(std::to_string(boundary) + " boundary edges");
  if (bad) dm_throw(DM_ERR_ARGUMENT, "Mesh quality check failed: " + os.str());
}

// build_rwg: src/basis/RWG.jl:5-22,55-143.  Interior edges only; T+ = the
// lower-index triangle (the first one to list the edge).  Basis order: sorted
// (v1 < v2) edge keys.
void build_rwg_canonical(const dm_mesh& m, dm_rwg& r) {
  std::vector<std::tuple<i64, i64, i64, int>> rec;  // (va, vb, t, le)
  rec.reserve(3 * m.nt);
  for (i64 t = 0; t < m.nt; ++t)
    for (int le = 0; le < 3; ++le) {
      i64 a = m.tri[3 * t + le], b = m.tri[3 * t + (le + 1) % 3];
      rec.emplace_back(std::min(a, b), std::max(a, b), t + 1, le);
    }
  std::sort(rec.begin(), rec.end());
  r.tplus.clear(); r.tminus.clear(); r.evert.clear(); r.vplus_opp.clear(); r.vminus_opp.clear();
  r.len.clear(); r.area_p.clear(); r.area_m.clear();
  size_t i = 0;
  while (i < rec.size()) {
    size_t j = i + 1;
    while (j < rec.size() && std::get<0>(rec[j]) == std::get<0>(rec[i]) &&
           std::get<1>(rec[j]) == std::get<1>(rec[i]))
      ++j;
    if (j - i == 2) {
      i64 va = std::get<0>(rec[i]), vb = std::get<1>(rec[i]);
      i64 t1 = std::get<2>(rec[i]), t2 = std::get<2>(rec[i + 1]);
      int l1 = std::get<3>(rec[i]), l2 = std::get<3>(rec[i + 1]);
      r.tplus.push_back(t1);
      r.tminus.push_back(t2);
      r.evert.push_back(va);
      r.evert.push_back(vb);
      r.vplus_opp.push_back(m.tri[3 * (t1 - 1) + (l1 + 2) % 3]);
      r.vminus_opp.push_back(m.tri[3 * (t2 - 1) + (l2 + 2) % 3]);
      P3 p1 = vert(m, va), p2 = vert(m, vb);
      double len = std::hypot(std::hypot(p2.x - p1.x, p2.y - p1.y), p2.z - p1.z);
      DM_REQUIRE(len > 0 && std::isfinite(len), DM_ERR_DOMAIN, "RWG edge has zero or unrepresentable length");
      r.len.push_back(len);
      r.area_p.push_back(triangle_area(m, t1 - 1));
      r.area_m.push_back(triangle_area(m, t2 - 1));
    }
    i = j;
  }
  r.n = (i64)r.tplus.size();
  r.coeff_p.assign(r.n, 1.0);
  r.coeff_m.assign(r.n, 1.0);
}

namespace {
// Recursive coordinate bisection of the bases in [b, e): `units` units of
// `unit` bases each (the last one possibly short), cut across the widest
// extent of their centres c so that the lower side holds a whole number of
// units; a unit larger than tb is then cut into blocks of tb the same way.
// Only the last block of the whole order can be short.
void bisect(int* b, int* e, i64 units, i64 unit, i64 tb, const std::array<const double*, 3>& c) {
  if (units <= 1) {
    if (unit > tb) bisect(b, e, (e - b + tb - 1) / tb, tb, tb, c);
    return;
  }
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int* p = b; p != e; ++p)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], c[k][*p]);
      hi[k] = std::max(hi[k], c[k][*p]);
    }
  int d = 0;
  for (int k = 1; k < 3; ++k)
    if (hi[k] - lo[k] > hi[d] - lo[d]) d = k;
  const double* cd = c[d];
  const i64 nl = units / 2, k = std::min<i64>(e - b, nl * unit);
  std::nth_element(b, b + k, e, [cd](int x, int y) { return cd[x] < cd[y] || (cd[x] == cd[y] && x < y); });
  bisect(b, b + k, nl, unit, tb, c);
  bisect(b + k, e, units - nl, unit, tb, c);
}
}  // namespace

void finalize_rwg(dm_rwg& r) {
  const dm_mesh& m = *r.mesh;
  i64 n = r.n, nt = m.nt;
  DM_REQUIRE(n >= 1, DM_ERR_ARGUMENT, "RWG data has no basis functions");
  DM_REQUIRE(n < (1LL << 31) && nt < (1LL << 31), DM_ERR_ARGUMENT, "mesh too large for 32-bit indices");
  r.eplus.resize(n);
  r.eminus.resize(n);
  std::vector<int> hrwg_ext(nt * 3, -1);
  std::vector<double> hsig(nt * 3, 0.0);
  auto locate = [&](i64 t, i64 vopp, const char* side, i64 idx) {
    DM_REQUIRE(t >= 1 && t <= nt, DM_ERR_ARGUMENT,
               std::string("RWG ") + std::to_string(idx + 1) + ": " + side + " triangle out of range");
    const i64* tr = &m.tri[3 * (t - 1)];
    for (int e = 0; e < 3; ++e)
      if (tr[e] == vopp) return e;
    dm_throw(DM_ERR_ARGUMENT, std::string("RWG ") + std::to_string(idx + 1) + ": opposite vertex not in " + side + " triangle");
  };
  for (i64 i = 0; i < n; ++i) {
    DM_REQUIRE(r.tplus[i] != r.tminus[i], DM_ERR_ARGUMENT, "RWG " + std::to_string(i + 1) + ": T+ == T-");
    int ep = locate(r.tplus[i], r.vplus_opp[i], "T+", i);
    int em = locate(r.tminus[i], r.vminus_opp[i], "T-", i);
    r.eplus[i] = ep;
    r.eminus[i] = em;
    DM_REQUIRE(std::isfinite(r.len[i]) && r.len[i] > 0, DM_ERR_DOMAIN, "RWG edge has zero or unrepresentable length");
    DM_REQUIRE(std::isfinite(r.area_p[i]) && r.area_p[i] > 0 && std::isfinite(r.area_m[i]) && r.area_m[i] > 0,
               DM_ERR_ARGUMENT, "RWG triangle areas must be positive and finite");
    DM_REQUIRE(std::isfinite(r.coeff_p[i]) && std::isfinite(r.coeff_m[i]), DM_ERR_ARGUMENT,
               "RWG coefficients must be finite");
    i64 hp = 3 * (r.tplus[i] - 1) + ep, hm = 3 * (r.tminus[i] - 1) + em;
    DM_REQUIRE(hrwg_ext[hp] < 0 && hrwg_ext[hm] < 0, DM_ERR_ARGUMENT,
               "two RWG functions share the same triangle edge (RWG " + std::to_string(i + 1) + ")");
    hrwg_ext[hp] = (int)i;
    hrwg_ext[hm] = (int)i;
    // f_n = sigma * len/(2 A_rwg) * (r - v_opp); quadrature weights use the mesh
    // area, so the half carries sigma * A_mesh / A_rwg (exactly sigma when the
    // caller's areas are the mesh's own).
    hsig[hp] = r.coeff_p[i] * (triangle_area(m, r.tplus[i] - 1) / r.area_p[i]);
    hsig[hm] = -r.coeff_m[i] * (triangle_area(m, r.tminus[i] - 1) / r.area_m[i]);
  }
  // Edge adjacency, src/assembly/EFIE.jl:140-231: triangles sharing a recorded edge.
  std::vector<std::tuple<i64, i64, int>> rec;
  rec.reserve(3 * nt);
  for (i64 t = 0; t < nt; ++t)
    for (int j = 0; j < 3; ++j) {
      i64 a = m.tri[3 * t + j], b = m.tri[3 * t + (j + 1) % 3];
      rec.emplace_back(std::min(a, b), std::max(a, b), (int)t);
    }
  std::sort(rec.begin(), rec.end());
  std::vector<std::pair<int, int>> pairs;
  for (size_t i = 0; i < rec.size();) {
    size_t j = i + 1;
    while (j < rec.size() && std::get<0>(rec[j]) == std::get<0>(rec[i]) && std::get<1>(rec[j]) == std::get<1>(rec[i])) ++j;
    for (size_t a = i; a < j; ++a)
      for (size_t b = a + 1; b < j; ++b) {
        int t1 = std::get<2>(rec[a]), t2 = std::get<2>(rec[b]);
        if (t1 != t2) pairs.emplace_back(std::min(t1, t2), std::max(t1, t2));
      }
    i = j;
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  r.adj_pairs = (i64)pairs.size();
  r.nb.assign(nt * 3, -1);
  std::vector<int> deg(nt, 0);
  for (auto& p : pairs) {
    for (int s = 0; s < 2; ++s) {
      int t = s ? p.second : p.first, o = s ? p.first : p.second;
      DM_REQUIRE(deg[t] < 3, DM_ERR_ARGUMENT,
                 "triangle " + std::to_string(t + 1) + " has more than three edge neighbours (non-manifold mesh)");
      r.nb[3 * t + deg[t]++] = o;
    }
  }
  // Spatial basis order: recursive coordinate bisection of the RWG centres
  // (midpoint of the two triangle centroids) into units of 2 TB bases (one
  // packed storage tile), each cut into its two blocks of TB.  The blocks
  // are compact, nearly convex surface patches: on the 30720-basis sphere a
  // block touches 52.6 triangles on average (a Morton order of the same
  // centres: 57.2), so a tile evaluates 15 % fewer triangle pairs, and the
  // bisection tree keeps nearby blocks at nearby indices.
  std::vector<double> cx(n), cy(n), cz(n);
  for (i64 i = 0; i < n; ++i) {
    double c[3] = {0, 0, 0};
    for (int s = 0; s < 2; ++s) {
      i64 t = (s ? r.tminus[i] : r.tplus[i]) - 1;
      for (int j = 0; j < 3; ++j) {
        P3 v = vert(m, m.tri[3 * t + j]);
        c[0] += v.x; c[1] += v.y; c[2] += v.z;
      }
    }
    cx[i] = c[0] / 6; cy[i] = c[1] / 6; cz[i] = c[2] / 6;
  }
  r.ext_of_int.resize(n);
  std::iota(r.ext_of_int.begin(), r.ext_of_int.end(), 0);
  if (n > 0) {
    const i64 unit = 2 * (i64)r.TB;
    bisect(r.ext_of_int.data(), r.ext_of_int.data() + n, (n + unit - 1) / unit, unit, r.TB,
           {cx.data(), cy.data(), cz.data()});
  }
  r.int_of_ext.resize(n);
  for (i64 i = 0; i < n; ++i) r.int_of_ext[r.ext_of_int[i]] = (int)i;
  // Tiles: blocks of TB consecutive internal bases and the triangles they touch.
  r.nblk = (int)((n + r.TB - 1) / r.TB);
  r.blk_off.assign(r.nblk + 1, 0);
  r.blk_tri.clear();
  r.max_blk_tris = 0;
  std::vector<int> mark(nt, -1);
  for (int b = 0; b < r.nblk; ++b) {
    i64 i0 = (i64)b * r.TB, i1 = std::min<i64>(n, i0 + r.TB);
    for (i64 i = i0; i < i1; ++i) {
      int e = r.ext_of_int[i];
      for (int s = 0; s < 2; ++s) {
        int t = (int)((s ? r.tminus[e] : r.tplus[e]) - 1);
        if (mark[t] != b) { mark[t] = b; r.blk_tri.push_back(t); }
      }
    }
    r.blk_off[b + 1] = (int)r.blk_tri.size();
    r.max_blk_tris = std::max(r.max_blk_tris, r.blk_off[b + 1] - r.blk_off[b]);
  }
  // half -> internal basis index
  std::vector<int> hrwg(nt * 3, -1);
  for (i64 h = 0; h < nt * 3; ++h)
    if (hrwg_ext[h] >= 0) hrwg[h] = r.int_of_ext[hrwg_ext[h]];
  r.cen_i.assign(3 * n, 0.0);
  double lsum = 0;
  for (i64 i = 0; i < n; ++i) {
    int e = r.ext_of_int[i];
    r.cen_i[3 * i] = cx[e];
    r.cen_i[3 * i + 1] = cy[e];
    r.cen_i[3 * i + 2] = cz[e];
    lsum += r.len[e];
  }
  r.mean_len = lsum / n;
  std::vector<double> len_i(n);
  std::vector<int> tp(n), tm(n), ep(n), em(n);
  for (i64 i = 0; i < n; ++i) {
    int e = r.ext_of_int[i];
    len_i[i] = r.len[e];
    tp[i] = (int)(r.tplus[e] - 1);
    tm[i] = (int)(r.tminus[e] - 1);
    ep[i] = r.eplus[e];
    em[i] = r.eminus[e];
  }
  hipStream_t s = r.ctx->stream;
  r.d_ext.upload(r.ext_of_int, s);
  r.d_iofe.upload(r.int_of_ext, s);
  r.d_len.upload(len_i, s);
  r.d_hrwg.upload(hrwg, s);
  r.d_hsig.upload(hsig, s);
  r.d_nb.upload(r.nb, s);
  r.d_blk_off.upload(r.blk_off, s);
  r.d_blk_tri.upload(r.blk_tri, s);
  r.d_tplus_i.upload(tp, s);
  r.d_tminus_i.upload(tm, s);
  r.d_eplus_i.upload(ep, s);
  r.d_eminus_i.upload(em, s);
  HIP_OK(hipStreamSynchronize(s));
}

// Near-field pattern for the sparse approximate inverse: J_j = the (at most
// maxJ nearest) bases with centre distance <= cutJ from basis j, I_j = bases
// within 1.6 x the radius actually covered by J_j (at most 512 nearest).
// Centres are the midpoints of the two triangle centroids
// (rwg_centers, src/solver/NearFieldPreconditioner.jl:247-257).
// The hash grid's cells are sized to the radius the maxJ nearest bases cover
// on a surface of this mesh's mean edge (~2.7 mean edges for 80), not to the
// cutoff, so a cutoff of a wavelength does not make every column scan the
// whole mesh; a column whose J or I radius reaches past the scanned cells is
// scanned again over a wider block of cells.
std::shared_ptr<const SpaiPattern> spai_pattern(dm_rwg& r, double cutJ, int maxJ) {
  std::lock_guard<std::mutex> lk(r.tab_mu);
  if (r.spai && r.spai->cutJ == cutJ) return r.spai;
  auto P = std::make_shared<SpaiPattern>();
  P->cutJ = cutJ;
  i64 n = r.n;
  const double* c = r.cen_i.data();
  double lo[3] = {INFINITY, INFINITY, INFINITY};
  for (i64 i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], c[3 * i + k]);
  const double reach = std::sqrt((double)maxJ / 3.0) * r.mean_len;  // ~ radius of maxJ nearest
  const double cell = 1.6 * std::max(std::min(cutJ, reach), 1e-300);
  std::map<std::array<long, 3>, std::vector<int>> grid;
  auto key = [&](i64 i) {
    return std::array<long, 3>{(long)std::floor((c[3 * i] - lo[0]) / cell), (long)std::floor((c[3 * i + 1] - lo[1]) / cell),
                               (long)std::floor((c[3 * i + 2] - lo[2]) / cell)};
  };
  for (i64 i = 0; i < n; ++i) grid[key(i)].push_back((int)i);
  P->jptr.assign(n + 1, 0);
  P->iptr.assign(n + 1, 0);
  std::vector<std::pair<double, int>> cand;
  double rsum = 0;
  i64 capped = 0;
  for (i64 j = 0; j < n; ++j) {
    auto kj = key(j);
    std::vector<int> J, I;
    double rJ = 0;
    bool capped_col = false;
    for (long w = 1;; ++w) {
      // every centre within w * cell of centre j lies in this block of cells
      cand.clear();
      for (long dx = -w; dx <= w; ++dx)
        for (long dy = -w; dy <= w; ++dy)
          for (long dz = -w; dz <= w; ++dz) {
            auto it = grid.find({kj[0] + dx, kj[1] + dy, kj[2] + dz});
            if (it == grid.end()) continue;
            for (int m : it->second) {
              double ddx = c[3 * m] - c[3 * j], ddy = c[3 * m + 1] - c[3 * j + 1], ddz = c[3 * m + 2] - c[3 * j + 2];
              double d = std::sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
              if (d <= w * cell) cand.emplace_back(d, m);
            }
          }
      std::sort(cand.begin(), cand.end());
      J.clear();
      rJ = 0;
      for (auto& pr : cand) {
        if (pr.first > cutJ || (int)J.size() >= maxJ) break;
        J.push_back(pr.second);
        rJ = pr.first;
      }
      const bool j_done = (int)J.size() >= maxJ || cutJ <= w * cell || (i64)cand.size() == n;
      // the cap cut the cutoff short: a further candidate lies within cutJ
      capped_col = (int)J.size() >= maxJ && J.size() < cand.size() && cand[J.size()].first <= cutJ;
      const double rI = std::max(1.6 * rJ, 1e-300);
      const bool i_done = rI <= w * cell || (i64)cand.size() == n;
      if (j_done && i_done) {
        I.clear();
        for (auto& pr : cand) {
          if (pr.first > rI || (int)I.size() >= 512) break;
          I.push_back(pr.second);
        }
        break;
      }
    }
    rsum += rJ;
    capped += capped_col ? 1 : 0;
    if (J.empty() || J[0] != (int)j) {  // the basis itself is always in its own support
      J.erase(std::remove(J.begin(), J.end(), (int)j), J.end());
      J.insert(J.begin(), (int)j);
      if ((int)J.size() > maxJ) J.resize(maxJ);
    }
    std::sort(J.begin(), J.end());
    std::sort(I.begin(), I.end());
    P->jidx.insert(P->jidx.end(), J.begin(), J.end());
    P->iidx.insert(P->iidx.end(), I.begin(), I.end());
    P->jptr[j + 1] = (int)P->jidx.size();
    P->iptr[j + 1] = (int)P->iidx.size();
  }
  P->mean_radius = rsum / (double)std::max<i64>(n, 1);
  P->capped = capped;
  P->nnz = (i64)P->jidx.size();
  // CSR of M by transposing the CSC support lists
  P->rowptr.assign(n + 1, 0);
  for (int rr : P->jidx) P->rowptr[rr + 1]++;
  for (i64 i = 0; i < n; ++i) P->rowptr[i + 1] += P->rowptr[i];
  P->colidx.resize(P->nnz);
  P->map.resize(P->nnz);
  std::vector<int> fill(P->rowptr.begin(), P->rowptr.end() - 1);
  for (i64 j = 0; j < n; ++j)
    for (int p = P->jptr[j]; p < P->jptr[j + 1]; ++p) {
      int rr = P->jidx[p];
      int q = fill[rr]++;
      P->colidx[q] = (int)j;
      P->map[q] = p;
    }
  hipStream_t s = r.ctx->stream;
  P->d_jptr.upload(P->jptr, s);
  P->d_jidx.upload(P->jidx, s);
  P->d_iptr.upload(P->iptr, s);
  P->d_iidx.upload(P->iidx, s);
  P->d_rowptr.upload(P->rowptr, s);
  P->d_colidx.upload(P->colidx, s);
  P->d_map.upload(P->map, s);
  HIP_OK(hipStreamSynchronize(s));
  r.spai = P;
  return P;
}

// Per-triangle tables for one quadrature order: quad points via the affine map
// of src/basis/Quadrature.jl:76-129 (muladd form), 7-point rule points for the
// singular terms, centroid-relative copies for the tiled far-pair kernels.
const TriTables& tables(dm_rwg& r, int quad_order) {
  std::lock_guard<std::mutex> lk(r.tab_mu);
  DM_REQUIRE(quad_order == 1 || quad_order == 3 || quad_order == 4 || quad_order == 7, DM_ERR_ARGUMENT,
             "Unsupported quadrature order " + std::to_string(quad_order) + ". Use 1, 3, 4, or 7.");
  TriTables& T = r.tabs[quad_order];
  if (T.quad_order == quad_order) return T;
  const dm_mesh& m = *r.mesh;
  i64 nt = m.nt;
  double xi[14], w[7], xi7[14], w7[7];
  int nq = quad_rule(quad_order, xi, w);
  quad_rule(7, xi7, w7);
  std::vector<double> cen(nt * 3), uq(nt * nq * 3), av(nt * 9), rq(nt * nq * 3), rq7(nt * 21), vtx(nt * 9), area(nt);
  auto affine = [](double a, double b, double c, double x1, double x2) {
    double w1 = 1.0 - x1 - x2;
    return std::fma(w1, a, std::fma(x1, b, x2 * c));
  };
  for (i64 t = 0; t < nt; ++t) {
    P3 v[3];
    for (int j = 0; j < 3; ++j) v[j] = vert(m, m.tri[3 * t + j]);
    double c[3] = {(v[0].x + v[1].x + v[2].x) / 3, (v[0].y + v[1].y + v[2].y) / 3, (v[0].z + v[1].z + v[2].z) / 3};
    for (int k = 0; k < 3; ++k) cen[3 * t + k] = c[k];
    for (int j = 0; j < 3; ++j) {
      double vv[3] = {v[j].x, v[j].y, v[j].z};
      for (int k = 0; k < 3; ++k) {
        vtx[9 * t + 3 * j + k] = vv[k];
        av[9 * t + 3 * j + k] = vv[k] - c[k];
      }
    }
    for (int q = 0; q < nq; ++q) {
      double p[3] = {affine(v[0].x, v[1].x, v[2].x, xi[2 * q], xi[2 * q + 1]),
                     affine(v[0].y, v[1].y, v[2].y, xi[2 * q], xi[2 * q + 1]),
                     affine(v[0].z, v[1].z, v[2].z, xi[2 * q], xi[2 * q + 1])};
      for (int k = 0; k < 3; ++k) {
        rq[(t * nq + q) * 3 + k] = p[k];
        uq[(t * nq + q) * 3 + k] = p[k] - c[k];
      }
    }
    for (int q = 0; q < 7; ++q) {
      rq7[(t * 7 + q) * 3 + 0] = affine(v[0].x, v[1].x, v[2].x, xi7[2 * q], xi7[2 * q + 1]);
      rq7[(t * 7 + q) * 3 + 1] = affine(v[0].y, v[1].y, v[2].y, xi7[2 * q], xi7[2 * q + 1]);
      rq7[(t * 7 + q) * 3 + 2] = affine(v[0].z, v[1].z, v[2].z, xi7[2 * q], xi7[2 * q + 1]);
    }
    area[t] = triangle_area(m, t);
  }
  hipStream_t s = r.ctx->stream;
  T.nq = nq;
  T.cen.upload(cen, s);
  T.uq.upload(uq, s);
  T.av.upload(av, s);
  T.rq.upload(rq, s);
  T.rq7.upload(rq7, s);
  T.vtx.upload(vtx, s);
  T.area.upload(area, s);
  T.h_rq = rq;
  T.h_area = area;
  HIP_OK(hipStreamSynchronize(s));
  T.quad_order = quad_order;
  return T;
}

}  // namespace dmh
