// oracle.cpp — CPU FP64 restatement of DiffMoM.jl's surface-EFIE path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests, the smoke entry
// and the CPU-baseline leg of the benchmark load it to compare against the HIP
// library.  Nothing in the product library links or calls it.
//
// Every routine follows a reference function and cites it (path:line relative
// to the DiffMoM.jl source tree).  Numbering conventions follow the reference:
// vertex/triangle indices are 1-based int64, arrays are column-major (3 x Nv,
// 3 x Nt), complex values are interleaved (re, im) doubles.
//
// The reference's BigFloat fallbacks (extreme exponents, cancellation) are not
// restated: they never trigger on the meshes and wavenumbers of the tests and
// benchmark (coordinates ~1e-3..1 m, k ~ 60 rad/m).  The fast-path formulas are
// restated exactly, including evaluation order where it is cheap to keep.
//
// Pinning status (the checks in tests/test_oracle.py; fixtures and the scripts
// that wrote them under tests/golden/):
//   quadrature            literal tables + polynomial exactness (and inexactness
//                         one degree higher)
//   greens / greens_smooth mpmath closed forms at 50 digits (green_mp.json), R=0
//                         special values, exprel branch continuity
//   int_1overR            reference golden values (Test 44, test44.json) +
//                         mpmath polar reduction / 2-D quadrature (int1overR_mp.json)
//   self / adjacent / far independent 30-digit mpmath restatement of the pair
//   pair terms            terms on a bent, jittered 2 x 2-cell plate, real and
//                         lossy k, every adjacent pair in both (outer, inner)
//                         orders (cells_mp.json, make_cells_mp.py; S there by
//                         polar reduction, not the closed form)
//   efie entry            the same fixture's entries Z[m,n] and Z[n,m] (outer =
//                         basis min(m,n)), separated from the swapped rule;
//                         symmetry, matrix-free == dense, quad-order 3->4 < 15 %
//   adjacency             40-pair known answer of the 4x4 plate (Test 45)
//   solves / physics      LU residual, GMRES == direct, sphere-vs-Mie RCS gate
//                         (Test 13)
//   plane-wave RHS        closed form at broadside on a single RWG (numpy)
//   radiation vectors     closed form at broadside on a single RWG + Mie gate

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

typedef std::complex<double> cd;
typedef int64_t i64;

namespace {

struct V3 { double x, y, z; };
inline V3 vsub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 vadd(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 vscale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double vdot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 vcross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline double vnorm(V3 a) { return std::sqrt(vdot(a, a)); }
// R = hypot(hypot(dx, dy), dz): src/basis/Greens.jl:138-143
inline double hyp3(V3 d) { return std::hypot(std::hypot(d.x, d.y), d.z); }

const double INV_FOUR_PI = 1.0 / (4.0 * M_PI);  // src/basis/Greens.jl:11

struct Mesh {
  i64 nv, nt;
  const double* xyz;  // 3 x nv
  const i64* tri;     // 3 x nt, 1-based
  V3 vert(i64 v1based) const {
    const double* p = xyz + 3 * (v1based - 1);
    return {p[0], p[1], p[2]};
  }
  V3 tv(i64 t1based, int j) const { return vert(tri[3 * (t1based - 1) + j]); }
};

struct Rwg {
  i64 n;
  const i64 *tplus, *tminus, *vplus_opp, *vminus_opp;  // 1-based
  const double *len, *area_p, *area_m, *coeff_p, *coeff_m;
};

// ---------------------------------------------------------------------------
// Quadrature: src/basis/Quadrature.jl:17-54 (literal tables kept alike: they
// are the reference's rule and parity needs the same points and weights).
int quad_rule(int order, double* xi /* 2 x nq */, double* w) {
  if (order == 1) {
    xi[0] = 1.0 / 3; xi[1] = 1.0 / 3; w[0] = 0.5;
    return 1;
  } else if (order == 3) {
    const double p[3][2] = {{1.0 / 6, 1.0 / 6}, {2.0 / 3, 1.0 / 6}, {1.0 / 6, 2.0 / 3}};
    for (int q = 0; q < 3; ++q) { xi[2 * q] = p[q][0]; xi[2 * q + 1] = p[q][1]; w[q] = 1.0 / 6; }
    return 3;
  } else if (order == 4) {
    const double p[4][2] = {{1.0 / 3, 1.0 / 3}, {0.6, 0.2}, {0.2, 0.6}, {0.2, 0.2}};
    const double ww[4] = {-27.0 / 96, 25.0 / 96, 25.0 / 96, 25.0 / 96};
    for (int q = 0; q < 4; ++q) { xi[2 * q] = p[q][0]; xi[2 * q + 1] = p[q][1]; w[q] = ww[q]; }
    return 4;
  } else if (order == 7) {
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

// Affine map, src/basis/Quadrature.jl:76-113,121-129:
// muladd(1 - xi1 - xi2, v1, muladd(xi1, v2, xi2 * v3))
inline double affine(double a, double b, double c, double x1, double x2) {
  double w1 = 1.0 - x1 - x2;
  return std::fma(w1, a, std::fma(x1, b, x2 * c));
}
V3 quad_point(const Mesh& m, i64 t, double x1, double x2) {
  V3 a = m.tv(t, 0), b = m.tv(t, 1), c = m.tv(t, 2);
  return {affine(a.x, b.x, c.x, x1, x2), affine(a.y, b.y, c.y, x1, x2),
          affine(a.z, b.z, c.z, x1, x2)};
}

// Compensated, exponent-scaled area: src/geometry/Mesh.jl:2940-2967,2966-3080,3205-3232
inline double comp_det(double a, double b, double c, double d) {
  double p1 = a * b, p2 = c * d;
  double lead = p1 - p2;
  double corr = std::fma(a, b, -p1) - std::fma(c, d, -p2);
  return lead + corr;
}
double triangle_area_v(V3 v1, V3 v2, V3 v3) {
  V3 e1 = vsub(v2, v1), e2 = vsub(v3, v1);
  double s1 = std::max(std::fabs(e1.x), std::max(std::fabs(e1.y), std::fabs(e1.z)));
  double s2 = std::max(std::fabs(e2.x), std::max(std::fabs(e2.y), std::fabs(e2.z)));
  if (!(s1 > 0 && s2 > 0)) return 0.0;
  int x1, x2;
  std::frexp(s1, &x1);
  std::frexp(s2, &x2);
  double a1 = std::ldexp(e1.x, -x1), a2 = std::ldexp(e1.y, -x1), a3 = std::ldexp(e1.z, -x1);
  double b1 = std::ldexp(e2.x, -x2), b2 = std::ldexp(e2.y, -x2), b3 = std::ldexp(e2.z, -x2);
  double cx = comp_det(a2, b3, a3, b2);
  double cy = comp_det(a3, b1, a1, b3);
  double cz = comp_det(a1, b2, a2, b1);
  double cn = std::hypot(std::hypot(cx, cy), cz);
  if (cn == 0.0) return 0.0;
  int xn;
  double fr = std::frexp(cn, &xn);
  return std::ldexp(0.5 * fr, xn + x1 + x2);
}
double triangle_area(const Mesh& m, i64 t) { return triangle_area_v(m.tv(t, 0), m.tv(t, 1), m.tv(t, 2)); }

// ---------------------------------------------------------------------------
// Green's functions: src/basis/Greens.jl:364-379 (G) and 189-201,381-400 (G_s)
cd greens(V3 r, V3 rp, cd k) {
  double R = hyp3(vsub(r, rp));
  if (R == 0.0) return cd(0.0, 0.0);
  cd q(k.imag(), -k.real());
  cd ph = q * R;
  return std::exp(ph) * (INV_FOUR_PI / R);
}

cd green_exprel(cd z) {  // Kahan-compensated series, src/basis/Greens.jl:189-201
  cd term(1.0, 0.0), total(1.0, 0.0), corr(0.0, 0.0);
  for (int order = 1; order <= 18; ++order) {
    term *= z / double(order + 1);
    cd ct = term - corr;
    cd upd = total + ct;
    corr = (upd - total) - ct;
    total = upd;
  }
  return total;
}

// Complex expm1 as Julia Base computes it for finite arguments:
// (expm1(x) - 2 e^x sin^2(y/2)) + i e^x sin(y)
cd cexpm1(cd z) {
  double zr = z.real(), zi = z.imag();
  double erm1 = std::expm1(zr);
  if (zi == 0.0) return cd(erm1, zi);
  double er = erm1 + 1.0;
  double s = std::sin(0.5 * zi);
  double wr = erm1 - 2.0 * er * (s * s);
  return cd(wr, er * std::sin(zi));
}

cd greens_smooth(V3 r, V3 rp, cd k) {
  double R = hyp3(vsub(r, rp));
  cd q(k.imag(), -k.real());
  if (R == 0.0) return q * INV_FOUR_PI;
  cd ph = q * R;
  if (std::abs(ph) <= 0.5) return (q * INV_FOUR_PI) * green_exprel(ph);
  return cexpm1(ph) * (INV_FOUR_PI / R);
}

// ---------------------------------------------------------------------------
// Closed-form static integral S(P) = int_T 1/|P - r'| dS'
// src/assembly/SingularIntegrals.jl:146-203 (unchecked form used by EFIE)
double int_1overR(V3 P, V3 V1, V3 V2, V3 V3_) {
  V3 nT = vcross(vsub(V2, V1), vsub(V3_, V1));
  double nn = vnorm(nT);
  if (nn == 0.0) return 0.0;
  nT = vscale(nT, 1.0 / nn);
  double h = vdot(vsub(P, V1), nT);
  double ah = std::fabs(h);
  V3 xi = vsub(P, vscale(nT, h));
  V3 A_[3] = {V1, V2, V3_}, B_[3] = {V2, V3_, V1};
  double log_sum = 0.0, atan_sum = 0.0;
  for (int e = 0; e < 3; ++e) {
    V3 A = A_[e], B = B_[e];
    V3 ev = vsub(B, A);
    double el = vnorm(ev);
    if (el == 0.0) continue;
    V3 lhat = vscale(ev, 1.0 / el);
    V3 nhat = vcross(lhat, nT);
    double d = vdot(vsub(A, xi), nhat);
    double sm = vdot(vsub(A, xi), lhat);
    double sp = vdot(vsub(B, xi), lhat);
    double Rm = vnorm(vsub(P, A));
    double Rp = vnorm(vsub(P, B));
    double R0sq = d * d + h * h;
    if (d != 0.0) {
      double den = sm + Rm, num = sp + Rp;
      if (den > 0.0 && num > 0.0) log_sum += d * std::log(num / den);
    }
    if (ah != 0.0 && d != 0.0) {
      double ap = std::atan(d * sp / (R0sq + ah * Rp));
      double am = std::atan(d * sm / (R0sq + ah * Rm));
      atan_sum += ap - am;
    }
  }
  return log_sum - ah * atan_sum;
}

// ---------------------------------------------------------------------------
// RWG evaluation: src/basis/RWG.jl:763-783,795-856
inline V3 eval_rwg_side(double coeff, double len, double area, V3 delta) {
  double ratio = len / (2.0 * area);
  return {coeff * (ratio * delta.x), coeff * (ratio * delta.y), coeff * (ratio * delta.z)};
}

// EFIE cache: src/assembly/EFIE.jl:261-279,332-398
struct Cache {
  Mesh mesh;
  Rwg rwg;
  cd k, inv_k2, omega_mu0;
  int nq, nq_hi;
  std::vector<double> wq, wq_hi;
  std::vector<V3> qp, qp_hi;   // nt * nq, nt * nq_hi
  std::vector<double> areas;   // nt
  std::vector<V3> fv, fv_hi;   // rwg values: n * 2 * nq, n * 2 * nq_hi
  std::vector<double> divv;    // n * 2
  std::vector<std::vector<i64>> adj;  // edge-adjacency rows (1-based tri ids)
};

// Compact edge adjacency: src/assembly/EFIE.jl:26-85,140-231 (triangles that
// share a recorded edge; pairs deduplicated).
std::vector<std::vector<i64>> build_adjacency(const Mesh& m) {
  std::vector<std::tuple<i64, i64, i64>> rec;
  rec.reserve(3 * m.nt);
  for (i64 t = 1; t <= m.nt; ++t) {
    i64 v[3] = {m.tri[3 * (t - 1)], m.tri[3 * (t - 1) + 1], m.tri[3 * (t - 1) + 2]};
    for (int j = 0; j < 3; ++j) {
      i64 a = v[j], b = v[(j + 1) % 3];
      rec.emplace_back(std::min(a, b), std::max(a, b), t);
    }
  }
  std::sort(rec.begin(), rec.end());
  std::vector<std::pair<i64, i64>> pairs;
  size_t i = 0;
  while (i < rec.size()) {
    size_t j = i + 1;
    while (j < rec.size() && std::get<0>(rec[j]) == std::get<0>(rec[i]) &&
           std::get<1>(rec[j]) == std::get<1>(rec[i]))
      ++j;
    for (size_t a = i; a < j; ++a)
      for (size_t b = a + 1; b < j; ++b) {
        i64 t1 = std::get<2>(rec[a]), t2 = std::get<2>(rec[b]);
        if (t1 == t2) continue;
        pairs.emplace_back(std::min(t1, t2), std::max(t1, t2));
      }
    i = j;
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  std::vector<std::vector<i64>> adj(m.nt + 1);
  for (auto& p : pairs) {
    adj[p.first].push_back(p.second);
    adj[p.second].push_back(p.first);
  }
  return adj;
}

void build_cache(Cache& c, const Mesh& m, const Rwg& r, cd k, double eta0, int quad_order) {
  c.mesh = m;
  c.rwg = r;
  c.k = k;
  c.inv_k2 = 1.0 / (k * k);
  c.omega_mu0 = k * eta0;
  double xi[14], w[7];
  c.nq = quad_rule(quad_order, xi, w);
  c.wq.assign(w, w + c.nq);
  std::vector<double> xiq(xi, xi + 2 * c.nq);
  c.nq_hi = quad_rule(7, xi, w);
  c.wq_hi.assign(w, w + 7);
  std::vector<double> xih(xi, xi + 14);
  c.qp.resize(m.nt * c.nq);
  c.qp_hi.resize(m.nt * 7);
  c.areas.resize(m.nt);
  for (i64 t = 1; t <= m.nt; ++t) {
    for (int q = 0; q < c.nq; ++q) c.qp[(t - 1) * c.nq + q] = quad_point(m, t, xiq[2 * q], xiq[2 * q + 1]);
    for (int q = 0; q < 7; ++q) c.qp_hi[(t - 1) * 7 + q] = quad_point(m, t, xih[2 * q], xih[2 * q + 1]);
    c.areas[t - 1] = triangle_area(m, t);
  }
  c.fv.resize(r.n * 2 * c.nq);
  c.fv_hi.resize(r.n * 2 * 7);
  c.divv.resize(r.n * 2);
  for (i64 n = 0; n < r.n; ++n) {
    i64 tp = r.tplus[n], tm = r.tminus[n];
    V3 pp = m.vert(r.vplus_opp[n]), pm = m.vert(r.vminus_opp[n]);
    c.divv[2 * n] = r.coeff_p[n] * (r.len[n] / r.area_p[n]);
    c.divv[2 * n + 1] = -r.coeff_m[n] * (r.len[n] / r.area_m[n]);
    for (int q = 0; q < c.nq; ++q) {
      c.fv[(n * 2 + 0) * c.nq + q] = eval_rwg_side(r.coeff_p[n], r.len[n], r.area_p[n], vsub(c.qp[(tp - 1) * c.nq + q], pp));
      c.fv[(n * 2 + 1) * c.nq + q] = eval_rwg_side(r.coeff_m[n], r.len[n], r.area_m[n], vsub(pm, c.qp[(tm - 1) * c.nq + q]));
    }
    for (int q = 0; q < 7; ++q) {
      c.fv_hi[(n * 2 + 0) * 7 + q] = eval_rwg_side(r.coeff_p[n], r.len[n], r.area_p[n], vsub(c.qp_hi[(tp - 1) * 7 + q], pp));
      c.fv_hi[(n * 2 + 1) * 7 + q] = eval_rwg_side(r.coeff_m[n], r.len[n], r.area_m[n], vsub(pm, c.qp_hi[(tm - 1) * 7 + q]));
    }
  }
  c.adj = build_adjacency(m);
}

bool is_adjacent(const Cache& c, i64 t1, i64 t2) {
  for (i64 x : c.adj[t1])
    if (x == t2) return true;
  return false;
}

// Self cell: src/assembly/EFIE.jl:534-606
cd self_cell(const Cache& c, i64 tm, const V3* fm, const V3* fn, double dvm, double dvn, double Am) {
  const Mesh& m = c.mesh;
  V3 V1 = m.tv(tm, 0), V2 = m.tv(tm, 1), V3_ = m.tv(tm, 2);
  const V3* qp = &c.qp_hi[(tm - 1) * 7];
  const double* w = c.wq_hi.data();
  double inv4pi = 1.0 / (4 * M_PI);
  cd dv = dvm * dvn * c.inv_k2;
  cd smooth(0, 0);
  for (int a = 0; a < 7; ++a)
    for (int b = 0; b < 7; ++b) {
      cd Gs = greens_smooth(qp[a], qp[b], c.k);
      cd vec = vdot(fm[a], fn[b]) * Gs;
      cd scl = dv * Gs;
      double wt = w[a] * w[b] * (2 * Am) * (2 * Am);
      smooth += (vec - scl) * wt;
    }
  cd sing(0, 0);
  for (int a = 0; a < 7; ++a) {
    V3 rm = qp[a];
    double S = int_1overR(rm, V1, V2, V3_);
    double inner = inv4pi * S;
    cd scl = dv * inner;
    V3 fn_at = fn[a];
    double lead = vdot(fm[a], fn_at) * inner;
    double rem = 0.0;
    for (int b = 0; b < 7; ++b) {
      double R = hyp3(vsub(rm, qp[b]));
      if (R == 0.0) continue;
      V3 dfn = vsub(fn[b], fn_at);
      rem += vdot(fm[a], dfn) * (inv4pi / R) * w[b] * (2 * Am);
    }
    double ow = w[a] * (2 * Am);
    sing += ((lead + rem) - scl) * ow;
  }
  return smooth + sing;
}

// Adjacent cell: src/assembly/EFIE.jl:611-684
cd adjacent_cell(const Cache& c, i64 tm, i64 tn, const V3* fm, const V3* fn, const V3* fmh,
                 const V3* fnh, double dvm, double dvn, double Am, double An) {
  const Mesh& m = c.mesh;
  V3 V1 = m.tv(tn, 0), V2 = m.tv(tn, 1), V3_ = m.tv(tn, 2);
  const V3* qm = &c.qp[(tm - 1) * c.nq];
  const V3* qn = &c.qp[(tn - 1) * c.nq];
  const V3* qmh = &c.qp_hi[(tm - 1) * 7];
  const V3* qnh = &c.qp_hi[(tn - 1) * 7];
  double inv4pi = 1.0 / (4 * M_PI);
  cd dv = dvm * dvn * c.inv_k2;
  cd smooth(0, 0);
  for (int a = 0; a < c.nq; ++a)
    for (int b = 0; b < c.nq; ++b) {
      cd Gs = greens_smooth(qm[a], qn[b], c.k);
      cd vec = vdot(fm[a], fn[b]) * Gs;
      cd scl = dv * Gs;
      double wt = c.wq[a] * c.wq[b] * (2 * Am) * (2 * An);
      smooth += (vec - scl) * wt;
    }
  cd sing(0, 0);
  for (int a = 0; a < 7; ++a) {
    V3 rm = qmh[a];
    double S = int_1overR(rm, V1, V2, V3_);
    double inner = inv4pi * S;
    cd scl = dv * inner;
    double vs = 0.0;
    for (int b = 0; b < 7; ++b) {
      double R = hyp3(vsub(rm, qnh[b]));
      if (R == 0.0) continue;
      vs += vdot(fmh[a], fnh[b]) * (inv4pi / R) * c.wq_hi[b] * (2 * An);
    }
    double ow = c.wq_hi[a] * (2 * Am);
    sing += (vs - scl) * ow;
  }
  return smooth + sing;
}

// The term of one triangle pair in _efie_entry (src/assembly/EFIE.jl:453-508):
// side itm of basis m (0-based) is the outer triangle, side itn of basis n the
// inner one; self, edge-adjacent or far branch by the triangles' relation.
cd pair_term(const Cache& c, i64 m, int itm, i64 n, int itn) {
  const Rwg& r = c.rwg;
  i64 tm = itm == 0 ? r.tplus[m] : r.tminus[m];
  double Am = c.areas[tm - 1];
  double dvm = c.divv[2 * m + itm];
  const V3* fm = &c.fv[(m * 2 + itm) * c.nq];
  const V3* fmh = &c.fv_hi[(m * 2 + itm) * 7];
  i64 tn = itn == 0 ? r.tplus[n] : r.tminus[n];
  double An = c.areas[tn - 1];
  double dvn = c.divv[2 * n + itn];
  const V3* fn = &c.fv[(n * 2 + itn) * c.nq];
  const V3* fnh = &c.fv_hi[(n * 2 + itn) * 7];
  if (tm == tn) return self_cell(c, tm, fmh, fnh, dvm, dvn, Am);
  if (is_adjacent(c, tm, tn)) return adjacent_cell(c, tm, tn, fm, fn, fmh, fnh, dvm, dvn, Am, An);
  cd val(0, 0);
  cd dv = dvm * dvn * c.inv_k2;
  const V3* qm = &c.qp[(tm - 1) * c.nq];
  const V3* qn = &c.qp[(tn - 1) * c.nq];
  for (int a = 0; a < c.nq; ++a)
    for (int b = 0; b < c.nq; ++b) {
      cd G = greens(qm[a], qn[b], c.k);
      cd vec = vdot(fm[a], fn[b]) * G;
      cd scl = dv * G;
      double wt = c.wq[a] * c.wq[b] * (2 * Am) * (2 * An);
      val += (vec - scl) * wt;
    }
  return val;
}

// One EFIE entry (0-based m, n): src/assembly/EFIE.jl:429-512.  Non-Bloch
// entries are evaluated as (min, max) (:443-445), so basis min(m, n) owns the
// outer triangles.
cd efie_entry(const Cache& c, i64 m, i64 n) {
  if (m > n) std::swap(m, n);
  cd val(0, 0);
  for (int itm = 0; itm < 2; ++itm)
    for (int itn = 0; itn < 2; ++itn) val += pair_term(c, m, itm, n, itn);
  return cd(0, -1) * c.omega_mu0 * val;  // _finalize_efie_entry, src/assembly/EFIE.jl:429-437
}

Mesh mk_mesh(i64 nv, const double* xyz, i64 nt, const i64* tri) { return Mesh{nv, nt, xyz, tri}; }

}  // namespace

extern "C" {

// ------------------------------- mesh generators -----------------------------
// make_rect_plate: src/geometry/Mesh.jl:309-396
void or_make_rect_plate(double Lx, double Ly, i64 Nx, i64 Ny, double* xyz, i64* tri) {
  double hx = Lx / 2, hy = Ly / 2;
  auto coord = [](double half, i64 cells, i64 idx) {
    if (idx == cells) return half;
    if (idx == 0) return -half;
    return half * (2 * (double(idx) / double(cells)) - 1);
  };
  i64 idx = 0;
  for (i64 jy = 0; jy <= Ny; ++jy)
    for (i64 jx = 0; jx <= Nx; ++jx) {
      xyz[3 * idx] = coord(hx, Nx, jx);
      xyz[3 * idx + 1] = coord(hy, Ny, jy);
      xyz[3 * idx + 2] = 0.0;
      ++idx;
    }
  auto vid = [Nx](i64 ix, i64 iy) { return iy * (Nx + 1) + ix + 1; };
  i64 t = 0;
  for (i64 jy = 0; jy < Ny; ++jy)
    for (i64 jx = 0; jx < Nx; ++jx) {
      i64 v1 = vid(jx, jy), v2 = vid(jx + 1, jy), v3 = vid(jx + 1, jy + 1), v4 = vid(jx, jy + 1);
      tri[3 * t] = v1; tri[3 * t + 1] = v2; tri[3 * t + 2] = v3; ++t;
      tri[3 * t] = v1; tri[3 * t + 1] = v3; tri[3 * t + 2] = v4; ++t;
    }
}

// Icosphere: test/runtests.jl:373-427 (vertex order = first-encounter order of
// edge midpoints, so it is deterministic). Counts: Nv = 10*4^s + 2, Nt = 20*4^s.
void or_make_icosphere(double radius, int subdiv, double* xyz, i64* tri) {
  const double phi = (1 + std::sqrt(5.0)) / 2;
  std::vector<V3> v = {{-1, phi, 0}, {1, phi, 0}, {-1, -phi, 0}, {1, -phi, 0},
                       {0, -1, phi}, {0, 1, phi}, {0, -1, -phi}, {0, 1, -phi},
                       {phi, 0, -1}, {phi, 0, 1}, {-phi, 0, -1}, {-phi, 0, 1}};
  for (auto& p : v) p = vscale(p, 1.0 / vnorm(p));
  std::vector<std::array<i64, 3>> f = {
      {1, 12, 6}, {1, 6, 2}, {1, 2, 8}, {1, 8, 11}, {1, 11, 12}, {2, 6, 10}, {6, 12, 5},
      {12, 11, 3}, {11, 8, 7}, {8, 2, 9}, {4, 10, 5}, {4, 5, 3}, {4, 3, 7}, {4, 7, 9},
      {4, 9, 10}, {5, 10, 6}, {3, 5, 12}, {7, 3, 11}, {9, 7, 8}, {10, 9, 2}};
  for (int s = 0; s < subdiv; ++s) {
    std::map<std::pair<i64, i64>, i64> mid;
    std::vector<std::array<i64, 3>> nf;
    auto midpoint = [&](i64 i, i64 j) {
      auto key = std::make_pair(std::min(i, j), std::max(i, j));
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      V3 m = vscale(vadd(v[i - 1], v[j - 1]), 0.5);
      m = vscale(m, 1.0 / vnorm(m));
      v.push_back(m);
      i64 id = (i64)v.size();
      mid[key] = id;
      return id;
    };
    for (auto& t : f) {
      i64 a = midpoint(t[0], t[1]);
      i64 b = midpoint(t[1], t[2]);
      i64 c = midpoint(t[2], t[0]);
      nf.push_back({t[0], a, c});
      nf.push_back({t[1], b, a});
      nf.push_back({t[2], c, b});
      nf.push_back({a, b, c});
    }
    f.swap(nf);
  }
  for (size_t i = 0; i < v.size(); ++i) {
    xyz[3 * i] = radius * v[i].x;
    xyz[3 * i + 1] = radius * v[i].y;
    xyz[3 * i + 2] = radius * v[i].z;
  }
  for (size_t t = 0; t < f.size(); ++t)
    for (int j = 0; j < 3; ++j) tri[3 * t + j] = f[t][j];
}

double or_triangle_area(const double* xyz, const i64* tri, i64 t1) {
  Mesh m{0, 0, xyz, tri};
  return triangle_area(m, t1);
}

// ------------------------------------ RWG ------------------------------------
// build_rwg: src/basis/RWG.jl:5-22,55-143.  T+ = first (lower-index) triangle
// listing the edge.  The reference's basis order is Julia Dict iteration order;
// this restatement emits interior edges sorted by (v1, v2) edge key.
i64 or_build_rwg(const double* xyz, i64 nt, const i64* tri, i64* tplus, i64* tminus, i64* evert,
                 i64* vplus_opp, i64* vminus_opp, double* len, double* area_p, double* area_m) {
  std::map<std::pair<i64, i64>, std::vector<std::pair<i64, int>>> et;
  for (i64 t = 1; t <= nt; ++t)
    for (int le = 0; le < 3; ++le) {
      i64 a = tri[3 * (t - 1) + le], b = tri[3 * (t - 1) + (le + 1) % 3];
      et[{std::min(a, b), std::max(a, b)}].push_back({t, le});
    }
  Mesh m{0, nt, xyz, tri};
  i64 n = 0;
  for (auto& kv : et) {
    if (kv.second.size() != 2) continue;
    if (tplus) {
      i64 t1 = kv.second[0].first, t2 = kv.second[1].first;
      int l1 = kv.second[0].second, l2 = kv.second[1].second;
      tplus[n] = t1;
      tminus[n] = t2;
      evert[2 * n] = kv.first.first;
      evert[2 * n + 1] = kv.first.second;
      vplus_opp[n] = tri[3 * (t1 - 1) + (l1 + 2) % 3];
      vminus_opp[n] = tri[3 * (t2 - 1) + (l2 + 2) % 3];
      V3 r1 = m.vert(kv.first.first), r2 = m.vert(kv.first.second);
      len[n] = hyp3(vsub(r2, r1));
      area_p[n] = triangle_area(m, t1);
      area_m[n] = triangle_area(m, t2);
    }
    ++n;
  }
  return n;
}

i64 or_adjacent_pair_count(i64 nv, const double* xyz, i64 nt, const i64* tri) {
  Mesh m = mk_mesh(nv, xyz, nt, tri);
  auto adj = build_adjacency(m);
  i64 s = 0;
  for (auto& a : adj) s += (i64)a.size();
  return s / 2;
}

// ------------------------------ scalar kernels -------------------------------
int or_quad_rule(int order, double* xi, double* w) { return quad_rule(order, xi, w); }

void or_greens(const double* r, const double* rp, double kr, double ki, double* out) {
  cd g = greens({r[0], r[1], r[2]}, {rp[0], rp[1], rp[2]}, cd(kr, ki));
  out[0] = g.real(); out[1] = g.imag();
}
void or_greens_smooth(const double* r, const double* rp, double kr, double ki, double* out) {
  cd g = greens_smooth({r[0], r[1], r[2]}, {rp[0], rp[1], rp[2]}, cd(kr, ki));
  out[0] = g.real(); out[1] = g.imag();
}
double or_int_1overR(const double* P, const double* A, const double* B, const double* C) {
  return int_1overR({P[0], P[1], P[2]}, {A[0], A[1], A[2]}, {B[0], B[1], B[2]}, {C[0], C[1], C[2]});
}

// ---------------------------------- EFIE -------------------------------------
struct OrCtx { Cache c; };

void* or_efie_create(i64 nv, const double* xyz, i64 nt, const i64* tri, i64 n, const i64* tplus,
                     const i64* tminus, const i64* vplus_opp, const i64* vminus_opp, const double* len,
                     const double* area_p, const double* area_m, const double* coeff_p,
                     const double* coeff_m, double kr, double ki, double eta0, int quad_order) {
  OrCtx* o = new OrCtx();
  Mesh m = mk_mesh(nv, xyz, nt, tri);
  Rwg r{n, tplus, tminus, vplus_opp, vminus_opp, len, area_p, area_m, coeff_p, coeff_m};
  build_cache(o->c, m, r, cd(kr, ki), eta0, quad_order);
  return o;
}
void or_efie_free(void* h) { delete (OrCtx*)h; }

// Entries Z[m_i, n_i] (0-based indices) into out (interleaved complex).
void or_efie_entries(void* h, const i64* ms, const i64* ns, i64 count, double* out) {
  const Cache& c = ((OrCtx*)h)->c;
#pragma omp parallel for schedule(dynamic, 64)
  for (i64 i = 0; i < count; ++i) {
    cd z = efie_entry(c, ms[i], ns[i]);
    out[2 * i] = z.real();
    out[2 * i + 1] = z.imag();
  }
}

// Triangle-pair terms (before -i k eta0): outer = side sm[i] of basis m[i],
// inner = side sn[i] of basis n[i] (0-based bases, side 0 = T+, 1 = T-).
void or_pair_terms(void* h, const i64* ms, const int* sm, const i64* ns, const int* sn, i64 count,
                   double* out) {
  const Cache& c = ((OrCtx*)h)->c;
  for (i64 i = 0; i < count; ++i) {
    cd z = pair_term(c, ms[i], sm[i], ns[i], sn[i]);
    out[2 * i] = z.real();
    out[2 * i + 1] = z.imag();
  }
}

// Dense Z, column-major N x N, rows [row0, row1) only (symmetric fill inside the
// block): src/assembly/EFIE.jl:699-753 with Threads.@threads -> OpenMP.
void or_assemble_z_rows(void* h, i64 row0, i64 row1, double* Z, i64 ldz) {
  const Cache& c = ((OrCtx*)h)->c;
  i64 N = c.rwg.n;
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 m = row0; m < row1; ++m)
    for (i64 n = m; n < N; ++n) {
      cd z = efie_entry(c, m, n);
      Z[2 * (m + n * ldz)] = z.real();
      Z[2 * (m + n * ldz) + 1] = z.imag();
      if (n < row1 && n != m) {
        Z[2 * (n + m * ldz)] = z.real();
        Z[2 * (n + m * ldz) + 1] = z.imag();
      }
    }
}

// CPU-baseline sample: entries (m, n >= m) for rows [row0, row1), reduced to a
// checksum (no storage).  Returns the number of entries evaluated.
i64 or_assemble_rows_sum(void* h, i64 row0, i64 row1, double* checksum) {
  const Cache& c = ((OrCtx*)h)->c;
  i64 N = c.rwg.n;
  double s = 0;
  i64 cnt = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : s, cnt)
  for (i64 m = row0; m < row1; ++m)
    for (i64 n = m; n < N; ++n) {
      cd z = efie_entry(c, m, n);
      s += z.real() + z.imag();
      ++cnt;
    }
  *checksum = s;
  return cnt;
}

// Matrix-free product y = A x recomputing every entry: src/assembly/EFIE.jl:867-909
// (the exact-reduction BigFloat fallback is not restated).
void or_mf_apply(void* h, const double* x, double* y) {
  const Cache& c = ((OrCtx*)h)->c;
  i64 N = c.rwg.n;
  const cd* xc = (const cd*)x;
#pragma omp parallel for schedule(dynamic, 4)
  for (i64 m = 0; m < N; ++m) {
    cd acc(0, 0);
    for (i64 n = 0; n < N; ++n) acc += efie_entry(c, m, n) * xc[n];
    y[2 * m] = acc.real();
    y[2 * m + 1] = acc.imag();
  }
}

// ------------------------------- excitation ----------------------------------
// assemble_plane_wave: src/assembly/Excitation.jl:2829-2896 with the field of
// src/assembly/Excitation.jl:1517-1547 (E = pol*E0*exp(-i k.r)).
void or_plane_wave(i64 nv, const double* xyz, i64 nt, const i64* tri, i64 n, const i64* tplus,
                   const i64* tminus, const i64* vplus_opp, const i64* vminus_opp, const double* len,
                   const double* area_p, const double* area_m, const double* coeff_p,
                   const double* coeff_m, const double* kvec, double E0, const double* pol,
                   int quad_order, double* v) {
  Mesh m = mk_mesh(nv, xyz, nt, tri);
  double xi[14], w[7];
  int nq = quad_rule(quad_order, xi, w);
  for (i64 e = 0; e < n; ++e) {
    cd acc(0, 0);
    for (int side = 0; side < 2; ++side) {
      i64 t = side == 0 ? tplus[e] : tminus[e];
      double A = triangle_area(m, t);
      V3 p = m.vert(side == 0 ? vplus_opp[e] : vminus_opp[e]);
      for (int q = 0; q < nq; ++q) {
        V3 rq = quad_point(m, t, xi[2 * q], xi[2 * q + 1]);
        V3 f = side == 0 ? eval_rwg_side(coeff_p[e], len[e], area_p[e], vsub(rq, p))
                         : eval_rwg_side(coeff_m[e], len[e], area_m[e], vsub(p, rq));
        double arg = 1.0 * (kvec[0] * rq.x + kvec[1] * rq.y + kvec[2] * rq.z);
        cd phase = std::exp(cd(0, -1) * arg);
        cd E[3] = {pol[0] * E0 * phase, pol[1] * E0 * phase, pol[2] * E0 * phase};
        cd inner = f.x * E[0] + f.y * E[1] + f.z * E[2];
        cd ap = A * inner;
        acc += -(2 * w[q]) * ap;
      }
    }
    v[2 * e] = acc.real();
    v[2 * e + 1] = acc.imag();
  }
}

// ------------------------------- far field -----------------------------------
// radiation_vectors: src/postprocessing/FarField.jl:425-574 (fast path);
// output G is (3*NΩ) x N column-major complex.
void or_radiation_vectors(i64 nv, const double* xyz, i64 nt, const i64* tri, i64 n,
                          const i64* tplus, const i64* tminus, const i64* vplus_opp,
                          const i64* vminus_opp, const double* len, const double* area_p,
                          const double* area_m, const double* coeff_p, const double* coeff_m,
                          const double* rhat, i64 nomega, double k, double eta0, int quad_order,
                          double* G) {
  Mesh m = mk_mesh(nv, xyz, nt, tri);
  double xi[14], w[7];
  int nq = quad_rule(quad_order, xi, w);
  cd pref = cd(0, 1) * k * eta0 / (4 * M_PI);
  i64 rows = 3 * nomega;
#pragma omp parallel for schedule(dynamic, 16)
  for (i64 e = 0; e < n; ++e) {
    for (i64 qd = 0; qd < nomega; ++qd) {
      V3 rh{rhat[3 * qd], rhat[3 * qd + 1], rhat[3 * qd + 2]};
      double rn2 = vdot(rh, rh);
      cd tot[3] = {0, 0, 0};
      for (int side = 0; side < 2; ++side) {
        i64 t = side == 0 ? tplus[e] : tminus[e];
        double A = triangle_area(m, t);
        V3 p = m.vert(side == 0 ? vplus_opp[e] : vminus_opp[e]);
        for (int q = 0; q < nq; ++q) {
          V3 rp = quad_point(m, t, xi[2 * q], xi[2 * q + 1]);
          V3 f = side == 0 ? eval_rwg_side(coeff_p[e], len[e], area_p[e], vsub(rp, p))
                           : eval_rwg_side(coeff_m[e], len[e], area_m[e], vsub(p, rp));
          double wt = w[q] * 2 * A;
          cd phase = std::exp(cd(0, 1) * k * vdot(rh, rp));
          cd c[3] = {f.x * (wt * phase), f.y * (wt * phase), f.z * (wt * phase)};
          // rh x (rh x c) / |rh|^2
          cd a1 = rh.y * c[2] - rh.z * c[1], a2 = rh.z * c[0] - rh.x * c[2], a3 = rh.x * c[1] - rh.y * c[0];
          cd b1 = rh.y * a3 - rh.z * a2, b2 = rh.z * a1 - rh.x * a3, b3 = rh.x * a2 - rh.y * a1;
          tot[0] += pref * (b1 / rn2);
          tot[1] += pref * (b2 / rn2);
          tot[2] += pref * (b3 / rn2);
        }
      }
      for (int j = 0; j < 3; ++j) {
        G[2 * (3 * qd + j + e * rows)] = tot[j].real();
        G[2 * (3 * qd + j + e * rows) + 1] = tot[j].imag();
      }
    }
  }
}

}  // extern "C"
