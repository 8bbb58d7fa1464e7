// k_assembly.hip — EFIE impedance-matrix assembly kernels for gfx950.
//
// Formulation.  For a basis m with support triangles T_m^+ / T_m^- write each
// half as f = sigma * len/(2A) * (r - v_e), v_e the vertex opposite the edge
// (src/basis/RWG.jl:819-856).  For a triangle pair (t, t') and opposite-vertex
// indices (e, e') let
//   J[e][e'] = sum_{q,q'} w_q w_q' G(r_q, r'_q') [ (r_q - v_e).(r'_q' - v'_e') - 4/k^2 ]
// Then Z_mn = -i k eta0 * len_m len_n * sum over the four support pairs of
// sigma_m sigma_n J[e_m][e_n]  — exactly _efie_entry of
// src/assembly/EFIE.jl:439-512 with the area factors cancelled.  One J block
// serves the 3 x 3 basis pairs of a triangle pair, so a triangle pair's nine
// Green evaluations are shared by nine matrix entries instead of being redone
// for each (the reference evaluates 36 Green values per entry).
//
// Self (t == t') and edge-adjacent pairs use the reference's singularity
// extraction (src/assembly/EFIE.jl:534-684): they are precomputed in FP64 into
// per-triangle tables by k_near_self / k_near_adj and looked up by the tile
// kernel.  The adjacent-cell integral is not symmetric in (outer, inner); the
// reference evaluates entry (min(m,n), max(m,n)) with the lower-index basis's
// triangle as the outer one, and the tile kernel follows the caller's basis
// numbering to pick the same table.
//
// Tiling.  Bases are in a spatial order (recursive coordinate bisection of
// their centres into compact patches); block b = TB consecutive
// bases, touching the triangle list blk_tri[blk_off[b]..blk_off[b+1]).  One
// workgroup owns tile (bi <= bj): it evaluates every (row triangle, column
// triangle) pair of the two lists, accumulates sigma*sigma*J into an LDS tile
// with ds_add atomics, and writes the tile and its mirror (Z = Z^T) to HBM.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <type_traits>

#include "dm_common.h"

namespace {

constexpr double INV4PI = 0.07957747154594767;  // 1/(4 pi)

struct c2 {
  double re, im;
};
__device__ inline c2 cmul(c2 a, c2 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline c2 cadd(c2 a, c2 b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline c2 csub(c2 a, c2 b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline c2 cscale(c2 a, double s) { return {a.re * s, a.im * s}; }

// ------------------------------------------------------------------ FP64 near
// int_T 1/|P - r'| dS' (Wilton/Graglia closed form), unchecked form of
// src/assembly/SingularIntegrals.jl:146-203.
__device__ double int_1overR(const double* P, const double* A0, const double* B0, const double* C0) {
  double e1[3] = {B0[0] - A0[0], B0[1] - A0[1], B0[2] - A0[2]};
  double e2[3] = {C0[0] - A0[0], C0[1] - A0[1], C0[2] - A0[2]};
  double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
  double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  if (nn == 0.0) return 0.0;
  n[0] /= nn; n[1] /= nn; n[2] /= nn;
  double h = (P[0] - A0[0]) * n[0] + (P[1] - A0[1]) * n[1] + (P[2] - A0[2]) * n[2];
  double ah = fabs(h);
  double xi[3] = {P[0] - h * n[0], P[1] - h * n[1], P[2] - h * n[2]};
  const double* As[3] = {A0, B0, C0};
  const double* Bs[3] = {B0, C0, A0};
  double log_sum = 0.0, atan_sum = 0.0;
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const double* A = As[e];
    const double* B = Bs[e];
    double ev[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
    double el = sqrt(ev[0] * ev[0] + ev[1] * ev[1] + ev[2] * ev[2]);
    if (el == 0.0) continue;
    double l[3] = {ev[0] / el, ev[1] / el, ev[2] / el};
    double nh[3] = {l[1] * n[2] - l[2] * n[1], l[2] * n[0] - l[0] * n[2], l[0] * n[1] - l[1] * n[0]};
    double ax[3] = {A[0] - xi[0], A[1] - xi[1], A[2] - xi[2]};
    double bx[3] = {B[0] - xi[0], B[1] - xi[1], B[2] - xi[2]};
    double d = ax[0] * nh[0] + ax[1] * nh[1] + ax[2] * nh[2];
    double sm = ax[0] * l[0] + ax[1] * l[1] + ax[2] * l[2];
    double sp = bx[0] * l[0] + bx[1] * l[1] + bx[2] * l[2];
    double pa[3] = {P[0] - A[0], P[1] - A[1], P[2] - A[2]};
    double pb[3] = {P[0] - B[0], P[1] - B[1], P[2] - B[2]};
    double Rm = sqrt(pa[0] * pa[0] + pa[1] * pa[1] + pa[2] * pa[2]);
    double Rp = sqrt(pb[0] * pb[0] + pb[1] * pb[1] + pb[2] * pb[2]);
    double R0sq = d * d + h * h;
    if (d != 0.0) {
      double den = sm + Rm, num = sp + Rp;
      if (den > 0.0 && num > 0.0) log_sum += d * log(num / den);
    }
    if (ah != 0.0 && d != 0.0) {
      double ap = atan(d * sp / (R0sq + ah * Rp));
      double am = atan(d * sm / (R0sq + ah * Rm));
      atan_sum += ap - am;
    }
  }
  return log_sum - ah * atan_sum;
}

__device__ inline double dist3(const double* a, const double* b) {
  double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return hypot(hypot(dx, dy), dz);
}

// G_smooth = (exp(-ikR) - 1)/(4 pi R): src/basis/Greens.jl:189-201,381-400.
// q = (Im k, -Re k); R = 0 -> q/(4 pi); |qR| <= 0.5 -> (q/4pi) * exprel(qR) by the
// Kahan-compensated 18-term series; otherwise expm1(qR)/(4 pi R).
__device__ c2 greens_smooth(double R, c2 q) {
  if (R == 0.0) return cscale(q, INV4PI);
  c2 z = cscale(q, R);
  double az = hypot(z.re, z.im);
  if (az <= 0.5) {
    c2 term = {1.0, 0.0}, total = {1.0, 0.0}, corr = {0.0, 0.0};
    for (int order = 1; order <= 18; ++order) {
      term = cscale(cmul(term, z), 1.0 / double(order + 1));
      c2 ct = csub(term, corr);
      c2 upd = cadd(total, ct);
      corr = csub(csub(upd, total), ct);
      total = upd;
    }
    return cmul(cscale(q, INV4PI), total);
  }
  // complex expm1 for finite z: (expm1(x) - 2 e^x sin^2(y/2)) + i e^x sin(y)
  double erm1 = expm1(z.re);
  c2 em1;
  if (z.im == 0.0) {
    em1 = {erm1, 0.0};
  } else {
    double er = erm1 + 1.0;
    double s = sin(0.5 * z.im);
    em1 = {erm1 - 2.0 * er * (s * s), er * sin(z.im)};
  }
  return cscale(em1, INV4PI / R);
}

struct NearArgs {
  i64 nt;
  int nq;
  const double* rq;   // nt*nq*3
  const double* rq7;  // nt*21
  const double* vtx;  // nt*9
  const double* area; // nt
  const int* nb;      // nt*3
  double w[7], w7[7];
  c2 q;               // (Im k, -Re k)
  c2 four_ik2;        // 4 / k^2
};

// Self cell (src/assembly/EFIE.jl:534-606), all nine (e, e') pairs of one
// triangle, normalised by (2A)^2:
//   J = sum_ab (f_e(a).f_e'(b) - 4/k^2) Gs_ab w_a w_b
//     + sum_a w_a [ (f_e(a).f_e'(a) - 4/k^2) S_a/(4 pi 2A) + sum_b f_e(a).(r_b - r_a) w_b/(4 pi R_ab) ]
// with f_e(r) = r - v_e.
__device__ void near_self_one(const NearArgs& a, i64 t, double* jself) {
  const double* P = a.rq7 + t * 21;
  const double* V = a.vtx + t * 9;
  double A = a.area[t];
  c2 J[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) J[i] = {0.0, 0.0};
  // smooth part, 7 x 7 with G_smooth
  for (int ia = 0; ia < 7; ++ia) {
    const double* ra = P + 3 * ia;
    double fa[3][3];
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
      for (int c = 0; c < 3; ++c) fa[e][c] = ra[c] - V[3 * e + c];
    for (int ib = 0; ib < 7; ++ib) {
      const double* rb = P + 3 * ib;
      c2 Gs = greens_smooth(dist3(ra, rb), a.q);
      double wt = a.w7[ia] * a.w7[ib];
#pragma unroll
      for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          double dd = fa[e][0] * (rb[0] - V[3 * f]) + fa[e][1] * (rb[1] - V[3 * f + 1]) + fa[e][2] * (rb[2] - V[3 * f + 2]);
          c2 coef = {dd - a.four_ik2.re, -a.four_ik2.im};
          J[3 * e + f] = cadd(J[3 * e + f], cscale(cmul(coef, Gs), wt));
        }
    }
  }
  // singular part: analytic inner integral + bounded remainder
  for (int ia = 0; ia < 7; ++ia) {
    const double* ra = P + 3 * ia;
    double S = int_1overR(ra, V, V + 3, V + 6);
    double sc = a.w7[ia] * S * INV4PI / (2.0 * A);
    double rem[3] = {0, 0, 0};
    for (int ib = 0; ib < 7; ++ib) {
      const double* rb = P + 3 * ib;
      double R = dist3(ra, rb);
      if (R == 0.0) continue;
      double k = a.w7[ia] * a.w7[ib] * INV4PI / R;
#pragma unroll
      for (int e = 0; e < 3; ++e)
        rem[e] += ((ra[0] - V[3 * e]) * (rb[0] - ra[0]) + (ra[1] - V[3 * e + 1]) * (rb[1] - ra[1]) +
                   (ra[2] - V[3 * e + 2]) * (rb[2] - ra[2])) * k;
    }
#pragma unroll
    for (int e = 0; e < 3; ++e)
#pragma unroll
      for (int f = 0; f < 3; ++f) {
        double dd = (ra[0] - V[3 * e]) * (ra[0] - V[3 * f]) + (ra[1] - V[3 * e + 1]) * (ra[1] - V[3 * f + 1]) +
                    (ra[2] - V[3 * e + 2]) * (ra[2] - V[3 * f + 2]);
        c2 coef = {dd - a.four_ik2.re, -a.four_ik2.im};
        J[3 * e + f] = cadd(J[3 * e + f], cscale(coef, sc));
        J[3 * e + f].re += rem[e];
      }
  }
  c2* out = reinterpret_cast<c2*>(jself) + t * 9;
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i] = J[i];
}

// Adjacent cell (src/assembly/EFIE.jl:611-684), outer triangle t, inner
// triangle u = nb[t][j], normalised by (2A_t)(2A_u):
//   J = sum_ab^{nq} (f_e(a).f'_e'(b) - 4/k^2) Gs_ab w_a w_b
//     + sum_a^7 w_a [ sum_b^7 f_e(a).f'_e'(b) w_b/(4 pi R_ab) - 4/k^2 S_u(a)/(4 pi 2A_u) ]
__device__ void near_adj_one(const NearArgs& a, i64 idx, double* jadj) {
  i64 t = idx / 3;
  int u = a.nb[idx];
  c2* out = reinterpret_cast<c2*>(jadj) + idx * 9;
  if (u < 0) {
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = {0.0, 0.0};
    return;
  }
  const double* Vt = a.vtx + t * 9;
  const double* Vu = a.vtx + (i64)u * 9;
  double Au = a.area[u];
  c2 J[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) J[i] = {0.0, 0.0};
  const int nq = a.nq;
  for (int ia = 0; ia < nq; ++ia) {
    const double* ra = a.rq + (t * nq + ia) * 3;
    for (int ib = 0; ib < nq; ++ib) {
      const double* rb = a.rq + ((i64)u * nq + ib) * 3;
      c2 Gs = greens_smooth(dist3(ra, rb), a.q);
      double wt = a.w[ia] * a.w[ib];
#pragma unroll
      for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          double dd = (ra[0] - Vt[3 * e]) * (rb[0] - Vu[3 * f]) + (ra[1] - Vt[3 * e + 1]) * (rb[1] - Vu[3 * f + 1]) +
                      (ra[2] - Vt[3 * e + 2]) * (rb[2] - Vu[3 * f + 2]);
          c2 coef = {dd - a.four_ik2.re, -a.four_ik2.im};
          J[3 * e + f] = cadd(J[3 * e + f], cscale(cmul(coef, Gs), wt));
        }
    }
  }
  for (int ia = 0; ia < 7; ++ia) {
    const double* ra = a.rq7 + t * 21 + 3 * ia;
    double S = int_1overR(ra, Vu, Vu + 3, Vu + 6);
    double sc = a.w7[ia] * S * INV4PI / (2.0 * Au);
    double vs[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) vs[i] = 0.0;
    for (int ib = 0; ib < 7; ++ib) {
      const double* rb = a.rq7 + (i64)u * 21 + 3 * ib;
      double R = dist3(ra, rb);
      if (R == 0.0) continue;
      double k = a.w7[ia] * a.w7[ib] * INV4PI / R;
#pragma unroll
      for (int e = 0; e < 3; ++e)
#pragma unroll
        for (int f = 0; f < 3; ++f)
          vs[3 * e + f] += ((ra[0] - Vt[3 * e]) * (rb[0] - Vu[3 * f]) + (ra[1] - Vt[3 * e + 1]) * (rb[1] - Vu[3 * f + 1]) +
                            (ra[2] - Vt[3 * e + 2]) * (rb[2] - Vu[3 * f + 2])) * k;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) {
      J[i].re += vs[i] - sc * a.four_ik2.re;
      J[i].im -= sc * a.four_ik2.im;
    }
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) out[i] = J[i];
}

__global__ __launch_bounds__(64) void k_near_self(NearArgs a, double* jself) {
  const i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (t < a.nt) near_self_one(a, t, jself);
}
__global__ __launch_bounds__(64) void k_near_adj(NearArgs a, double* jadj) {
  const i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx < a.nt * 3) near_adj_one(a, idx, jadj);
}

// Many meshes in one launch: item g of the concatenated lists (per mesh m,
// items per triangle x its triangles starting at toff[m]) belongs to the mesh
// found by binary search; each mesh's tables start at per x 9 complex x toff[m].
__device__ inline int mesh_of(const i64* __restrict__ toff, int nmesh, i64 g) {
  int lo = 0, hi = nmesh - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (toff[mid] <= g)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}
__global__ __launch_bounds__(64) void k_near_self_multi(const NearArgs* __restrict__ args,
                                                        const i64* __restrict__ toff, int nmesh, double* jself) {
  const i64 g = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (g >= toff[nmesh]) return;
  const int m = mesh_of(toff, nmesh, g);
  near_self_one(args[m], g - toff[m], jself + 18 * toff[m]);
}
__global__ __launch_bounds__(64) void k_near_adj_multi(const NearArgs* __restrict__ args,
                                                       const i64* __restrict__ toff, int nmesh, double* jadj) {
  const i64 g = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (g >= 3 * toff[nmesh]) return;
  const int m = mesh_of(toff, nmesh, g / 3);
  near_adj_one(args[m], g - 3 * toff[m], jadj + 54 * toff[m]);
}

// ------------------------------------------------------------------ far tiles
template <typename T>
struct TriLds;

// LDS image of one tile triangle (structure of arrays, `cap` triangles).
template <typename T, int NQ>
struct TileSide {
  double* c;    // cap*3 centroid (FP64: the centroid difference is formed in FP64)
  T* u;         // cap*NQ*3 quad points - centroid
  T* av;        // cap*9 vertices - centroid
  double* sig;  // cap*3 half sign (0 when the half's basis is outside this block)
  int* slot;    // cap*3 local basis index in this block or -1
  int* gid;     // cap
  int* nb;      // cap*3 edge neighbours
};

// Rounds p up to 16 bytes as an offset from the 16-byte aligned LDS base:
// pointer arithmetic keeps the LDS address space visible to the compiler
// (an integer round trip would turn every access through it into a flat one).
__device__ inline char* lds_align16(char* base, char* p) {
  return base + (((p - base) + 15) & ~(ptrdiff_t)15);
}

template <typename T, int NQ>
__device__ inline char* carve_side(char* base, char* p, int cap, TileSide<T, NQ>& s) {
  s.c = reinterpret_cast<double*>(p); p += sizeof(double) * cap * 3;
  s.sig = reinterpret_cast<double*>(p); p += sizeof(double) * cap * 3;
  s.u = reinterpret_cast<T*>(p); p += sizeof(T) * cap * NQ * 3;
  s.av = reinterpret_cast<T*>(p); p += sizeof(T) * cap * 9;
  p = lds_align16(base, p);
  s.slot = reinterpret_cast<int*>(p); p += sizeof(int) * cap * 3;
  s.gid = reinterpret_cast<int*>(p); p += sizeof(int) * cap;
  s.nb = reinterpret_cast<int*>(p); p += sizeof(int) * cap * 3;
  return lds_align16(base, p);
}

template <typename T, int NQ>
__host__ __device__ inline size_t side_bytes(int cap) {
  size_t b = sizeof(double) * cap * 6 + sizeof(T) * cap * (NQ * 3 + 9);
  b = (b + 15) & ~size_t(15);
  b += sizeof(int) * cap * 7;
  b = (b + 15) & ~size_t(15);
  return b;
}

template <typename T, int NQ>
__device__ inline void load_side(const AsmArgs& a, int blk, int nt_side, TileSide<T, NQ>& s, int tid, int nthr) {
  const int* tl = a.blk_tri + a.blk_off[blk];
  const int lo = blk * a.TB;
  for (int i = tid; i < nt_side; i += nthr) {
    int g = tl[i];
    s.gid[i] = g;
#pragma unroll
    for (int c = 0; c < 3; ++c) s.c[3 * i + c] = a.cen[3 * (i64)g + c];
#pragma unroll
    for (int k = 0; k < NQ * 3; ++k) s.u[i * NQ * 3 + k] = (T)a.uq[(i64)g * NQ * 3 + k];
#pragma unroll
    for (int k = 0; k < 9; ++k) s.av[9 * i + k] = (T)a.av[9 * (i64)g + k];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      int r = a.hrwg[3 * (i64)g + e];
      int loc = (r >= lo && r < lo + a.TB) ? r - lo : -1;
      s.slot[3 * i + e] = loc;
      s.sig[3 * i + e] = loc >= 0 ? a.hsig[3 * (i64)g + e] : 0.0;
      s.nb[3 * i + e] = a.nb[3 * (i64)g + e];
    }
  }
}

// ------------------------------------------------------------ fast kernels
// sin/cos of a phase kR (|kR| << 2^20) in FP64: Cody-Waite reduction by pi/2
// with a two-term split applied through FMA, then Taylor polynomials on
// |r| <= pi/4 (sin to r^15, cos to r^16: truncation < 5e-17).  The split is
// pi/2 rounded to double plus the double nearest the remainder, both computed
// from the exact value of pi.
__device__ inline void sincos_phase(double x, double* s, double* c) {
  const double PIO2_HI = 1.5707963267948966, PIO2_LO = 6.123233995736766e-17, TWO_OVER_PI = 0.6366197723675814;
  double n = rint(x * TWO_OVER_PI);
  double r = fma(-n, PIO2_HI, x);
  r = fma(-n, PIO2_LO, r);
  double z = r * r;
  double ps = fma(z, -7.647163731819816e-13, 1.6059043836821613e-10);
  ps = fma(z, ps, -2.505210838544172e-08);
  ps = fma(z, ps, 2.7557319223985893e-06);
  ps = fma(z, ps, -0.0001984126984126984);
  ps = fma(z, ps, 0.008333333333333333);
  ps = fma(z, ps, -0.16666666666666666);
  double sr = fma(r * z, ps, r);
  double pc = fma(z, 4.779477332387385e-14, -1.1470745597729725e-11);
  pc = fma(z, pc, 2.08767569878681e-09);
  pc = fma(z, pc, -2.755731922398589e-07);
  pc = fma(z, pc, 2.48015873015873e-05);
  pc = fma(z, pc, -0.001388888888888889);
  pc = fma(z, pc, 0.041666666666666664);
  pc = fma(z, pc, -0.5);
  double cr = fma(z, pc, 1.0);
  int q = (int)n & 3;
  double ss = (q & 1) ? cr : sr;
  double cc = (q & 1) ? sr : cr;
  *s = (q & 2) ? -ss : ss;
  *c = ((q + 1) & 2) ? -cc : cc;
}

// 1/sqrt in FP64: the hardware estimate (relative error ~2^-23) refined by
// one third-order step, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2 (error
// O(e^3), below FP64 rounding).
__device__ inline double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// sin/cos of a phase x >= 0 through a table of the 256 angles 2 pi j / 256
// (tab[j] = (cos, sin), built in LDS with sincos_phase): x = 2 pi n / 256 + r
// with |r| <= pi / 256 (two-term reduction through FMA), sin r and cos r by
// Taylor polynomials (sin to r^5, cos to r^6: truncation < 1e-17), then the
// angle-addition formulas.  About half the instructions of sincos_phase.
constexpr int kPhaseTable = 256;
__device__ inline void sincos_table(double x, const double2* __restrict__ tab, double* s, double* c) {
  const double STEP_HI = 0.02454369260617026, STEP_LO = 9.567553118338697e-19;  // pi / 128 = hi + lo
  const double INV_STEP = 40.74366543152521;                                      // 128 / pi
  const double n = rint(x * INV_STEP);
  double r = fma(-n, STEP_HI, x);
  r = fma(-n, STEP_LO, r);
  const double z = r * r;
  const double sr = fma(r * z, fma(z, 0.008333333333333333, -0.16666666666666666), r);
  const double cr = fma(z, fma(z, fma(z, -0.001388888888888889, 0.041666666666666664), -0.5), 1.0);
  const double2 t = tab[(int)n & (kPhaseTable - 1)];
  *s = fma(t.y, cr, t.x * sr);
  *c = fma(t.x, cr, -(t.y * sr));
}

// Far-pair J block in compute type T (moment form in centroid-relative
// coordinates, so FP32 keeps its precision for nearby pairs).
// LOSSY: complex wavenumber (the exp(-Im(k) R) attenuation is applied); the
// choice is made once per tile so the real-k loop carries no test for it.
template <typename T, int NQ, bool LOSSY>
__device__ inline void far_block(const TileSide<T, NQ>& R, int i, const TileSide<T, NQ>& C, int j,
                                 const T* ww, const double2* __restrict__ ptab, T kr, T ki, T fik_re, T fik_im,
                                 T Jr[9], T Ji[9]) {
  const T kr_turns = kr * (T)0.15915494309189535;  // k / (2 pi)
  T dc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) dc[c] = (T)(R.c[3 * i + c] - C.c[3 * j + c]);
  T uc[NQ][3];
#pragma unroll
  for (int q = 0; q < NQ; ++q)
#pragma unroll
    for (int c = 0; c < 3; ++c) uc[q][c] = C.u[(j * NQ + q) * 3 + c];
  // Moments of G over the quadrature pairs: M0 = sum G, Mu = sum G u,
  // Mv = sum G v, Muv = sum G u.v.  For rules of up to 4 points Mu is formed
  // from the row sums of G (one per outer point) and Mv from its column sums
  // (one per inner point), so each pair adds only to the row, column and u.v
  // sums; the 7-point rule accumulates every moment per pair (its column sums
  // would not fit the register budget).
  constexpr bool kSums = NQ <= 4;
  T m0r = 0, m0i = 0, muvr = 0, muvi = 0;
  T mur[3] = {0, 0, 0}, mui[3] = {0, 0, 0}, mvr[3] = {0, 0, 0}, mvi[3] = {0, 0, 0};
  T cpr[kSums ? NQ : 1], cpi[kSums ? NQ : 1];
#pragma unroll
  for (int p = 0; p < (kSums ? NQ : 1); ++p) cpr[p] = cpi[p] = 0;
  // FP64: the outer quadrature loop stays rolled (fewer live registers, so
  // more waves per SIMD); FP32 unrolls it
  constexpr int UQ = sizeof(T) == 8 ? 1 : NQ;
#pragma unroll UQ
  for (int q = 0; q < NQ; ++q) {
    T urq[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) urq[c] = R.u[(i * NQ + q) * 3 + c];
    T sr = 0, si = 0, wr = 0, wi = 0;
#pragma unroll
    for (int p = 0; p < NQ; ++p) {
      T dx = dc[0] + (urq[0] - uc[p][0]);
      T dy = dc[1] + (urq[1] - uc[p][1]);
      T dz = dc[2] + (urq[2] - uc[p][2]);
      T R2 = dx * dx + dy * dy + dz * dz;
      T sn, cs, Rd, g;
      if constexpr (sizeof(T) == 8) {
        T iR = rsqrt_nr(R2);
        Rd = R2 * iR;
        g = ww[q * NQ + p] * iR;
        sincos_table(kr * Rd, ptab, &sn, &cs);
      } else {
        // hardware rsq / sin / cos; the phase goes in revolutions (turns)
        T iR = __builtin_amdgcn_rsqf(R2);
        Rd = R2 * iR;
        g = ww[q * NQ + p] * iR;
        T turns = __builtin_amdgcn_fractf(kr_turns * Rd);
        sn = __builtin_amdgcn_sinf(turns);
        cs = __builtin_amdgcn_cosf(turns);
      }
      if constexpr (LOSSY) g *= exp(ki * Rd);
      T gr = g * cs, gi = -g * sn;
      T uv = urq[0] * uc[p][0] + urq[1] * uc[p][1] + urq[2] * uc[p][2];
      wr += gr * uv; wi += gi * uv;
      if constexpr (kSums) {
        sr += gr; si += gi;
        cpr[p] += gr; cpi[p] += gi;
      } else {
        m0r += gr; m0i += gi;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          mur[c] += gr * urq[c]; mui[c] += gi * urq[c];
          mvr[c] += gr * uc[p][c]; mvi[c] += gi * uc[p][c];
        }
      }
    }
    muvr += wr; muvi += wi;
    if constexpr (kSums) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        mur[c] += sr * urq[c]; mui[c] += si * urq[c];
      }
    }
  }
  if constexpr (kSums) {
#pragma unroll
    for (int p = 0; p < NQ; ++p) {
      m0r += cpr[p]; m0i += cpi[p];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        mvr[c] += cpr[p] * uc[p][c]; mvi[c] += cpi[p] * uc[p][c];
      }
    }
  }
  // (u - a).(u' - a') = u.u' - a.u' - u.a' + a.a'
  T bvr[3], bvi[3];  // a'_f . Mu
#pragma unroll
  for (int f = 0; f < 3; ++f) {
    const T* af = C.av + 9 * j + 3 * f;
    bvr[f] = af[0] * mur[0] + af[1] * mur[1] + af[2] * mur[2];
    bvi[f] = af[0] * mui[0] + af[1] * mui[1] + af[2] * mui[2];
  }
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const T* ae = R.av + 9 * i + 3 * e;
    T avr = ae[0] * mvr[0] + ae[1] * mvr[1] + ae[2] * mvr[2];
    T avi = ae[0] * mvi[0] + ae[1] * mvi[1] + ae[2] * mvi[2];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const T* af = C.av + 9 * j + 3 * f;
      T aa = ae[0] * af[0] + ae[1] * af[1] + ae[2] * af[2];
      T cr = aa - fik_re, ci = -fik_im;
      Jr[3 * e + f] = muvr - avr - bvr[f] + (cr * m0r - ci * m0i);
      Ji[3 * e + f] = muvi - avi - bvi[f] + (cr * m0i + ci * m0r);
    }
  }
}

__device__ inline void store_c(double* o, double vr, double vi) {
  *reinterpret_cast<double2*>(o) = make_double2(vr, vi);
}
__device__ inline void store_c(float* o, double vr, double vi) {
  *reinterpret_cast<float2*>(o) = make_float2((float)vr, (float)vi);
}

__device__ inline void tile_coords(int nblk, int& bi, int& bj, bool& valid) {
  // XCD-aware order: workgroup w runs on XCD w % 8; give each XCD a contiguous
  // range of tiles so tiles sharing a column block share that XCD's L2.
  i64 T = (i64)nblk * (nblk + 1) / 2;
  i64 per = (T + 7) / 8;
  i64 w = blockIdx.x;
  i64 l = (w % 8) * per + w / 8;
  valid = l < T;
  if (!valid) return;
  // l = bj (bj + 1) / 2 + bi, 0 <= bi <= bj
  i64 b = (i64)((sqrt(8.0 * (double)l + 1.0) - 1.0) * 0.5);
  while (b * (b + 1) / 2 > l) --b;
  while ((b + 1) * (b + 2) / 2 <= l) ++b;
  bj = (int)b;
  bi = (int)(l - b * (b + 1) / 2);
}

// MODE 0: write Z (AccT = storage scalar), MODE 1: matrix-free y += Z x.
// One tile (bi <= bj) of one mesh: the whole workgroup's work.
template <typename T, typename AccT, int NQ, int NTHR, int MODE>
__device__ inline void tile_body(const AsmArgs& a, int cap, AccT* __restrict__ zout, const double* __restrict__ x,
                                 double* __restrict__ y, int bi, int bj) {
  extern __shared__ __align__(16) char smem[];
  const int TB = a.TB;
  const int ldt = TB + 1;
  AccT* zt = reinterpret_cast<AccT*>(smem);
  char* p = smem + sizeof(AccT) * 2 * TB * ldt;
  p = lds_align16(smem, p);
  TileSide<T, NQ> R, C;
  p = carve_side<T, NQ>(smem, p, cap, R);
  p = carve_side<T, NQ>(smem, p, cap, C);
  int* extr = reinterpret_cast<int*>(p);
  int* extc = extr + TB;
  T* ww = reinterpret_cast<T*>(extc + TB + 2);
  // FP64 phase table (cos, sin of 2 pi j / 256), used by the FP64 far pairs
  double2* ptab = reinterpret_cast<double2*>(lds_align16(smem, reinterpret_cast<char*>(ww + 49)));
  const int tid = threadIdx.x;
  const int nI = a.blk_off[bi + 1] - a.blk_off[bi];
  const int nJ = a.blk_off[bj + 1] - a.blk_off[bj];
  for (int k = tid; k < 2 * TB * ldt; k += NTHR) zt[k] = (AccT)0;
  {
    // both sides' images are contiguous in LDS (R then C): 16-byte copies
    const i64 stride = sizeof(T) == 8 ? a.side_stride64 : a.side_stride32;
    const char* img = sizeof(T) == 8 ? a.side_img64 : a.side_img32;
    const int4* sr = reinterpret_cast<const int4*>(img + (i64)bi * stride);
    const int4* sc = reinterpret_cast<const int4*>(img + (i64)bj * stride);
    int4* dr = reinterpret_cast<int4*>(R.c);
    int4* dc = reinterpret_cast<int4*>(reinterpret_cast<char*>(R.c) + stride);
    const int nv = (int)(stride / 16);
    for (int k = tid; k < nv; k += NTHR) {
      dr[k] = sr[k];
      dc[k] = sc[k];
    }
  }
  for (int k = tid; k < TB; k += NTHR) {
    i64 gi = (i64)bi * TB + k, gj = (i64)bj * TB + k;
    extr[k] = gi < a.n ? a.ext[gi] : 0x7fffffff;
    extc[k] = gj < a.n ? a.ext[gj] : 0x7fffffff;
  }
  if (tid < NQ * NQ) ww[tid] = (T)(a.w[tid / NQ] * a.w[tid % NQ] * INV4PI);
  if constexpr (sizeof(T) == 8) {
    for (int k = tid; k < kPhaseTable; k += NTHR) {
      double sn, cs;
      // k pi / 128 with pi / 128 split in two parts (hi + lo)
      sincos_phase(fma((double)k, 0.02454369260617026, (double)k * 9.567553118338697e-19), &sn, &cs);
      ptab[k] = make_double2(cs, sn);
    }
  }
  __syncthreads();

  const T kr = (T)a.k_re, ki = (T)a.k_im;
  // 4/k^2
  double kk_re = a.k_re * a.k_re - a.k_im * a.k_im, kk_im = 2.0 * a.k_re * a.k_im;
  double kd = kk_re * kk_re + kk_im * kk_im;
  const double f4r = 4.0 * kk_re / kd, f4i = -4.0 * kk_im / kd;
  const c2* jself = reinterpret_cast<const c2*>(a.jself);
  const c2* jadj = reinterpret_cast<const c2*>(a.jadj);

  const int npairs = nI * nJ;
  auto pair_loop = [&](auto lossy) {
    for (int pidx = tid; pidx < npairs; pidx += NTHR) {
      const int i = pidx / nJ;
      const int j = pidx - i * nJ;
      const int gi = R.gid[i], gj = C.gid[j];
      int jn = -1;
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (R.nb[3 * i + k] == gj) jn = k;
      if (gi == gj || jn >= 0) {
        // self / edge-adjacent: FP64 table lookup with the caller-order rule
        int jn2 = 0;
#pragma unroll
        for (int k = 0; k < 3; ++k)
          if (C.nb[3 * j + k] == gi) jn2 = k;
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          int rs = R.slot[3 * i + e];
          if (rs < 0) continue;
#pragma unroll
          for (int f = 0; f < 3; ++f) {
            int cs = C.slot[3 * j + f];
            if (cs < 0) continue;
            bool fwd = extr[rs] <= extc[cs];
            c2 v;
            if (gi == gj)
              v = fwd ? jself[(i64)gi * 9 + 3 * e + f] : jself[(i64)gi * 9 + 3 * f + e];
            else
              v = fwd ? jadj[((i64)gi * 3 + jn) * 9 + 3 * e + f] : jadj[((i64)gj * 3 + jn2) * 9 + 3 * f + e];
            double s = R.sig[3 * i + e] * C.sig[3 * j + f];
            atomicAdd(zt + rs * ldt + cs, (AccT)(v.re * s));
            atomicAdd(zt + TB * ldt + rs * ldt + cs, (AccT)(v.im * s));
          }
        }
        continue;
      }
      T Jr[9], Ji[9];
      far_block<T, NQ, decltype(lossy)::value>(R, i, C, j, ww, ptab, kr, ki, (T)f4r, (T)f4i, Jr, Ji);
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        int rs = R.slot[3 * i + e];
        if (rs < 0) continue;
        T sr = (T)R.sig[3 * i + e];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
          int cs = C.slot[3 * j + f];
          if (cs < 0) continue;
          T s = sr * (T)C.sig[3 * j + f];
          atomicAdd(zt + rs * ldt + cs, (AccT)(Jr[3 * e + f] * s));
          atomicAdd(zt + TB * ldt + rs * ldt + cs, (AccT)(Ji[3 * e + f] * s));
        }
      }
    }
  };
  if (a.k_im != 0.0)
    pair_loop(std::true_type{});
  else
    pair_loop(std::false_type{});
  __syncthreads();

  // Z_mn = -i k eta0 len_m len_n * tile ; P = -i k eta0 = (k_im eta0, -k_re eta0)
  const double Pr = a.k_im * a.eta0, Pi = -a.k_re * a.eta0;
  const i64 r0 = (i64)bi * TB, c0 = (i64)bj * TB;
  float* const dual = sizeof(AccT) == 8 ? a.z32_dual : nullptr;
  const bool diag = bi == bj;
  if constexpr (MODE == 0) {
    int bad = 0;
    for (int k = tid; k < TB * TB; k += NTHR) {
      int r = k % TB, c = k / TB;
      i64 gr = r0 + r, gc = c0 + c;
      if (gr >= a.n || gc >= a.n) continue;
      int rr = r, cc = c;
      if (diag && r > c) { rr = c; cc = r; }
      double tr = (double)zt[rr * ldt + cc], ti = (double)zt[TB * ldt + rr * ldt + cc];
      double l2 = a.len[gr] * a.len[gc];
      double vr = (Pr * tr - Pi * ti) * l2, vi = (Pr * ti + Pi * tr) * l2;
      bad |= !isfinite(vr) || !isfinite(vi);
      const i64 at = gr + gc * a.ld;
      store_c(zout + 2 * at, vr, vi);
      if (dual) store_c(dual + 2 * at, vr, vi);
    }
    if (!diag) {
      for (int k = tid; k < TB * TB; k += NTHR) {
        int c = k % TB, r = k / TB;
        i64 gr = r0 + r, gc = c0 + c;
        if (gr >= a.n || gc >= a.n) continue;
        double tr = (double)zt[r * ldt + c], ti = (double)zt[TB * ldt + r * ldt + c];
        double l2 = a.len[gr] * a.len[gc];
        double vr = (Pr * tr - Pi * ti) * l2, vi = (Pr * ti + Pi * tr) * l2;
        const i64 at = gc + gr * a.ld;
        store_c(zout + 2 * at, vr, vi);
        if (dual) store_c(dual + 2 * at, vr, vi);
      }
    }
    if (bad) atomicOr(a.flag, 1);
  } else if constexpr (MODE == 2) {
    // packed symmetric storage: assembly tile (bi, bj) is the (bi & 1, bj & 1)
    // quarter of storage tile (bi / 2, bj / 2); a storage tile on the diagonal
    // also receives the mirror of its off-diagonal quarter
    int bad = 0;
    const int BI = bi >> 1, BJ = bj >> 1;
    const int ro = (bi & 1) * TB, co = (bj & 1) * TB;
    const i64 tbase = ((i64)BJ * (BJ + 1) / 2 + BI) * kPT * kPT;
    const bool mirror = BI == BJ && !diag;
    for (int k = tid; k < TB * TB; k += NTHR) {
      int r = k % TB, c = k / TB;
      i64 gr = r0 + r, gc = c0 + c;
      if (gr >= a.n || gc >= a.n) continue;
      int rr = r, cc = c;
      if (diag && r > c) { rr = c; cc = r; }
      double tr = (double)zt[rr * ldt + cc], ti = (double)zt[TB * ldt + rr * ldt + cc];
      double l2 = a.len[gr] * a.len[gc];
      double vr = (Pr * tr - Pi * ti) * l2, vi = (Pr * ti + Pi * tr) * l2;
      bad |= !isfinite(vr) || !isfinite(vi);
      const i64 at = tbase + (i64)(co + c) * kPT + ro + r;
      store_c(zout + 2 * at, vr, vi);
      if (dual) store_c(dual + 2 * at, vr, vi);
      if (mirror) {
        const i64 am = tbase + (i64)(ro + r) * kPT + co + c;
        store_c(zout + 2 * am, vr, vi);
        if (dual) store_c(dual + 2 * am, vr, vi);
      }
    }
    if (bad) atomicOr(a.flag, 1);
  } else {
    // y[rows] += Z_tile x[cols]; y[cols] += Z_tile^T x[rows] (off-diagonal tiles)
    for (int k = tid; k < 2 * TB; k += NTHR) {
      bool rows = k < TB;
      int r = rows ? k : k - TB;
      if (!rows && diag) continue;
      i64 go = rows ? r0 + r : c0 + r;
      if (go >= a.n) continue;
      double accr = 0, acci = 0;
      for (int c = 0; c < TB; ++c) {
        i64 gi2 = rows ? c0 + c : r0 + c;
        if (gi2 >= a.n) break;
        int rr = rows ? r : c, cc = rows ? c : r;
        if (diag && rr > cc) { int t2 = rr; rr = cc; cc = t2; }
        double tr = (double)zt[rr * ldt + cc], ti = (double)zt[TB * ldt + rr * ldt + cc];
        double l2 = a.len[go] * a.len[gi2];
        double vr = (Pr * tr - Pi * ti) * l2, vi = (Pr * ti + Pi * tr) * l2;
        double xr = x[2 * gi2], xi = x[2 * gi2 + 1];
        accr += vr * xr - vi * xi;
        acci += vr * xi + vi * xr;
      }
      atomicAdd(y + 2 * go, accr);
      atomicAdd(y + 2 * go + 1, acci);
    }
  }
}

template <typename T, typename AccT, int NQ, int NTHR, int MODE>
__global__ __launch_bounds__(NTHR) void k_tile(AsmArgs a, int cap, AccT* __restrict__ zout,
                                                const double* __restrict__ x, double* __restrict__ y) {
  int bi, bj;
  bool valid;
  tile_coords(a.nblk, bi, bj, valid);
  if (!valid) return;
  tile_body<T, AccT, NQ, NTHR, MODE>(a, cap, zout, x, y, bi, bj);
}

// Persistent assembly of several meshes in one launch: every workgroup takes
// the next tile from a global counter over the concatenated tile lists
// (meshes in the caller's order, largest first for balance) until all
// cum[nmesh] tiles are done.
template <typename T, typename AccT, int NQ, int NTHR, int MODE>
__global__ __launch_bounds__(NTHR) void k_tile_multi(const AsmArgs* __restrict__ args, AccT* const* __restrict__ zouts,
                                                      const i64* __restrict__ cum, int nmesh, int cap,
                                                      int* __restrict__ counter) {
  __shared__ int item;
  const i64 total = cum[nmesh];
  for (;;) {
    if (threadIdx.x == 0) item = atomicAdd(counter, 1);
    __syncthreads();
    const i64 w = item;
    __syncthreads();
    if (w >= total) break;
    int m = 0;
    while (m + 1 < nmesh && cum[m + 1] <= w) ++m;
    const i64 l = w - cum[m];
    i64 b = (i64)((sqrt(8.0 * (double)l + 1.0) - 1.0) * 0.5);
    while (b * (b + 1) / 2 > l) --b;
    while ((b + 1) * (b + 2) / 2 <= l) ++b;
    const AsmArgs& a = args[m];  // read in place: a private copy would live in scratch
    tile_body<T, AccT, NQ, NTHR, MODE>(a, cap, zouts[m], nullptr, nullptr, (int)(l - b * (b + 1) / 2), (int)b);
    __syncthreads();
  }
}

__global__ void k_fill_int(int* p, int v) { *p = v; }

// One block's tile-side image in global memory, laid out exactly as the
// tile kernel's LDS (carve_side), so tiles copy it with 16-byte loads.
template <typename T, int NQ>
__global__ void k_build_sides(AsmArgs a, int cap, char* img, i64 stride) {
  const int b = blockIdx.x;
  TileSide<T, NQ> S;
  char* base = img + (i64)b * stride;  // stride: a multiple of 16 bytes
  carve_side<T, NQ>(base, base, cap, S);
  load_side<T, NQ>(a, b, a.blk_off[b + 1] - a.blk_off[b], S, threadIdx.x, blockDim.x);
}

template <typename T, typename AccT>
size_t tile_smem(int nq, int TB, int cap) {
  size_t b = sizeof(AccT) * 2 * TB * (TB + 1);
  b = (b + 15) & ~size_t(15);
  size_t sb = 0;
  switch (nq) {
    case 1: sb = side_bytes<T, 1>(cap); break;
    case 3: sb = side_bytes<T, 3>(cap); break;
    case 4: sb = side_bytes<T, 4>(cap); break;
    default: sb = side_bytes<T, 7>(cap); break;
  }
  b += 2 * sb;
  b += sizeof(int) * (2 * TB + 2) + sizeof(T) * 49;
  if (sizeof(T) == 8) b = ((b + 15) & ~size_t(15)) + sizeof(double) * 2 * 256;  // phase table
  return b;
}

template <typename T, typename AccT, int MODE, int NTHR>
void launch_tile_n(const AsmArgs& a, int cap, AccT* z, const double* x, double* y, hipStream_t s) {
  size_t smem = tile_smem<T, AccT>(a.nq, a.TB, cap);
  DM_REQUIRE(smem <= 160 * 1024, DM_ERR_ARGUMENT, "tile working set exceeds LDS");
  i64 T_ = (i64)a.nblk * (a.nblk + 1) / 2;
  i64 grid = ((T_ + 7) / 8) * 8;
  DM_REQUIRE(grid < (1LL << 31), DM_ERR_ARGUMENT, "tile grid too large");
  switch (a.nq) {
#define DM_CASE(NQ)                                                                              \
  case NQ:                                                                                       \
    HIP_OK(hipFuncSetAttribute((const void*)k_tile<T, AccT, NQ, NTHR, MODE>,                     \
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));          \
    k_tile<T, AccT, NQ, NTHR, MODE><<<dim3((unsigned)grid), dim3(NTHR), smem, s>>>(a, cap, z, x, y); \
    break;
    DM_CASE(1)
    DM_CASE(3)
    DM_CASE(4)
    DM_CASE(7)
#undef DM_CASE
    default:
      dm_throw(DM_ERR_ARGUMENT, "unsupported quadrature size");
  }
  HIP_OK(hipGetLastError());
}

// FP64: 1024 threads (16 waves, <= 128 VGPRs each) fill the CU that the ~98 KB
// of LDS per tile limits to one workgroup; 512 is the tuner's other candidate.
// FP32: 512.
template <typename T, typename AccT, int MODE>
void launch_tile(const AsmArgs& a, int cap, AccT* z, const double* x, double* y, hipStream_t s, int nthr = 0) {
  if (sizeof(T) == 8 && nthr != 512) {
    DM_REQUIRE(nthr == 0 || nthr == 1024, DM_ERR_ARGUMENT, "FP64 tile threads must be 512 or 1024");
    launch_tile_n<T, AccT, MODE, 1024>(a, cap, z, x, y, s);
  } else {
    DM_REQUIRE(nthr == 0 || nthr == 512, DM_ERR_ARGUMENT, "FP32 tile threads must be 512");
    launch_tile_n<T, AccT, MODE, 512>(a, cap, z, x, y, s);
  }
}

template <typename T, typename AccT, int MODE>
void launch_tile_multi(const std::vector<AsmArgs>& args, const std::vector<AccT*>& zouts, int cap, hipStream_t s) {
  constexpr int NTHR = 512;
  const int nmesh = (int)args.size();
  DM_REQUIRE(nmesh >= 1, DM_ERR_ARGUMENT, "no meshes");
  const int nq = args[0].nq;
  for (const auto& a : args) DM_REQUIRE(a.nq == nq && a.TB == args[0].TB, DM_ERR_ARGUMENT, "meshes differ in quadrature");
  size_t smem = tile_smem<T, AccT>(nq, args[0].TB, cap);
  DM_REQUIRE(smem <= 160 * 1024, DM_ERR_ARGUMENT, "tile working set exceeds LDS");
  std::vector<i64> cum(nmesh + 1, 0);
  for (int m = 0; m < nmesh; ++m) cum[m + 1] = cum[m] + (i64)args[m].nblk * (args[m].nblk + 1) / 2;
  DM_REQUIRE(cum[nmesh] < (1LL << 31), DM_ERR_ARGUMENT, "too many tiles for one launch");
  DBuf<AsmArgs> dargs;
  DBuf<AccT*> dz;
  DBuf<i64> dcum;
  DBuf<int> counter;
  dargs.upload(args, s);
  dz.upload(zouts, s);
  dcum.upload(cum, s);
  counter.alloc(1);
  k_fill_int<<<1, 1, 0, s>>>(counter.p, 0);
  int dev = 0, ncu = 0;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  switch (nq) {
#define DM_CASE(NQ)                                                                                        \
  case NQ: {                                                                                               \
    auto kfn = k_tile_multi<T, AccT, NQ, NTHR, MODE>;                                                      \
    HIP_OK(hipFuncSetAttribute((const void*)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem)); \
    int per_cu = 0;                                                                                        \
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, NTHR, smem));                        \
    i64 grid = std::min<i64>(cum[nmesh], (i64)std::max(per_cu, 1) * ncu);                                  \
    kfn<<<dim3((unsigned)grid), dim3(NTHR), smem, s>>>(dargs.p, dz.p, dcum.p, nmesh, cap, counter.p);      \
    break;                                                                                                 \
  }
    DM_CASE(1)
    DM_CASE(3)
    DM_CASE(4)
    DM_CASE(7)
#undef DM_CASE
    default:
      dm_throw(DM_ERR_ARGUMENT, "unsupported quadrature size");
  }
  HIP_OK(hipGetLastError());
  HIP_OK(hipStreamSynchronize(s));  // the argument tables are freed on return
}

}  // namespace

static NearArgs near_args(const AsmArgs& a, const TriTables& t) {
  NearArgs n;
  n.nt = a.nt;
  n.nq = t.nq;
  n.rq = t.rq.p;
  n.rq7 = t.rq7.p;
  n.vtx = t.vtx.p;
  n.area = t.area.p;
  n.nb = a.nb;
  for (int i = 0; i < 7; ++i) n.w[i] = a.w[i];
  double xi[14], w7[7];
  dmh::quad_rule(7, xi, w7);
  for (int i = 0; i < 7; ++i) n.w7[i] = w7[i];
  n.q = {a.k_im, -a.k_re};
  double kk_re = a.k_re * a.k_re - a.k_im * a.k_im, kk_im = 2.0 * a.k_re * a.k_im;
  double kd = kk_re * kk_re + kk_im * kk_im;
  n.four_ik2 = {4.0 * kk_re / kd, -4.0 * kk_im / kd};
  return n;
}

void launch_near_tables(const AsmArgs& a, const TriTables& t, double* jself, double* jadj, hipStream_t s) {
  NearArgs n = near_args(a, t);
  unsigned g1 = (unsigned)((a.nt + 63) / 64), g2 = (unsigned)((3 * a.nt + 63) / 64);
  k_near_self<<<g1, 64, 0, s>>>(n, jself);
  k_near_adj<<<g2, 64, 0, s>>>(n, jadj);
  HIP_OK(hipGetLastError());
}

void launch_near_tables_multi(const std::vector<AsmArgs>& a, const std::vector<const TriTables*>& t, double* jself,
                              double* jadj, DBuf<char>& staging, hipStream_t s) {
  const int nmesh = (int)a.size();
  DM_REQUIRE(nmesh >= 1 && t.size() == a.size(), DM_ERR_ARGUMENT, "one table set per mesh");
  std::vector<NearArgs> h(nmesh);
  std::vector<i64> toff(nmesh + 1, 0);
  for (int m = 0; m < nmesh; ++m) {
    h[m] = near_args(a[m], *t[m]);
    toff[m + 1] = toff[m] + a[m].nt;
  }
  const size_t ab = sizeof(NearArgs) * nmesh, ob = sizeof(i64) * (nmesh + 1);
  const size_t obase = (ab + 15) & ~size_t(15);
  staging.alloc(obase + ob);
  HIP_OK(hipMemcpyAsync(staging.p, h.data(), ab, hipMemcpyHostToDevice, s));
  HIP_OK(hipMemcpyAsync(staging.p + obase, toff.data(), ob, hipMemcpyHostToDevice, s));
  const NearArgs* dargs = reinterpret_cast<const NearArgs*>(staging.p);
  const i64* doff = reinterpret_cast<const i64*>(staging.p + obase);
  const i64 total = toff[nmesh];
  k_near_self_multi<<<(unsigned)((total + 63) / 64), 64, 0, s>>>(dargs, doff, nmesh, jself);
  k_near_adj_multi<<<(unsigned)((3 * total + 63) / 64), 64, 0, s>>>(dargs, doff, nmesh, jadj);
  HIP_OK(hipGetLastError());
}

void zero_packed_padding(i64 n, void* z, size_t scalar_bytes, hipStream_t s) {
  const i64 nb = (n + kPT - 1) / kPT;
  if (n % kPT == 0 || nb == 0) return;
  const size_t tile_bytes = (size_t)kPT * kPT * 2 * scalar_bytes;
  const size_t first = (size_t)((nb - 1) * nb / 2) * tile_bytes;  // tiles (I, nb - 1), I = 0 .. nb - 1
  HIP_OK(hipMemsetAsync(static_cast<char*>(z) + first, 0, (size_t)nb * tile_bytes, s));
}

void launch_assemble_z64(const AsmArgs& a, double* z, int cap, hipStream_t s, int nthr) {
  launch_tile<double, double, 0>(a, cap, z, nullptr, nullptr, s, nthr);
}
void launch_assemble_z32(const AsmArgs& a, float* z, int cap, hipStream_t s) {
  launch_tile<float, float, 0>(a, cap, z, nullptr, nullptr, s);
}
void launch_assemble_z64p(const AsmArgs& a, double* z, int cap, hipStream_t s, int nthr) {
  DM_REQUIRE(2 * a.TB == kPT, DM_ERR_ARGUMENT, "packed storage needs assembly tiles of half the storage tile");
  launch_tile<double, double, 2>(a, cap, z, nullptr, nullptr, s, nthr);
}
void launch_assemble_z32p(const AsmArgs& a, float* z, int cap, hipStream_t s) {
  DM_REQUIRE(2 * a.TB == kPT, DM_ERR_ARGUMENT, "packed storage needs assembly tiles of half the storage tile");
  launch_tile<float, float, 2>(a, cap, z, nullptr, nullptr, s);
}
void launch_mf_apply(const AsmArgs& a, const double* x, double* y, int cap, hipStream_t s) {
  launch_tile<double, double, 1>(a, cap, nullptr, x, y, s);
}

void launch_assemble_multi(const std::vector<AsmArgs>& args, const std::vector<void*>& z, int form, int max_tris,
                           hipStream_t s) {
  if (form == DM_Z64 || form == DM_Z64P) {
    std::vector<double*> zz;
    for (void* p : z) zz.push_back(static_cast<double*>(p));
    if (form == DM_Z64)
      launch_tile_multi<double, double, 0>(args, zz, max_tris, s);
    else
      launch_tile_multi<double, double, 2>(args, zz, max_tris, s);
  } else {
    std::vector<float*> zz;
    for (void* p : z) zz.push_back(static_cast<float*>(p));
    if (form == DM_Z32)
      launch_tile_multi<float, float, 0>(args, zz, max_tris, s);
    else
      launch_tile_multi<float, float, 2>(args, zz, max_tris, s);
  }
}

size_t side_image_bytes(int nq, bool f64, int cap) {
  switch (nq) {
    case 1: return f64 ? side_bytes<double, 1>(cap) : side_bytes<float, 1>(cap);
    case 3: return f64 ? side_bytes<double, 3>(cap) : side_bytes<float, 3>(cap);
    case 4: return f64 ? side_bytes<double, 4>(cap) : side_bytes<float, 4>(cap);
    case 7: return f64 ? side_bytes<double, 7>(cap) : side_bytes<float, 7>(cap);
    default: dm_throw(DM_ERR_ARGUMENT, "unsupported quadrature size");
  }
}

void launch_build_sides(const AsmArgs& a, int cap, bool f64, char* img, hipStream_t s) {
  const i64 stride = (i64)side_image_bytes(a.nq, f64, cap);
  switch (a.nq) {
#define DM_CASE(NQ)                                                                          \
  case NQ:                                                                                   \
    if (f64)                                                                                 \
      k_build_sides<double, NQ><<<a.nblk, 64, 0, s>>>(a, cap, img, stride);                  \
    else                                                                                     \
      k_build_sides<float, NQ><<<a.nblk, 64, 0, s>>>(a, cap, img, stride);                   \
    break;
    DM_CASE(1)
    DM_CASE(3)
    DM_CASE(4)
    DM_CASE(7)
#undef DM_CASE
    default:
      dm_throw(DM_ERR_ARGUMENT, "unsupported quadrature size");
  }
  HIP_OK(hipGetLastError());
}
