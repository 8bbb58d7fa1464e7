// k_fields.hip — plane-wave right-hand sides and radiation vectors.
//
// With f = sigma * len/(2A) * (r - v_e) on each half of a basis (sigma already
// folded with the mesh/RWG area ratio in hsig), the reference's per-point
// terms reduce to
//   v_n   = -len_n sum_halves hsig sum_q w_q (r_q - v_e) . E_inc(r_q)
//           (assemble_plane_wave, src/assembly/Excitation.jl:2829-2896;
//            E_inc = pol E0 exp(-i k.r), src/assembly/Excitation.jl:1517-1547)
//   g_n(r^) = (i k eta0/4pi) len_n sum_halves hsig sum_q w_q
//             r^ x (r^ x (r_q - v_e)) exp(+i k r^.r_q) / |r^|^2
//           (radiation_vectors, src/postprocessing/FarField.jl:425-574)
#include <hip/hip_runtime.h>

#include "dm_common.h"
#include "k_linalg.h"

namespace {

struct PwArgs {
  i64 n;
  int nq, M;
  const int* tp;
  const int* tm;
  const int* ep;
  const int* em;
  const double* len;
  const double* hsig;
  const double* rq;
  const double* vtx;
  const double* kvec;  // 3*M
  const double* pol;   // 3*M
  const double* E0;    // M
  const int* out_idx;
  double w[7];
};

__global__ __launch_bounds__(256) void k_planewave(PwArgs a, double2* __restrict__ V, i64 ldv) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= a.n * a.M) return;
  i64 i = idx % a.n;
  int m = (int)(idx / a.n);
  const double* kv = a.kvec + 3 * m;
  const double* pl = a.pol + 3 * m;
  double E0 = a.E0[m];
  double accr = 0, acci = 0;
  for (int side = 0; side < 2; ++side) {
    i64 t = side ? a.tm[i] : a.tp[i];
    int e = side ? a.em[i] : a.ep[i];
    double s = a.hsig[3 * t + e];
    const double* v = a.vtx + 9 * t + 3 * e;
    double hr = 0, hi = 0;
    for (int q = 0; q < a.nq; ++q) {
      const double* r = a.rq + (t * a.nq + q) * 3;
      double arg = kv[0] * r[0] + kv[1] * r[1] + kv[2] * r[2];
      double sn, cs;
      sincos(arg, &sn, &cs);
      double fe = (r[0] - v[0]) * pl[0] + (r[1] - v[1]) * pl[1] + (r[2] - v[2]) * pl[2];
      double c = a.w[q] * fe * E0;
      hr += c * cs;
      hi -= c * sn;
    }
    accr += s * hr;
    acci += s * hi;
  }
  double l = a.len[i];
  V[(a.out_idx ? (i64)a.out_idx[i] : i) + (i64)m * ldv] = make_double2(-l * accr, -l * acci);
}

struct RadArgs {
  i64 n, nomega;
  int nq;
  const int* tp;
  const int* tm;
  const int* ep;
  const int* em;
  const int* ext;
  const double* len;
  const double* hsig;
  const double* rq;
  const double* vtx;
  const double* rhat;
  double k, eta0;
  double w[7];
};

__global__ __launch_bounds__(256) void k_radiation(RadArgs a, double2* __restrict__ G) {
  i64 idx = blockIdx.x * (i64)blockDim.x + threadIdx.x;
  if (idx >= a.n * a.nomega) return;
  i64 i = idx / a.nomega;
  i64 qd = idx % a.nomega;
  double rh[3] = {a.rhat[3 * qd], a.rhat[3 * qd + 1], a.rhat[3 * qd + 2]};
  double rn2 = rh[0] * rh[0] + rh[1] * rh[1] + rh[2] * rh[2];
  double tr[3] = {0, 0, 0}, ti[3] = {0, 0, 0};
  for (int side = 0; side < 2; ++side) {
    i64 t = side ? a.tm[i] : a.tp[i];
    int e = side ? a.em[i] : a.ep[i];
    double s = a.hsig[3 * t + e];
    const double* v = a.vtx + 9 * t + 3 * e;
    for (int q = 0; q < a.nq; ++q) {
      const double* r = a.rq + (t * a.nq + q) * 3;
      double sn, cs;
      sincos(a.k * (rh[0] * r[0] + rh[1] * r[1] + rh[2] * r[2]), &sn, &cs);
      double f[3] = {r[0] - v[0], r[1] - v[1], r[2] - v[2]};
      // r^ x (r^ x f) = r^ (r^.f) - f |r^|^2
      double rf = rh[0] * f[0] + rh[1] * f[1] + rh[2] * f[2];
      double c = s * a.w[q] / rn2;
      for (int j = 0; j < 3; ++j) {
        double b = (rh[j] * rf - f[j] * rn2) * c;
        tr[j] += b * cs;
        ti[j] += b * sn;
      }
    }
  }
  // prefactor i k eta0 / (4 pi) * len
  double pf = a.k * a.eta0 * 0.07957747154594767 * a.len[i];
  i64 col = a.ext[i];
  i64 rows = 3 * a.nomega;
  for (int j = 0; j < 3; ++j) G[3 * qd + j + col * rows] = make_double2(-pf * ti[j], pf * tr[j]);
}

}  // namespace

namespace dmf {

void planewave(const dm_rwg& r, const TriTables& t, const double* kvec, const double* pol, const double* E0, int M,
               const int* out_idx, double* V, i64 ldv, hipStream_t s) {
  PwArgs a;
  a.n = r.n;
  a.nq = t.nq;
  a.M = M;
  a.tp = r.d_tplus_i.p;
  a.tm = r.d_tminus_i.p;
  a.ep = r.d_eplus_i.p;
  a.em = r.d_eminus_i.p;
  a.len = r.d_len.p;
  a.hsig = r.d_hsig.p;
  a.rq = t.rq.p;
  a.vtx = t.vtx.p;
  a.kvec = kvec;
  a.pol = pol;
  a.E0 = E0;
  a.out_idx = out_idx;
  double xi[14], w[7];
  dmh::quad_rule(t.quad_order, xi, w);
  for (int i = 0; i < 7; ++i) a.w[i] = i < t.nq ? w[i] : 0.0;
  i64 total = r.n * M;
  k_planewave<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(a, reinterpret_cast<double2*>(V), ldv);
  HIP_OK(hipGetLastError());
}

void radiation_vectors(const dm_rwg& r, const TriTables& t, const double* rhat, i64 nomega, double k, double eta0,
                       double* G, hipStream_t s) {
  RadArgs a;
  a.n = r.n;
  a.nomega = nomega;
  a.nq = t.nq;
  a.tp = r.d_tplus_i.p;
  a.tm = r.d_tminus_i.p;
  a.ep = r.d_eplus_i.p;
  a.em = r.d_eminus_i.p;
  a.ext = r.d_ext.p;
  a.len = r.d_len.p;
  a.hsig = r.d_hsig.p;
  a.rq = t.rq.p;
  a.vtx = t.vtx.p;
  a.rhat = rhat;
  a.k = k;
  a.eta0 = eta0;
  double xi[14], w[7];
  dmh::quad_rule(t.quad_order, xi, w);
  for (int i = 0; i < 7; ++i) a.w[i] = i < t.nq ? w[i] : 0.0;
  i64 total = r.n * nomega;
  k_radiation<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(a, reinterpret_cast<double2*>(G));
  HIP_OK(hipGetLastError());
}

}  // namespace dmf
