// k_mpnn.h — the learned preconditioner's message-passing network: parameter
// layout and the launchers of its forward, backward and Adam kernels.
//
// Nodes are RWG bases, edges join bases that share a triangle (at most 4 per
// basis), signed by the product of the two bases' orientations on that
// triangle.  For a vector r (a residual or a Krylov basis vector) with
// near-field preconditioned z = S r:
//   rho = sqrt(mean |r|^2), q = sum r |r|, c = conj(q) / (|q| rho)
//   tau = sqrt(mean |c z|^2)
//   u_i = [Re c z_i / tau, Im c z_i / tau, Re c r_i, Im c r_i, g1_i, g2_i]
//   h^0 = W_e u;  l = 0..3:  m_i = sum_j s_ij h_j / max(deg_i, 1),
//                            h <- h + tanh(W_s h + W_n m + b)
//   y = W_d h^4;  d = (y_0 + i y_1) / c   (a correction in residual units)
//   M r = z + S d = S (r + d)
// g1 = k len / (2 pi / 10), g2 = (A+ + A-) / len^2.  Parameters are kept in
// FP32 (master copy for Adam) and used rounded to bfloat16; activations and
// accumulation are FP32.  M is scale- and phase-equivariant: M(a r) = a M(r)
// for complex a, through c.
//
// Training objective of a sample: the relative residual that one application
// of M leaves, ||Z M r - r||^2 / ||r||^2.  With A = Z S (the sample's mesh,
// node order, a complex64 copy on the device) and p = r - A r it is
// ||A y - c p||^2 / (|c|^2 ||r||^2), whose gradient in y is
// 2 A^H (A y - c p) / (|c|^2 ||r||^2).
#pragma once
#include <cstdint>

#include "dm_common.h"

namespace dmn {

constexpr int H = 64;       // hidden width
constexpr int L = 4;        // message-passing layers
constexpr int FIN = 6;      // input features
constexpr int DEG = 4;      // neighbours per node (padded with -1)
constexpr int OFF_WE = 0;
constexpr int LAYER_SZ = 2 * H * H + H;
constexpr int OFF_LAYER = H * FIN;
constexpr int OFF_WD = OFF_LAYER + L * LAYER_SZ;
constexpr int NPARAM = OFF_WD + 2 * H;
inline int off_ws(int l) { return OFF_LAYER + l * LAYER_SZ; }
inline int off_wn(int l) { return OFF_LAYER + l * LAYER_SZ + H * H; }
inline int off_b(int l) { return OFF_LAYER + l * LAYER_SZ + 2 * H * H; }

// Graph of one or more meshes over nodes [0, n): nbr (n x DEG, global node ids,
// -1 padding), sgn (n x DEG), dinv = 1 / max(deg, 1).
struct Graph {
  i64 n;
  const int* nbr;
  const float* sgn;
  const float* dinv;
};

// Per-node network activations kept for the backward pass (n nodes).
struct Cache {
  float* u;   // n x FIN
  float* h;   // (L + 1) x n x H
  float* t;   // L x n x H   tanh outputs
  float* m;   // L x n x H   aggregated messages
  float* y;   // n x 2
};

// Per-sample normaliser c (complex), den = |c|^2 ||r||^2 and tau for the
// samples [s0, s0 + ns) whose nodes are ptr[s] .. ptr[s + 1] (global node
// ids): cden[4 s + {0, 1, 2, 3}] = Re c, Im c, den, tau.
void stats(int ns, const int64_t* ptr, int64_t node0, const double* z, const double* r, double* cden, hipStream_t s);
// Input features u of nodes [0, n) of a batch whose first sample is s0;
// sample of local node i is smp[i] (global sample id).
void features(i64 n, const int* smp, int s0, const double* cden, const double* z, const double* r, const float* geom,
              float* u, hipStream_t s);
// Forward through the network; bf16: NPARAM parameters (bfloat16 bits).
// gnode0: global id of local node 0 (neighbour ids are global).
void forward(const Graph& g, i64 gnode0, const uint16_t* bf16, const Cache& c, float* hbuf, hipStream_t s);
// out[0] = sum of parts[0 .. n) (fixed order).
void sum(i64 n, const double* parts, double* out, hipStream_t s);
// Gradient of the loss w.r.t. the (rounded) parameters into grad (NPARAM floats).
// work: backward scratch of backward_floats(n) floats; part: outer-product
// partials of outer_floats() floats.
size_t backward_floats(i64 n);
size_t outer_floats();
void backward(const Graph& g, i64 gnode0, const uint16_t* bf16, const Cache& c, const float* gy, float* grad,
              float* work, float* part, hipStream_t s);
// One Adam step on the FP32 master parameters; refreshes the bfloat16 copy.
void adam(float* p, const float* grad, float* m, float* v, int step, float lr, uint16_t* bf16, hipStream_t s);
// bfloat16 copy of the parameters.
void to_bf16(const float* p, uint16_t* bf16, hipStream_t s);
// The correction d = (y0 + i y1) / c of one sample (c from cden[0..1]).
void correction(i64 n, const float* y, const double* cden, double* d, hipStream_t s);
// The n x n complex128 identity (column-major).
void identity(i64 n, double* out, hipStream_t s);
// out = a + b (complex128, n).
void add(i64 n, const double* a, const double* b, double* out, hipStream_t s);
// Node-order complex64 copy of an internal-order complex128 matrix (n x n,
// column-major): out[a + b n] = A[iofe[a] + iofe[b] n].
void gather_c64(i64 n, const double* a_int, const int* iofe, float* out, hipStream_t s);
// p = r - A r for k samples of one matrix (columns of n nodes, consecutive).
void res_rhs(i64 n, int k, const float* A, const double* r, double* p, hipStream_t s);
// For k samples of one matrix (columns of y, p consecutive; cden of those
// samples): q = A y - c p and loss parts |q_i|^2 / (den ns); when gy is not
// null also gy = 2 A^H q / (den ns).  work: 2 n k doubles.
void res_loss(i64 n, int k, const float* A, const float* y, const double* p, const double* cden, int ns, float* q,
              double* parts, float* gy, double* work, hipStream_t s);

}  // namespace dmn
