// k_mpnn.h — the learned preconditioner's message-passing network: parameter
// layout and the launchers of its forward, backward and Adam kernels.
//
// Nodes are RWG bases, edges join bases that share a triangle (at most 4 per
// basis), signed by the product of the two bases' orientations on that
// triangle.  For a residual r with near-field preconditioned z = S r:
//   rho = sqrt(mean |z|^2), q = sum z |z|, c = conj(q) / (|q| rho)
//   u_i = [Re c z_i, Im c z_i, Re c r_i, Im c r_i, g1_i, g2_i]
//   h^0 = W_e u;  l = 0..3:  m_i = sum_j s_ij h_j / max(deg_i, 1),
//                            h <- h + tanh(W_s h + W_n m + b)
//   y = W_d h^4;  M r = z + (y_0 + i y_1) / c
// g1 = k len / (2 pi / 10), g2 = (A+ + A-) / len^2.  Parameters are kept in
// FP32 (master copy for Adam) and used rounded to bfloat16; activations and
// accumulation are FP32.  The network is scale- and phase-equivariant in r
// through c.
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

// Per-sample normaliser c (complex) and loss denominator |c|^2 ||e||^2 for the
// samples [s0, s0 + ns) whose nodes are ptr[s] .. ptr[s + 1] (global node ids,
// e may be null).
void stats(int ns, const int64_t* ptr, int64_t node0, const double* z, const double* e, double* cden, hipStream_t s);
// u and (when e != null) target c (e - z) for nodes [0, n) of a batch whose
// first sample is s0; sample of local node i is smp[i] (global sample id).
void features(i64 n, const int* smp, int s0, const double* cden, const double* z, const double* r, const float* geom,
              const double* e, float* u, float* tgt, hipStream_t s);
// Forward through the network; bf16: NPARAM parameters (bfloat16 bits).
// gnode0: global id of local node 0 (neighbour ids are global).
void forward(const Graph& g, i64 gnode0, const uint16_t* bf16, const Cache& c, float* hbuf, hipStream_t s);
// Loss of the batch (mean over ns samples) and dL/dy; loss_parts: n doubles.
void loss(i64 n, const int* smp, int s0, const double* cden, int ns, const float* y, const float* tgt, float* gy,
          double* loss_parts, double* loss_out, hipStream_t s);
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
// Preconditioner output out = z + (y0 + i y1) / c (c from cden[0]).
void output(i64 n, const double* z, const float* y, const double* cden, double* out, hipStream_t s);

}  // namespace dmn
