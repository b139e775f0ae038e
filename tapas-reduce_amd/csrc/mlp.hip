// mlp.hip -- the learned (ncrit, theta) picker: tree statistics of a trial
// tree, and a 16 -> 64 -> 64 -> 6 ReLU network trained and evaluated on the
// GPU by the kernels of this file.  It has no counterpart in the reference,
// whose knobs ncrit and theta are fixed command-line values
// (sample/exafmm-dev-13274dd4ac68/include/args.h:101-158); the nearest
// reference code is the leaf depth histogram of ShowHistogram
// (cpp/include/tapas/hot/buildtree.h:89-141).
//
// Outputs: 5 logits over ncrit in {16, 32, 64, 128, 256} and one value mapped
// to theta = 0.2 + 0.4 sigmoid(o5).  Loss per sample: cross entropy of the
// ncrit class + (theta - theta_label)^2 / 0.04.
//
// Training step (deterministic): one wave64 per sample computes its forward
// pass (lane j = hidden unit j) and writes its parameter gradient to its own
// row; a second kernel sums the rows in sample order and applies Adam.
#include <cmath>

#include "tfmm_device.h"
#include "tfmm_kernels.h"

namespace tfmm {

constexpr int MLP_IN = 16, MLP_H = 64, MLP_OUT = 6, MLP_CLASSES = 5;
// parameter layout: W1[64][16] b1[64] W2[64][64] b2[64] W3[6][64] b3[6]
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_OUT * MLP_H;
constexpr int MLP_NPARAM = OFF_B3 + MLP_OUT;  // 5638

int mlp_nparam() { return MLP_NPARAM; }

namespace {

__device__ inline float sigmoidf_(float x) { return 1.f / (1.f + expf(-x)); }

// forward of one sample by one wave; lane j = hidden unit j.  Returns the
// outputs in o[0..5] (valid in every lane) and keeps h1, h2 in LDS.
__device__ inline void mlp_forward(const float *W, const float *x, float *h1, float *h2,
                                   float *o, int lane) {
  float a = W[OFF_B1 + lane];
  for (int k = 0; k < MLP_IN; k++) a += W[OFF_W1 + lane * MLP_IN + k] * x[k];
  h1[lane] = a > 0.f ? a : 0.f;
  __syncthreads();
  float b = W[OFF_B2 + lane];
  for (int k = 0; k < MLP_H; k++) b += W[OFF_W2 + lane * MLP_H + k] * h1[k];
  h2[lane] = b > 0.f ? b : 0.f;
  __syncthreads();
  for (int i = 0; i < MLP_OUT; i++) {
    float v = W[OFF_W3 + i * MLP_H + lane] * h2[lane];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    o[i] = v + W[OFF_B3 + i];
  }
}

__global__ __launch_bounds__(64) void k_mlp_grad(const float *W, const float *X,
                                                 const int32_t *cls, const float *theta, int n,
                                                 float *G, float *loss) {
  const int s = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ float x[MLP_IN], h1[MLP_H], h2[MLP_H], dh2[MLP_H], dout[MLP_OUT];
  if (lane < MLP_IN) x[lane] = X[(int64_t)s * MLP_IN + lane];
  __syncthreads();
  float o[MLP_OUT];
  mlp_forward(W, x, h1, h2, o, lane);
  // loss and output gradient
  float mx = o[0];
  for (int i = 1; i < MLP_CLASSES; i++) mx = fmaxf(mx, o[i]);
  float se = 0.f;
  for (int i = 0; i < MLP_CLASSES; i++) se += expf(o[i] - mx);
  const int c = cls[s];
  const float sg = sigmoidf_(o[5]);
  const float th = 0.2f + 0.4f * sg;
  const float dth = th - theta[s];
  if (lane == 0) {
    loss[s] = -(o[c] - mx - logf(se)) + dth * dth / 0.04f;
    for (int i = 0; i < MLP_CLASSES; i++) dout[i] = expf(o[i] - mx) / se - (i == c ? 1.f : 0.f);
    dout[5] = 2.f * dth / 0.04f * 0.4f * sg * (1.f - sg);
  }
  __syncthreads();
  float *g = G + (int64_t)s * MLP_NPARAM;
  // layer 3
  float d2 = 0.f;
  for (int i = 0; i < MLP_OUT; i++) {
    g[OFF_W3 + i * MLP_H + lane] = dout[i] * h2[lane];
    d2 += W[OFF_W3 + i * MLP_H + lane] * dout[i];
  }
  if (lane < MLP_OUT) g[OFF_B3 + lane] = dout[lane];
  d2 = h2[lane] > 0.f ? d2 : 0.f;
  dh2[lane] = d2;
  __syncthreads();
  // layer 2
  for (int k = 0; k < MLP_H; k++) g[OFF_W2 + lane * MLP_H + k] = d2 * h1[k];
  g[OFF_B2 + lane] = d2;
  float d1 = 0.f;
  for (int k = 0; k < MLP_H; k++) d1 += W[OFF_W2 + k * MLP_H + lane] * dh2[k];
  d1 = h1[lane] > 0.f ? d1 : 0.f;
  // layer 1
  for (int k = 0; k < MLP_IN; k++) g[OFF_W1 + lane * MLP_IN + k] = d1 * x[k];
  g[OFF_B1 + lane] = d1;
}

// mean gradient over the n samples (in sample order) and one Adam step
__global__ void k_mlp_adam(float *W, const float *G, int n, float *m, float *v, float lr,
                           float b1, float b2, float eps, int step, float *gmean) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= MLP_NPARAM) return;
  float g = 0.f;
  for (int s = 0; s < n; s++) g += G[(int64_t)s * MLP_NPARAM + p];
  g /= (float)n;
  if (gmean) gmean[p] = g;
  if (lr == 0.f) return;
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  const float mh = mm / (1.f - powf(b1, (float)step));
  const float vh = vv / (1.f - powf(b2, (float)step));
  W[p] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ __launch_bounds__(64) void k_mlp_forward(const float *W, const float *X, int n,
                                                    float *out) {
  const int s = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ float x[MLP_IN], h1[MLP_H], h2[MLP_H];
  if (lane < MLP_IN) x[lane] = X[(int64_t)s * MLP_IN + lane];
  __syncthreads();
  float o[MLP_OUT];
  mlp_forward(W, x, h1, h2, o, lane);
  if (lane < MLP_OUT) {
    float v = 0.f;
    for (int i = 0; i < MLP_OUT; i++)
      if (i == lane) v = o[i];
    out[(int64_t)s * MLP_OUT + lane] = v;
  }
}

// Tree statistics of a built tree (features of the picker):
//  acc[0..7]  leaves (non-empty) per depth bin: <=3, 4, ..., 9, >=10
//  acc[8]     non-empty leaves      acc[9]  empty leaves
//  acc[10..12] sum of occupancy^1..3 (occupancy = nb / ncrit)
//  acc[13]    sum of leaf depths    acc[14] bodies in leaves at depth >= 8
__global__ void k_tree_stats(Cells c, int32_t ncells, int ncrit, double *acc) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncells || c.child[i] >= 0) return;
  const int nb = c.nb[i];
  if (nb == 0) {
    atomicAdd(&acc[9], 1.0);
    return;
  }
  const int d = c.depth[i];
  const int bin = d <= 3 ? 0 : (d >= 10 ? 7 : d - 3);
  atomicAdd(&acc[bin], 1.0);
  atomicAdd(&acc[8], 1.0);
  const double o = (double)nb / ncrit;
  atomicAdd(&acc[10], o);
  atomicAdd(&acc[11], o * o);
  atomicAdd(&acc[12], o * o * o);
  atomicAdd(&acc[13], (double)d);
  if (d >= 8) atomicAdd(&acc[14], (double)nb);
}

}  // namespace

hipError_t launch_mlp_grad(const float *W, const float *X, const int32_t *cls,
                           const float *theta, int n, float *G, float *loss, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_mlp_grad<<<n, 64, 0, s>>>(W, X, cls, theta, n, G, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float *W, const float *G, int n, float *m, float *v, float lr,
                           int step, float *gmean, hipStream_t s) {
  k_mlp_adam<<<(MLP_NPARAM + 255) / 256, 256, 0, s>>>(W, G, n, m, v, lr, 0.9f, 0.999f, 1e-8f,
                                                      step, gmean);
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const float *W, const float *X, int n, float *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  k_mlp_forward<<<n, 64, 0, s>>>(W, X, n, out);
  return hipGetLastError();
}

hipError_t launch_tree_stats(Cells c, int32_t ncells, int ncrit, double *acc, hipStream_t s) {
  if (ncells <= 0) return hipSuccess;
  k_tree_stats<<<(ncells + 255) / 256, 256, 0, s>>>(c, ncells, ncrit, acc);
  return hipGetLastError();
}

}  // namespace tfmm
