// picker.hip -- C ABI of the learned (ncrit, theta) picker: an FP32 MLP whose
// forward, backward and Adam update run in the kernels of mlp.hip.
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/tfmm.h"
#include "tfmm_internal.h"
#include "tfmm_kernels.h"

using namespace tfmm;

struct tfmm_picker {
  int device = 0;
  float theta_max = 0.6f;  // largest theta tfmm_pick_params may return
  hipStream_t stream = nullptr;
  std::string err;
  DevBuf W, m, v, X, cls, theta, G, loss, gmean, out;
  int step = 0;
};

namespace {

int pfail(tfmm_picker *p, int code, const std::string &msg) {
  if (p) p->err = msg;
  return code;
}

#define PTRY(expr)                                                              \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
      return pfail(p, _e == hipErrorOutOfMemory ? TFMM_ENOMEM : TFMM_EHIP,      \
                   std::string(#expr) + ": " + hipGetErrorString(_e));          \
  } while (0)

int upload_batch(tfmm_picker *p, const float *X, const int32_t *cls, const float *theta, int n) {
  PTRY(dev_reserve(p->X, (size_t)n * 16 * 4));
  PTRY(hipMemcpyAsync(p->X.ptr, X, (size_t)n * 16 * 4, hipMemcpyHostToDevice, p->stream));
  if (cls) {
    PTRY(dev_reserve(p->cls, (size_t)n * 4));
    PTRY(hipMemcpyAsync(p->cls.ptr, cls, (size_t)n * 4, hipMemcpyHostToDevice, p->stream));
  }
  if (theta) {
    PTRY(dev_reserve(p->theta, (size_t)n * 4));
    PTRY(hipMemcpyAsync(p->theta.ptr, theta, (size_t)n * 4, hipMemcpyHostToDevice, p->stream));
  }
  return TFMM_OK;
}

int check_labels(tfmm_picker *p, const int32_t *cls, const float *theta, int n) {
  for (int i = 0; i < n; i++) {
    if (cls[i] < 0 || cls[i] >= 5) return pfail(p, TFMM_EINVAL, "ncrit class out of range");
    if (!(theta[i] >= 0.2f && theta[i] <= 0.6f))
      return pfail(p, TFMM_EINVAL, "theta label outside [0.2, 0.6]");
  }
  return TFMM_OK;
}

}  // namespace

namespace tfmm {
const float *picker_weights(tfmm_picker *p) { return p ? dp<float>(p->W) : nullptr; }
float picker_theta_max(const tfmm_picker *p) { return p ? p->theta_max : 0.6f; }
}  // namespace tfmm

extern "C" {

int64_t tfmm_picker_nparam(void) { return mlp_nparam(); }

int tfmm_picker_create(int device, uint64_t seed, tfmm_picker **out) {
  if (!out) return TFMM_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return TFMM_EHIP;
  if (device < 0 || device >= ndev) return TFMM_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return TFMM_EHIP;
  tfmm_picker *p = new tfmm_picker();
  p->device = device;
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    delete p;
    return TFMM_EHIP;
  }
  // Glorot-uniform weights, zero biases, from a seeded generator
  const int np = mlp_nparam();
  std::vector<float> w(np, 0.f);
  std::mt19937_64 rng(seed);
  auto fill = [&](int off, int fan_out, int fan_in) {
    const float a = std::sqrt(6.f / (fan_in + fan_out));
    std::uniform_real_distribution<float> u(-a, a);
    for (int i = 0; i < fan_out * fan_in; i++) w[off + i] = u(rng);
  };
  fill(0, 64, 16);
  fill(64 * 16 + 64, 64, 64);
  fill(64 * 16 + 64 + 64 * 64 + 64, 6, 64);
  if (dev_reserve(p->W, np * 4) != hipSuccess || dev_reserve(p->m, np * 4) != hipSuccess ||
      dev_reserve(p->v, np * 4) != hipSuccess || dev_reserve(p->gmean, np * 4) != hipSuccess ||
      hipMemcpy(p->W.ptr, w.data(), np * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(p->m.ptr, 0, np * 4) != hipSuccess || hipMemset(p->v.ptr, 0, np * 4) != hipSuccess) {
    tfmm_picker_destroy(p);
    return TFMM_ENOMEM;
  }
  *out = p;
  return TFMM_OK;
}

void tfmm_picker_destroy(tfmm_picker *p) {
  if (!p) return;
  (void)hipSetDevice(p->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  DevBuf *b[] = {&p->W, &p->m, &p->v, &p->X, &p->cls, &p->theta, &p->G, &p->loss, &p->gmean, &p->out};
  for (DevBuf *x : b) dev_release(*x);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

const char *tfmm_picker_last_error(const tfmm_picker *p) { return p ? p->err.c_str() : "null picker"; }

int tfmm_picker_set_weights(tfmm_picker *p, const float *w, int64_t n) {
  if (!p) return TFMM_EINVAL;
  if (!w || n != mlp_nparam()) return pfail(p, TFMM_EINVAL, "weights must have tfmm_picker_nparam() entries");
  PTRY(hipMemcpyAsync(p->W.ptr, w, n * 4, hipMemcpyHostToDevice, p->stream));
  PTRY(hipMemsetAsync(p->m.ptr, 0, n * 4, p->stream));
  PTRY(hipMemsetAsync(p->v.ptr, 0, n * 4, p->stream));
  PTRY(hipStreamSynchronize(p->stream));
  p->step = 0;
  return TFMM_OK;
}

int tfmm_picker_set_theta_max(tfmm_picker *p, float theta_max) {
  if (!p) return TFMM_EINVAL;
  if (!(theta_max >= 0.2f && theta_max <= 0.6f))
    return pfail(p, TFMM_EINVAL, "theta_max outside [0.2, 0.6]");
  p->theta_max = theta_max;
  return TFMM_OK;
}

int tfmm_picker_get_weights(tfmm_picker *p, float *w, int64_t n) {
  if (!p) return TFMM_EINVAL;
  if (!w || n != mlp_nparam()) return pfail(p, TFMM_EINVAL, "weights must have tfmm_picker_nparam() entries");
  PTRY(hipMemcpyAsync(w, p->W.ptr, n * 4, hipMemcpyDeviceToHost, p->stream));
  PTRY(hipStreamSynchronize(p->stream));
  return TFMM_OK;
}

int tfmm_picker_gradient(tfmm_picker *p, const float *X, const int32_t *cls, const float *theta,
                         int n, float *grad, float *loss) {
  if (!p) return TFMM_EINVAL;
  if (n < 1 || !X || !cls || !theta) return pfail(p, TFMM_EINVAL, "empty or null batch");
  int rc = check_labels(p, cls, theta, n);
  if (!rc) rc = upload_batch(p, X, cls, theta, n);
  if (rc) return rc;
  const int np = mlp_nparam();
  PTRY(dev_reserve(p->G, (size_t)n * np * 4));
  PTRY(dev_reserve(p->loss, (size_t)n * 4));
  PTRY(launch_mlp_grad(dp<float>(p->W), dp<float>(p->X), dp<int32_t>(p->cls), dp<float>(p->theta),
                       n, dp<float>(p->G), dp<float>(p->loss), p->stream));
  PTRY(launch_mlp_adam(dp<float>(p->W), dp<float>(p->G), n, dp<float>(p->m), dp<float>(p->v), 0.f,
                       1, dp<float>(p->gmean), p->stream));
  if (grad) PTRY(hipMemcpyAsync(grad, p->gmean.ptr, np * 4, hipMemcpyDeviceToHost, p->stream));
  if (loss) PTRY(hipMemcpyAsync(loss, p->loss.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, p->stream));
  PTRY(hipStreamSynchronize(p->stream));
  return TFMM_OK;
}

int tfmm_picker_train(tfmm_picker *p, const float *X, const int32_t *cls, const float *theta,
                      int n, int steps, float lr, float *loss_per_step) {
  if (!p) return TFMM_EINVAL;
  if (n < 1 || !X || !cls || !theta || steps < 0 || !(lr > 0))
    return pfail(p, TFMM_EINVAL, "bad training arguments");
  int rc = check_labels(p, cls, theta, n);
  if (!rc) rc = upload_batch(p, X, cls, theta, n);
  if (rc) return rc;
  const int np = mlp_nparam();
  PTRY(dev_reserve(p->G, (size_t)n * np * 4));
  PTRY(dev_reserve(p->loss, (size_t)n * 4));
  std::vector<float> lh(n);
  for (int it = 0; it < steps; it++) {
    PTRY(launch_mlp_grad(dp<float>(p->W), dp<float>(p->X), dp<int32_t>(p->cls),
                         dp<float>(p->theta), n, dp<float>(p->G), dp<float>(p->loss), p->stream));
    p->step++;
    PTRY(launch_mlp_adam(dp<float>(p->W), dp<float>(p->G), n, dp<float>(p->m), dp<float>(p->v),
                         lr, p->step, nullptr, p->stream));
    if (loss_per_step) {
      PTRY(hipMemcpyAsync(lh.data(), p->loss.ptr, (size_t)n * 4, hipMemcpyDeviceToHost, p->stream));
      PTRY(hipStreamSynchronize(p->stream));
      double s = 0;
      for (int i = 0; i < n; i++) s += lh[i];
      loss_per_step[it] = (float)(s / n);
    }
  }
  PTRY(hipStreamSynchronize(p->stream));
  return TFMM_OK;
}

int tfmm_picker_predict(tfmm_picker *p, const float *X, int n, float *out) {
  if (!p) return TFMM_EINVAL;
  if (n < 1 || !X || !out) return pfail(p, TFMM_EINVAL, "empty or null batch");
  int rc = upload_batch(p, X, nullptr, nullptr, n);
  if (rc) return rc;
  PTRY(dev_reserve(p->out, (size_t)n * 6 * 4));
  PTRY(launch_mlp_forward(dp<float>(p->W), dp<float>(p->X), n, dp<float>(p->out), p->stream));
  PTRY(hipMemcpyAsync(out, p->out.ptr, (size_t)n * 6 * 4, hipMemcpyDeviceToHost, p->stream));
  PTRY(hipStreamSynchronize(p->stream));
  return TFMM_OK;
}

}  // extern "C"
