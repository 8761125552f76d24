// Caller-side steps of the batched growth inversion, on the device, so that a
// whole Gauss-Newton iteration of every instance runs without the host:
//   growth : G_p = I + theta_i dG_p for the live slots of instance i
//            (the caller's growth parameterisation; the reference's is
//            Fg = g I, python/morphompm/integrate.py:31-34);
//   loss   : L_i = sum_p det F_p in FP64 and its cotangent det F F^-T
//            (python/morphompm/integrate.py:41-49);
//   grad   : dL_i / dtheta_i = sum_p <dG_p, G_bar_p> in FP64
//            (python/morphompm/integrate.py:37-38 for dG = I);
//   update : one damped Gauss-Newton iteration per instance
//            (python/morphompm/diff.py:14-23).
// Every sum has a fixed order -- a tree over the 256 slots of a chunk, then
// a strided sum and a tree over an instance's chunks -- so results are
// bitwise reproducible.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/morphompm_hip.h"
#include "mpm_internal.h"

namespace {

using mpm_internal::check_launch;
using mpm_internal::set_err;

constexpr int kChunk = MPM_SLOT_ALIGN;  // slots per chunk; one block per chunk

// Fixed-order sum of v over the 256 threads of a block; valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int off = kChunk / 2; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  return red[0];
}

// Instance of block b's chunk and whether this thread's slot holds a particle.
__device__ __forceinline__ int chunk_instance(const int32_t* slot_inst) {
  return __builtin_amdgcn_readfirstlane(slot_inst[blockIdx.x]);
}

template <typename T>
__global__ __launch_bounds__(kChunk) void k_growth_from_rates(const mpm_instance* inst,
                                                              const int32_t* slot_inst,
                                                              int64_t n_slots,
                                                              const double* theta,
                                                              const double* dG, T* G) {
  const int ii = chunk_instance(slot_inst);
  const int64_t p = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  const bool live = p - inst[ii].p_off < inst[ii].n;
  const double th = theta[ii];
  for (int k = 0; k < 9; ++k) {
    const double e = (k % 4 == 0) ? 1.0 : 0.0;
    G[(int64_t)k * n_slots + p] = (T)(live ? e + th * dG[(int64_t)k * n_slots + p] : e);
  }
}

// Per slot: det F and cof F = det F F^-T in FP64 from the F rows (15..23) of
// dyn; writes the whole cotangent state bar (cof F in the F rows of live
// slots, zero elsewhere) and the chunk's sum of det F.
template <typename T>
__global__ __launch_bounds__(kChunk) void k_loss_det(const mpm_instance* inst,
                                                     const int32_t* slot_inst, int64_t n_slots,
                                                     const T* dyn, T* bar, double* chunk_acc) {
  __shared__ double red[kChunk];
  const int ii = chunk_instance(slot_inst);
  const int64_t p = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  const bool live = p - inst[ii].p_off < inst[ii].n;
  double f[9], c[9], det = 0.0;
  for (int k = 0; k < 9; ++k) f[k] = live ? (double)dyn[(int64_t)(15 + k) * n_slots + p] : 0.0;
  c[0] = f[4] * f[8] - f[5] * f[7];
  c[1] = f[5] * f[6] - f[3] * f[8];
  c[2] = f[3] * f[7] - f[4] * f[6];
  c[3] = f[2] * f[7] - f[1] * f[8];
  c[4] = f[0] * f[8] - f[2] * f[6];
  c[5] = f[1] * f[6] - f[0] * f[7];
  c[6] = f[1] * f[5] - f[2] * f[4];
  c[7] = f[2] * f[3] - f[0] * f[5];
  c[8] = f[0] * f[4] - f[1] * f[3];
  if (live) det = f[0] * c[0] + f[1] * c[1] + f[2] * c[2];
  for (int r = 0; r < 15; ++r) bar[(int64_t)r * n_slots + p] = T(0);
  for (int k = 0; k < 9; ++k) bar[(int64_t)(15 + k) * n_slots + p] = live ? (T)c[k] : T(0);
  const double s = block_sum(det, red);
  if (threadIdx.x == 0) chunk_acc[blockIdx.x] = s;
}

// Per slot: <dG_p, G_bar_p> in FP64; the chunk's sum.
template <typename T>
__global__ __launch_bounds__(kChunk) void k_growth_grad(const mpm_instance* inst,
                                                        const int32_t* slot_inst,
                                                        int64_t n_slots, const T* G_bar,
                                                        const double* dG, double* chunk_acc) {
  __shared__ double red[kChunk];
  const int ii = chunk_instance(slot_inst);
  const int64_t p = (int64_t)blockIdx.x * kChunk + threadIdx.x;
  double s = 0.0;
  if (p - inst[ii].p_off < inst[ii].n)
    for (int k = 0; k < 9; ++k)
      s += dG[(int64_t)k * n_slots + p] * (double)G_bar[(int64_t)k * n_slots + p];
  s = block_sum(s, red);
  if (threadIdx.x == 0) chunk_acc[blockIdx.x] = s;
}

// out[i] = sum of instance i's chunk sums: thread t adds chunks t, t + 256,
// ... in order, then a tree over the threads.
__global__ __launch_bounds__(kChunk) void k_instance_sum(const mpm_instance* inst,
                                                         const double* chunk_acc, double* out) {
  __shared__ double red[kChunk];
  const mpm_instance& I = inst[blockIdx.x];
  const int64_t c0 = I.p_off / kChunk, nc = ((int64_t)I.n + kChunk - 1) / kChunk;
  double s = 0.0;
  for (int64_t c = threadIdx.x; c < nc; c += kChunk) s += chunk_acc[c0 + c];
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[blockIdx.x] = s;
}

// One Gauss-Newton iteration of every instance that has not converged
// (python/morphompm/diff.py:14-23):
//   resid = L - target; |resid| < tol: converged (theta kept);
//   |L'| < 1e-12 or L' not finite: theta += 1e-3 (resid < 0 ? 1 : -1);
//   else theta -= clip(resid / L', -0.5, 0.5) and theta is appended to the
//   instance's history.  iters[i] counts the iterations an instance took
//   part in, the reference's returned it + 1.
__global__ void k_gauss_newton(int32_t n_inst, const double* loss, const double* dloss,
                               const double* target, double tol, double* theta, int32_t* done,
                               int32_t* iters, double* hist, int32_t hist_cap,
                               int32_t* hist_len) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_inst || done[i]) return;
  iters[i] += 1;
  const double resid = loss[i] - target[i];
  const double d = dloss[i];
  if (fabs(resid) < tol) {
    done[i] = 1;
    return;
  }
  if (fabs(d) < 1e-12 || !isfinite(d)) {
    theta[i] += resid < 0.0 ? 1e-3 : -1e-3;
    return;
  }
  // fmin / fmax return the other operand for a NaN ratio, as Python's
  // min(0.5, nan) and max(-0.5, 0.5) do
  const double g = theta[i] - fmax(-0.5, fmin(0.5, resid / d));
  theta[i] = g;
  if (hist != nullptr && hist_len[i] < hist_cap) {
    hist[(int64_t)i * hist_cap + hist_len[i]] = g;
    hist_len[i] += 1;
  }
}

template <typename T>
int growth_impl(const mpm_batch* b, const double* theta, const double* dG, T* G, void* stream) {
  int rc = mpm_internal::validate_batch(b);
  if (rc) return rc;
  if (!theta || !dG || !G) return set_err(MPM_EINVAL, "growth_from_rates: NULL buffer");
  k_growth_from_rates<T><<<dim3((unsigned)(b->n_slots / kChunk)), kChunk, 0,
                           (hipStream_t)stream>>>(b->inst, b->slot_inst, b->n_slots, theta, dG,
                                                  G);
  return check_launch("growth_from_rates");
}

template <typename T>
int loss_impl(const mpm_batch* b, const T* dyn, T* bar, double* chunk_acc, double* loss,
              void* stream) {
  int rc = mpm_internal::validate_batch(b);
  if (rc) return rc;
  if (!dyn || !bar || !chunk_acc || !loss) return set_err(MPM_EINVAL, "loss_det: NULL buffer");
  if (bar == dyn) return set_err(MPM_EINVAL, "loss_det: bar must not alias dyn");
  hipStream_t st = (hipStream_t)stream;
  k_loss_det<T><<<dim3((unsigned)(b->n_slots / kChunk)), kChunk, 0, st>>>(
      b->inst, b->slot_inst, b->n_slots, dyn, bar, chunk_acc);
  k_instance_sum<<<dim3((unsigned)b->n_inst), kChunk, 0, st>>>(b->inst, chunk_acc, loss);
  return check_launch("loss_det");
}

template <typename T>
int grad_impl(const mpm_batch* b, const T* G_bar, const double* dG, double* chunk_acc,
              double* grad, void* stream) {
  int rc = mpm_internal::validate_batch(b);
  if (rc) return rc;
  if (!G_bar || !dG || !chunk_acc || !grad) return set_err(MPM_EINVAL, "growth_grad: NULL buffer");
  hipStream_t st = (hipStream_t)stream;
  k_growth_grad<T><<<dim3((unsigned)(b->n_slots / kChunk)), kChunk, 0, st>>>(
      b->inst, b->slot_inst, b->n_slots, G_bar, dG, chunk_acc);
  k_instance_sum<<<dim3((unsigned)b->n_inst), kChunk, 0, st>>>(b->inst, chunk_acc, grad);
  return check_launch("growth_grad");
}

}  // namespace

extern "C" {

int32_t mpm_growth_from_rates_f32(const mpm_batch* b, const double* theta, const double* dG,
                                  float* G, void* stream) {
  return growth_impl<float>(b, theta, dG, G, stream);
}
int32_t mpm_growth_from_rates_f64(const mpm_batch* b, const double* theta, const double* dG,
                                  double* G, void* stream) {
  return growth_impl<double>(b, theta, dG, G, stream);
}

int32_t mpm_loss_det_f32(const mpm_batch* b, const float* dyn, float* bar, double* chunk_acc,
                         double* loss, void* stream) {
  return loss_impl<float>(b, dyn, bar, chunk_acc, loss, stream);
}
int32_t mpm_loss_det_f64(const mpm_batch* b, const double* dyn, double* bar, double* chunk_acc,
                         double* loss, void* stream) {
  return loss_impl<double>(b, dyn, bar, chunk_acc, loss, stream);
}

int32_t mpm_growth_grad_f32(const mpm_batch* b, const float* G_bar, const double* dG,
                            double* chunk_acc, double* grad, void* stream) {
  return grad_impl<float>(b, G_bar, dG, chunk_acc, grad, stream);
}
int32_t mpm_growth_grad_f64(const mpm_batch* b, const double* G_bar, const double* dG,
                            double* chunk_acc, double* grad, void* stream) {
  return grad_impl<double>(b, G_bar, dG, chunk_acc, grad, stream);
}

int32_t mpm_gauss_newton(int32_t n_inst, const double* loss, const double* dloss,
                         const double* target, double tol, double* theta, int32_t* done,
                         int32_t* iters, double* hist, int32_t hist_cap, int32_t* hist_len,
                         void* stream) {
  if (n_inst <= 0) return set_err(MPM_EINVAL, "gauss_newton: n_inst must be positive");
  if (!loss || !dloss || !target || !theta || !done || !iters)
    return set_err(MPM_EINVAL, "gauss_newton: NULL buffer");
  if (hist != nullptr && (hist_cap <= 0 || hist_len == nullptr))
    return set_err(MPM_EINVAL, "gauss_newton: history needs hist_cap > 0 and hist_len");
  if (!(tol >= 0.0)) return set_err(MPM_EINVAL, "gauss_newton: tol must be >= 0");
  k_gauss_newton<<<dim3((unsigned)((n_inst + 255) / 256)), 256, 0, (hipStream_t)stream>>>(
      n_inst, loss, dloss, target, tol, theta, done, iters, hist, hist_cap, hist_len);
  return check_launch("gauss_newton");
}

}  // extern "C"
