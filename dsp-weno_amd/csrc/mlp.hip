// mlp.hip — DSP-WENO network kernels (placeholder implementations are
// replaced by the matrix-core inference and training kernels).
#include "mlp.hpp"
#include "fv_mixed.hpp"
#include "mlp_mfma.hpp"

namespace tecno {

// Neural_Network::forward on n raw stencils (no zero-jump skip), features and
// head in FP32 from FP64 cell differences, the network on the matrix cores.
// Every wave runs whole; lanes past n repeat the last stencil.
template <bool H>
__global__ __launch_bounds__(256) void k_mlp_mfma(const unsigned short* A, const float* bias,
                                                  const double* u4, int64_t n, double* lo,
                                                  double* ro) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t q = i < n ? i : n - 1;
  mfma::Net5<H> net;
  net.load(A, bias, threadIdx.x & 63);
  const double u0 = u4[4 * q], u1 = u4[4 * q + 1], u2 = u4[4 * q + 2], u3 = u4[4 * q + 3];
  const float d0 = (float)(u1 - u0), d1 = (float)(u2 - u1), d2 = (float)(u3 - u2);
  const float mx = (float)fmax(fmax(fabs(u0), fabs(u1)), fmax(fabs(u2), fabs(u3)));
  float st[5], xv[5], yv[5], lg[5], C1, C2;
  fast::features(d0, d1, d2, mx, st, xv, yv);
#pragma unroll
  for (int k = 0; k < 5; ++k) st[k] = (fabsf(st[k]) < INFINITY) ? st[k] : 0.0f;
  net.eval(st, lg);
  fast::head_c(lg, xv, yv, C1, C2);
  const double w1 = 0.25 - 2.0 * (double)C1, tw0 = 0.25 - 2.0 * (double)C2;
  if (i < n) {
    lo[i] = u1 + 0.5 * (1.0 - w1) * (u2 - u1) + 0.5 * w1 * (u1 - u0);
    ro[i] = u2 - 0.5 * (1.0 - tw0) * (u2 - u1) - 0.5 * tw0 * (u3 - u2);
  }
}

hipError_t mlp_mfma_batch(hipStream_t s, const unsigned short* A, const float* bias, bool fp16,
                          const double* u4, int64_t n, double* l, double* r) {
  if (!A) return hipErrorInvalidValue;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (fp16)
    hipLaunchKernelGGL((k_mlp_mfma<true>), dim3(nb), dim3(256), 0, s, A, bias, u4, n, l, r);
  else
    hipLaunchKernelGGL((k_mlp_mfma<false>), dim3(nb), dim3(256), 0, s, A, bias, u4, n, l, r);
  return hipGetLastError();
}

int MlpTrainer::init(hipStream_t, int, const int*, const double*, const double*, double, double,
                     double, double, double, std::string& err) {
  err = "training kernels not built";
  return -1;
}
int MlpTrainer::step(const double*, int64_t, bool, bool, double*, std::vector<double>*,
                     std::string& err) {
  err = "training kernels not built";
  return -1;
}
void MlpTrainer::get_weights(std::vector<double>&, std::vector<double>&) {}
MlpTrainer::~MlpTrainer() {}

int make_train_rows_host(const double*, const double*, int64_t, double*) { return -1; }

}  // namespace tecno
