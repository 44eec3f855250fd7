// mlp.hpp — DSP-WENO network kernels outside the stage kernel: batched MFMA
// inference and the training forward/backward/Adam kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace tecno {

// Neural_Network::forward on n raw stencils with the 5-5-5-5-5 network on
// matrix cores (bf16 or fp16 operands, fp32 accumulation).
hipError_t mlp_mfma_batch(hipStream_t s, const unsigned short* packedA, const float* bias,
                          bool fp16, const double* u4, int64_t n, double* l, double* r);

// Training state on the device: fp32 master weights, Adam moments.
class MlpTrainer {
 public:
  int init(hipStream_t s, int nl, const int* widths, const double* W, const double* b, double lr,
           double beta1, double beta2, double eps, double wd, std::string& err);
  // forward + backward on n rows (21 doubles each); if update, one Adam step.
  // grad (if not null) receives the gradient, layout [4][8x8] then [4][8].
  int step(const double* rows, int64_t n, bool bf16, bool update, double* loss,
           std::vector<double>* grad, std::string& err);
  void get_weights(std::vector<double>& W, std::vector<double>& b);
  ~MlpTrainer();

 private:
  hipStream_t s_ = nullptr;
  int nl_ = 0, w_[5] = {0};
  float* dparam_ = nullptr;   // master weights [4][64] + biases [4][8]
  float* dm_ = nullptr;       // Adam first moment
  float* dv_ = nullptr;       // Adam second moment
  double* dgrad_ = nullptr;   // gradient accumulation (fp64)
  double* drows_ = nullptr;
  size_t rows_cap_ = 0;
  int64_t t_ = 0;
  double lr_ = 1e-3, b1_ = 0.5, b2_ = 0.9, eps_ = 1e-8, wd_ = 1e-5;
};

// rows (n x 21) = 15 features (inputs, x-vertices, y-vertices), 4 raw cells,
// 2 targets, as the trainer's data file holds them
int make_train_rows_host(const double* u4, const double* target, int64_t n, double* rows);

}  // namespace tecno
