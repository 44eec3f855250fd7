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
hipError_t mlp_mfma_batch(hipStream_t s, const unsigned short* packedA, bool fp16,
                          const double* u4, int64_t n, double* l, double* r);

// Training state on the device: fp32 master weights, Adam moments.  The
// trainer's parameter block is padded to 4 layers of 8x8 weights (row =
// output neuron) followed by 4 x 8 biases; padding stays exactly zero.
class MlpTrainer {
 public:
  static constexpr int PW = 8, PL = 4, NP = PL * PW * PW + PL * PW;
  int init(hipStream_t s, int nl, const int* widths, const double* W, const double* b, double lr,
           double beta1, double beta2, double eps, double wd, std::string& err);
  // forward + backward on n host rows (21 doubles each); if update, one Adam
  // step.  grad (if not null) receives the gradient in the padded layout.
  int step(const double* rows, int64_t n, bool bf16, bool update, double* loss,
           std::vector<double>* grad, std::string& err);
  // keep a data set of n host rows on the device
  int set_data(const double* rows, int64_t n, std::string& err);
  // one pass over the device data set in mini-batches of `batch` rows in a
  // seeded random order (the last batch may be shorter); mean batch loss
  int epoch(int64_t batch, uint64_t seed, bool bf16, double* mean_loss, std::string& err);
  void get_weights(std::vector<double>& W, std::vector<double>& b);
  ~MlpTrainer();

 private:
  int run_batch(const double* drows, const int* didx, int64_t row0, int64_t n, bool bf16,
                bool update, double* dgrad, std::string& err);
  hipStream_t s_ = nullptr;
  int nl_ = 0, w_[5] = {0};
  float* dparam_ = nullptr;   // master weights, padded layout
  float* dm_ = nullptr;       // Adam first moment
  float* dv_ = nullptr;       // Adam second moment
  float* dpart_ = nullptr;    // per-block gradient and loss partial sums
  size_t part_cap_ = 0;       // blocks
  double* dloss_ = nullptr;   // loss accumulator
  double* dgrad_ = nullptr;   // summed gradient (fp64)
  double* drows_ = nullptr;   // rows of a single step
  size_t rows_cap_ = 0;
  double* ddata_ = nullptr;   // device data set
  int* didx_ = nullptr;       // epoch order
  int64_t ndata_ = 0;
  int64_t t_ = 0;
  double lr_ = 1e-3, b1_ = 0.5, b2_ = 0.9, eps_ = 1e-8, wd_ = 1e-5;
};

// training rows (n x 21: 5 inputs, 5 x-vertices, 5 y-vertices, 4 cells,
// 2 targets) from host stencils u4 (n x 4) and targets (n x 2), on the device
int make_train_rows(hipStream_t s, const double* u4, const double* target, int64_t n,
                    double* rows, std::string& err);

}  // namespace tecno
