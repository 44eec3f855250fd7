// mlp.hip — DSP-WENO network kernels (placeholder implementations are
// replaced by the matrix-core inference and training kernels).
#include "mlp.hpp"
#include "fv_device.hpp"

namespace tecno {

hipError_t mlp_mfma_batch(hipStream_t, const float*, const float*, int, bool, const double*,
                          int64_t, double*, double*) {
  return hipErrorNotSupported;
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
