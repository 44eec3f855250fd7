// ens_fp32.hip — FP32 ensemble stage kernels (built with FP contraction on,
// as the FP32 single-problem kernels).
#include "ens_kernels.hpp"

namespace tecno {

template hipError_t ens_launch<ScalarPh<float, 2>>(hipStream_t, int, int, const void*,
                                                    const EnsWork*, int, int, int);
template int ens_grid_of<ScalarPh<float, 2>>(int, int, int);
template hipError_t ens_launch<ScalarPh<float, 2, 1>>(hipStream_t, int, int, const void*,
                                                       const EnsWork*, int, int, int);
template int ens_grid_of<ScalarPh<float, 2, 1>>(int, int, int);

}  // namespace tecno
