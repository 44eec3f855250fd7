// FP32 instantiation: scalar FP32 kernels and the mixed-precision 2D Euler
// path (FP64 state, FP32 flux and reconstruction).
#include "fv_launch.hpp"

namespace tecno {
template struct Launch<float>;
template struct StateOps<float>;
}
