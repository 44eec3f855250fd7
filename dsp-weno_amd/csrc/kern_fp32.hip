// FP32 instantiation of the TeCNO kernels (FMA contraction allowed).
#include "fv_launch.hpp"

namespace tecno {
template struct Launch<float>;
}
