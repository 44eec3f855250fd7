// FP64 instantiation of the TeCNO kernels.  Built with -ffp-contract=off so
// that every product and sum is rounded as in the reference build.
#define TECNO_DEFINE_CONTROL
#include "fv_launch.hpp"

namespace tecno {
template struct Launch<double>;
template struct StateOps<double>;
}
