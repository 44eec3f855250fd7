// ensemble.hpp — many independent scalar problems of different sizes advanced
// together: one persistent launch per (precision, dimension) group and stage,
// blocks striding over a flat list of (problem, tile) work items.  Every
// problem keeps its own state buffers, dt, time and output stamps in its own
// control block, so problems with different meshes and end times share the
// launches; a problem that has reached its end time skips its tiles.  The
// per-tile work is the single-problem stage code (Tile<...>), so an ensemble
// member's state is bitwise the state of the same problem run alone.
#pragma once
#include <hip/hip_runtime.h>

#include "fv_kernels.hpp"

namespace tecno {

// per-problem step control (CFL dt and stamp bookkeeping), any precision
struct EnsCtl {
  Ctl* ctl;
  double hmin, cfl, speed;
  int const_speed, pad;
};

// per-problem stage inputs of a group of precision TP
template <typename TP>
struct EnsProb {
  Phys<TP> P;
  GridDev g;
  Ctl* ctl;
  void* U[3];  // U^n, U^(1), U^(2)
};

struct EnsWork {
  int prob, bx, by, pad;
};

// tile shapes the autotuner chooses from (every shape fills whole 256-lane
// face rounds to >= 88%)
constexpr int ENS_NV2 = 5, ENS_NV1 = 3;
__host__ __device__ constexpr int ens_tx(int dims, int v) {
  return dims == 1 ? (v == 0 ? 255 : v == 1 ? 511 : 1023)
                   : (v == 0 ? 32 : v == 1 ? 64 : v == 2 ? 16 : v == 3 ? 24 : 48);
}
__host__ __device__ constexpr int ens_ty(int dims, int v) {
  return dims == 1 ? 1 : (v == 0 ? 15 : v == 1 ? 7 : v == 2 ? 30 : v == 3 ? 20 : 10);
}

// dt of every active problem (k_dt per problem) / end of step (k_finish)
hipError_t ens_control(hipStream_t s, const EnsCtl* c, int n, int finish);
// one SSPRK3 stage (0, 1, 2) of every problem of a group of precision
// kind prec (ENS_FP64, ENS_FP32, ENS_FP32_STATE64); spw: every member
// reconstructs with the SP-WENO family (the kernel with that
// reconstruction only)
enum { ENS_FP64 = 0, ENS_FP32 = 1, ENS_FP32_STATE64 = 2 };
hipError_t ens_stage(hipStream_t s, int prec, int dims, int variant, int spw, const void* probs,
                     const EnsWork* work, int nwork, int stage, int grid);
// persistent grid size for a group's kernel: resident blocks on the device
int ens_grid(int prec, int dims, int variant, int spw, int n_cu);
// number of problems still advancing (written to *out on the device)
hipError_t ens_count_active(hipStream_t s, const EnsCtl* c, int n, int* out);

// per-physics instantiations (ens_fp64.hip: ScalarPh<double> and
// ScalarMixedPh; ens_fp32.hip: ScalarPh<float>)
template <class PH>
hipError_t ens_launch(hipStream_t s, int dims, int v, const void* probs, const EnsWork* work,
                      int nwork, int stage, int grid);
template <class PH>
int ens_grid_of(int dims, int v, int n_cu);

}  // namespace tecno
