// ens_fp64.hip — FP64 ensemble stage kernels, the ensemble control kernels
// and the dispatch (built with FP contraction off).
#include "ens_kernels.hpp"

namespace tecno {

template hipError_t ens_launch<ScalarPh<double, 2>>(hipStream_t, int, int, const void*,
                                                     const EnsWork*, int, int, int);
template int ens_grid_of<ScalarPh<double, 2>>(int, int, int);
template hipError_t ens_launch<ScalarPh<double, 2, 1>>(hipStream_t, int, int, const void*,
                                                        const EnsWork*, int, int, int);
template int ens_grid_of<ScalarPh<double, 2, 1>>(int, int, int);
template hipError_t ens_launch<ScalarMixedPh<2>>(hipStream_t, int, int, const void*,
                                                 const EnsWork*, int, int, int);
template int ens_grid_of<ScalarMixedPh<2>>(int, int, int);
// the FP32 kernels are built in ens_fp32.hip
extern template hipError_t ens_launch<ScalarPh<float, 2>>(hipStream_t, int, int, const void*,
                                                          const EnsWork*, int, int, int);
extern template int ens_grid_of<ScalarPh<float, 2>>(int, int, int);
extern template hipError_t ens_launch<ScalarPh<float, 2, 1>>(hipStream_t, int, int, const void*,
                                                             const EnsWork*, int, int, int);
extern template int ens_grid_of<ScalarPh<float, 2, 1>>(int, int, int);

__global__ void k_ens_control(const EnsCtl* cs, int n, int finish) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const EnsCtl e = cs[i];
  Ctl* c = e.ctl;
  if (finish) {  // k_finish
    if (!c->active) return;
    c->time += c->dt;
    c->iter += 1;
    if (c->time == c->t_save) c->t_save += c->t_int;
    return;
  }
  // k_dt: FV::compute_dt (enflo_scalar/src_scalar/fv.cc:286-311) and the
  // output-stamp clamp
  if (!(c->time < c->final_time && c->iter < c->max_iter)) {
    c->active = 0;
    c->dt = 0.0;
    return;
  }
  c->active = 1;
  const double smax = e.const_speed ? e.speed : __longlong_as_double((long long)c->smax_bits);
  double dt = 1.0e20;
  const double q = e.hmin / smax;
  dt = (q < dt) ? q : dt;
  dt *= e.cfl;
  if (c->time + dt > c->t_save) dt = c->t_save - c->time;
  c->dt = dt;
  c->smax_bits = 0ull;
}

__global__ void k_ens_count(const EnsCtl* cs, int n, int* out) {
  // one wave: count problems that will still advance
  int cnt = 0;
  for (int i = threadIdx.x; i < n; i += 64) {
    const Ctl* c = cs[i].ctl;
    cnt += (c->time < c->final_time && c->iter < c->max_iter) ? 1 : 0;
  }
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
  if (threadIdx.x == 0) *out = cnt;
}

hipError_t ens_control(hipStream_t s, const EnsCtl* c, int n, int finish) {
  hipLaunchKernelGGL(k_ens_control, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s, c, n, finish);
  return hipGetLastError();
}

hipError_t ens_count_active(hipStream_t s, const EnsCtl* c, int n, int* out) {
  hipLaunchKernelGGL(k_ens_count, dim3(1), dim3(64), 0, s, c, n, out);
  return hipGetLastError();
}

hipError_t ens_stage(hipStream_t s, int prec, int dims, int variant, int spw, const void* probs,
                     const EnsWork* work, int nwork, int stage, int grid) {
  if (nwork <= 0) return hipSuccess;
  switch (prec) {
    case ENS_FP32:
      return spw ? ens_launch<ScalarPh<float, 2, 1>>(s, dims, variant, probs, work, nwork, stage, grid)
                 : ens_launch<ScalarPh<float, 2>>(s, dims, variant, probs, work, nwork, stage, grid);
    case ENS_FP32_STATE64:
      return ens_launch<ScalarMixedPh<2>>(s, dims, variant, probs, work, nwork, stage, grid);
    default:
      return spw ? ens_launch<ScalarPh<double, 2, 1>>(s, dims, variant, probs, work, nwork, stage, grid)
                 : ens_launch<ScalarPh<double, 2>>(s, dims, variant, probs, work, nwork, stage, grid);
  }
}

int ens_grid(int prec, int dims, int variant, int spw, int n_cu) {
  switch (prec) {
    case ENS_FP32:
      return spw ? ens_grid_of<ScalarPh<float, 2, 1>>(dims, variant, n_cu)
                 : ens_grid_of<ScalarPh<float, 2>>(dims, variant, n_cu);
    case ENS_FP32_STATE64: return ens_grid_of<ScalarMixedPh<2>>(dims, variant, n_cu);
    default:
      return spw ? ens_grid_of<ScalarPh<double, 2, 1>>(dims, variant, n_cu)
                 : ens_grid_of<ScalarPh<double, 2>>(dims, variant, n_cu);
  }
}

}  // namespace tecno
