// ens_kernels.hpp — the persistent ensemble stage kernel and its launchers,
// instantiated once per state precision (ens_fp64.hip with FP contraction
// off, as the FP64 single-problem kernels; ens_fp32.hip with contraction on,
// as the FP32 ones), so that ensemble members are bitwise single runs.
#pragma once
#include "ensemble.hpp"
#include "fv_mixed.hpp"

namespace tecno {

template <class PH, int DIMS, int TX, int TY, int G>
__global__ __launch_bounds__(256) void ens_stage_kernel(const EnsProb<typename PH::TP>* probs,
                                                        const EnsWork* work, int nwork,
                                                        int stage) {
  using TL = Tile<PH, DIMS, TX, TY, G>;
  using TS = typename PH::TS;
  __shared__ __attribute__((aligned(16))) unsigned char lds[TL::LDS_BYTES];
  __shared__ double red[4];
  const int tid = threadIdx.x;
  // SSPRK3 (ext_constants.h:7-8): a = {0, 3/4, 1/3}, b = {1, 1/4, 2/3}
  const double a = stage == 0 ? 0.0 : stage == 1 ? 3.0 / 4.0 : 1.0 / 3.0;
  const double b = stage == 0 ? 1.0 : stage == 1 ? 1.0 / 4.0 : 2.0 / 3.0;
  const int last = stage == 2;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const EnsWork it = work[w];
    const EnsProb<typename PH::TP>& pr = probs[it.prob];
    Ctl* ctl = pr.ctl;
    if (!ctl->active) continue;  // the whole block agrees: same work item
    TS* U0 = (TS*)pr.U[0];
    TS* U1 = (TS*)pr.U[1];
    TS* U2 = (TS*)pr.U[2];
    StageArgs<TS> A{stage == 0 ? U0 : stage == 1 ? U1 : U2, U0,
                    stage == 0 ? U1 : stage == 1 ? U2 : U0, ctl, a, b, b, last, nullptr, 0.0, 0};
    TL::load(A, pr.P, pr.g, lds, tid, it.bx, it.by);
    __syncthreads();
    typename PH::TF Fr[TL::ROUNDS][TL::NV];
    unsigned nfb = 0;
    TL::faces(pr.P, lds, tid, Fr, nfb);
    if (nfb) atomicAdd(&ctl->nfallback, (unsigned long long)nfb);
    __syncthreads();
    TL::store(lds, tid, Fr);
    __syncthreads();
    TS smax = TS(0);
    unsigned nbad = 0;
    TL::update(A, pr.P, pr.g, lds, tid, it.bx, it.by, ctl->dt, smax, nbad);
    if (last) {
      double s = (double)smax;
      for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off));
      if ((tid & 63) == 0) red[tid >> 6] = s;
      __syncthreads();
      if (tid == 0) {
        double m = red[0];
        for (int q = 1; q < 4; ++q) m = fmax(m, red[q]);
        atomic_max_pos(&ctl->smax_bits, m);
      }
      if (nbad) atomicAdd(&ctl->nbad, (unsigned long long)nbad);
    }
    __syncthreads();  // LDS and red are reused by the next work item
  }
}

template <class PH, int DIMS, int V>
hipError_t ens_launch_v(hipStream_t s, const void* probs, const EnsWork* work, int nwork, int stage,
                        int grid) {
  constexpr int TX = ens_tx(DIMS, V), TY = ens_ty(DIMS, V);
  hipLaunchKernelGGL((ens_stage_kernel<PH, DIMS, TX, TY, 2>), dim3(grid), dim3(256), 0, s,
                     (const EnsProb<typename PH::TP>*)probs, work, nwork, stage);
  return hipGetLastError();
}

template <class PH, int DIMS, int V>
int ens_grid_v(int n_cu) {
  int nb = 0;
  constexpr int TX = ens_tx(DIMS, V), TY = ens_ty(DIMS, V);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, ens_stage_kernel<PH, DIMS, TX, TY, 2>, 256,
                                                   0) != hipSuccess ||
      nb < 1)
    nb = 1;
  return nb * n_cu;
}

template <class PH>
hipError_t ens_launch(hipStream_t s, int dims, int v, const void* probs, const EnsWork* work,
                      int nwork, int stage, int grid) {
  if (dims == 1) {
    switch (v) {
      case 0: return ens_launch_v<PH, 1, 0>(s, probs, work, nwork, stage, grid);
      case 1: return ens_launch_v<PH, 1, 1>(s, probs, work, nwork, stage, grid);
      case 2: return ens_launch_v<PH, 1, 2>(s, probs, work, nwork, stage, grid);
    }
    return hipErrorInvalidValue;
  }
  switch (v) {
    case 0: return ens_launch_v<PH, 2, 0>(s, probs, work, nwork, stage, grid);
    case 1: return ens_launch_v<PH, 2, 1>(s, probs, work, nwork, stage, grid);
    case 2: return ens_launch_v<PH, 2, 2>(s, probs, work, nwork, stage, grid);
    case 3: return ens_launch_v<PH, 2, 3>(s, probs, work, nwork, stage, grid);
    case 4: return ens_launch_v<PH, 2, 4>(s, probs, work, nwork, stage, grid);
  }
  return hipErrorInvalidValue;
}

template <class PH>
int ens_grid_of(int dims, int v, int n_cu) {
  if (dims == 1) {
    switch (v) {
      case 0: return ens_grid_v<PH, 1, 0>(n_cu);
      case 1: return ens_grid_v<PH, 1, 1>(n_cu);
      default: return ens_grid_v<PH, 1, 2>(n_cu);
    }
  }
  switch (v) {
    case 0: return ens_grid_v<PH, 2, 0>(n_cu);
    case 1: return ens_grid_v<PH, 2, 1>(n_cu);
    case 2: return ens_grid_v<PH, 2, 2>(n_cu);
    case 3: return ens_grid_v<PH, 2, 3>(n_cu);
    default: return ens_grid_v<PH, 2, 4>(n_cu);
  }
}

}  // namespace tecno
