// mlp.hip — DSP-WENO network kernels outside the stage kernel: batched
// matrix-core inference, and the trainer (forward, backward, Adam).
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <random>
#include "fv_mixed.hpp"
#include "mlp_mfma.hpp"
#include "netinit.hpp"

namespace tecno {

// Neural_Network::forward on n raw stencils (no zero-jump skip), features and
// head in FP32 from FP64 cell differences, the network on the matrix cores.
// Every wave runs whole; lanes past n repeat the last stencil.
template <bool H>
__global__ __launch_bounds__(256) void k_mlp_mfma(const unsigned short* A, const double* u4,
                                                  int64_t n, double* lo, double* ro) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t q = i < n ? i : n - 1;
  mfma::Net5<H> net;
  net.load(A, threadIdx.x & 63);
  const double u0 = u4[4 * q], u1 = u4[4 * q + 1], u2 = u4[4 * q + 2], u3 = u4[4 * q + 3];
  const float d0 = (float)(u1 - u0), d1 = (float)(u2 - u1), d2 = (float)(u3 - u2);
  const float mx = (float)fmax(fmax(fabs(u0), fabs(u1)), fmax(fabs(u2), fabs(u3)));
  float st[5], xv[5], yv[5], lg[5], C1, C2;
  fast::features(d0, d1, d2, mx, st, xv, yv);
#pragma unroll
  for (int k = 0; k < 5; ++k) st[k] = (fabsf(st[k]) < INFINITY) ? st[k] : 0.0f;
  net.eval(st, lg);
  fast::head_c(lg, xv, yv, C1, C2);
  const double w1 = 0.25 - 2.0 * (double)C1, tw0 = 0.25 - 2.0 * (double)C2;
  if (i < n) {
    lo[i] = u1 + 0.5 * (1.0 - w1) * (u2 - u1) + 0.5 * w1 * (u1 - u0);
    ro[i] = u2 - 0.5 * (1.0 - tw0) * (u2 - u1) - 0.5 * tw0 * (u3 - u2);
  }
}

hipError_t mlp_mfma_batch(hipStream_t s, const unsigned short* A, bool fp16, const double* u4,
                          int64_t n, double* l, double* r) {
  if (!A) return hipErrorInvalidValue;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (fp16)
    hipLaunchKernelGGL((k_mlp_mfma<true>), dim3(nb), dim3(256), 0, s, A, u4, n, l, r);
  else
    hipLaunchKernelGGL((k_mlp_mfma<false>), dim3(nb), dim3(256), 0, s, A, u4, n, l, r);
  return hipGetLastError();
}

// ---------------------------------------------------------------- training
// Semantics of DSP-WENO_trainer/DSP-WENO_network_trainer.py: MLP.forward
// (63-91), inter_evaluation (94-106), loss (MSE(left) + MSE(right)) / 2
// (355-356), torch.optim.Adam with L2 weight decay (353).  One thread per
// row; every parameter's gradient is summed over the wave with a butterfly,
// over the block in LDS and over the blocks in FP64 by the update kernel, in
// a fixed order (deterministic).
namespace {
constexpr int PW = MlpTrainer::PW, PL = MlpTrainer::PL, NP = MlpTrainer::NP, NPL = NP + 1;
constexpr int ROW = 21;

__device__ __forceinline__ int wi(int l, int j, int k) { return l * PW * PW + j * PW + k; }
__device__ __forceinline__ int bi(int l, int j) { return PL * PW * PW + l * PW + j; }

__device__ __forceinline__ float bf16r(float x) {  // nearest-even bfloat16 of a finite x
  unsigned u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return __uint_as_float(u & 0xffff0000u);
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct Widths {
  int w[PL + 1];
};

template <int NL, bool BF>
__global__ __launch_bounds__(256) void k_train_fb(const float* prm, Widths W, const double* rows,
                                                  const int* idx, int64_t row0, int64_t nrows,
                                                  double inv_n, float* part) {
  __shared__ float sW[NP];
  __shared__ float sred[4][NPL];
  for (int q = threadIdx.x; q < NP; q += 256) sW[q] = BF ? bf16r(prm[q]) : prm[q];
  for (int q = threadIdx.x; q < 4 * NPL; q += 256) (&sred[0][0])[q] = 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool act = r < nrows;
  const int64_t rr = row0 + (act ? r : 0);
  const double* row = rows + (idx ? (int64_t)idx[rr] : rr) * ROW;

  // forward
  float h[NL + 1][PW];
#pragma unroll
  for (int k = 0; k < PW; ++k) h[0][k] = k < 5 ? (BF ? bf16r((float)row[k]) : (float)row[k]) : 0.0f;
#pragma unroll
  for (int l = 0; l < NL; ++l)
#pragma unroll
    for (int j = 0; j < PW; ++j) {
      float z = sW[bi(l, j)];
#pragma unroll
      for (int k = 0; k < PW; ++k) z = fmaf(sW[wi(l, j, k)], h[l][k], z);
      if (l < NL - 1) {
        z = fmaxf(z, 0.0f);
        if (BF) z = bf16r(z);
      }
      h[l + 1][j] = z;
    }
  float mx = h[NL][0];
#pragma unroll
  for (int i = 1; i < 5; ++i) mx = fmaxf(mx, h[NL][i]);
  float s[5], sum = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    s[i] = expf(h[NL][i] - mx);
    sum += s[i];
  }
  float C1 = 0.0f, C2 = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    s[i] /= sum;
    C1 = fmaf(s[i], (float)row[5 + i], C1);
    C2 = fmaf(s[i], (float)row[10 + i], C2);
  }
  const double u0 = row[15], u1 = row[16], u2 = row[17], u3 = row[18];
  const double w1 = 0.25 - 2.0 * (double)C1, tw0 = 0.25 - 2.0 * (double)C2;
  const double left = (1.0 - w1) * 0.5 * (u1 + u2) + w1 * (1.5 * u1 - 0.5 * u0);
  const double right = tw0 * (1.5 * u2 - 0.5 * u3) + (1.0 - tw0) * 0.5 * (u1 + u2);
  const double eL = act ? left - row[19] : 0.0, eR = act ? right - row[20] : 0.0;
  const float lrow = (float)(0.5 * (eL * eL + eR * eR));

  // backward
  const float gC1 = (float)(eL * inv_n * -2.0 * (u1 - 0.5 * u0 - 0.5 * u2));
  const float gC2 = (float)(eR * inv_n * -2.0 * (u2 - 0.5 * u3 - 0.5 * u1));
  float gz[PW];
#pragma unroll
  for (int i = 0; i < PW; ++i)
    gz[i] = i < 5 ? s[i] * (((float)row[5 + i] - C1) * gC1 + ((float)row[10 + i] - C2) * gC2)
                  : 0.0f;
#pragma unroll
  for (int l = NL - 1; l >= 0; --l) {
    const int wo = W.w[l + 1], wn = W.w[l];
#pragma unroll
    for (int j = 0; j < PW; ++j) {
      if (j >= wo) continue;  // wave-uniform
      const float gb = wave_sum(gz[j]);
      if (lane == 0) sred[wave][bi(l, j)] = gb;
#pragma unroll
      for (int k = 0; k < PW; ++k) {
        if (k >= wn) continue;
        const float g = wave_sum(gz[j] * h[l][k]);
        if (lane == 0) sred[wave][wi(l, j, k)] = g;
      }
    }
    if (l > 0) {
      float gh[PW];
#pragma unroll
      for (int k = 0; k < PW; ++k) {
        float a = 0.0f;
#pragma unroll
        for (int j = 0; j < PW; ++j) a = fmaf(sW[wi(l, j, k)], gz[j], a);
        gh[k] = a;
      }
#pragma unroll
      for (int k = 0; k < PW; ++k) gz[k] = h[l][k] > 0.0f ? gh[k] : 0.0f;
    }
  }
  const float ls = wave_sum(lrow);
  if (lane == 0) sred[wave][NP] = ls;
  __syncthreads();
  for (int q = threadIdx.x; q < NPL; q += 256)
    part[(size_t)blockIdx.x * NPL + q] = ((sred[0][q] + sred[1][q]) + sred[2][q]) + sred[3][q];
}

// sum of the block partials (fp64, fixed order), optional Adam step
__global__ __launch_bounds__(320) void k_train_update(const float* part, int nb, float* prm,
                                                      float* m, float* v, double* grad,
                                                      double* loss_acc, int update, float lr,
                                                      float b1, float b2, float eps, float wd,
                                                      float bc1, float bc2) {
  const int q = threadIdx.x;
  if (q >= NPL) return;
  double g = 0.0;
  for (int b = 0; b < nb; ++b) g += (double)part[(size_t)b * NPL + q];
  if (q == NP) {
    loss_acc[0] += g;
    loss_acc[1] += 1.0;
    return;
  }
  if (grad) grad[q] = g;
  if (!update) return;
  float p = prm[q];
  const float gg = (float)g + wd * p;
  const float mm = b1 * m[q] + (1.0f - b1) * gg;
  const float vv = b2 * v[q] + (1.0f - b2) * gg * gg;
  m[q] = mm;
  v[q] = vv;
  p -= (lr / bc1) * mm / (sqrtf(vv) / sqrtf(bc2) + eps);
  prm[q] = p;
}

__global__ void k_train_rows(const double* u4, const double* tgt, int64_t n, double* rows) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double st[5], xv[5], yv[5];
  nn_features(u4[4 * i], u4[4 * i + 1], u4[4 * i + 2], u4[4 * i + 3], st, xv, yv);
  double* r = rows + i * ROW;
  for (int k = 0; k < 5; ++k) {
    r[k] = st[k];
    r[5 + k] = xv[k];
    r[10 + k] = yv[k];
  }
  for (int k = 0; k < 4; ++k) r[15 + k] = u4[4 * i + k];
  r[19] = tgt[2 * i];
  r[20] = tgt[2 * i + 1];
}

template <int NL>
hipError_t launch_fb(hipStream_t s, bool bf, unsigned nb, const float* prm, Widths W,
                     const double* rows, const int* idx, int64_t row0, int64_t n, double inv_n,
                     float* part) {
  if (bf)
    hipLaunchKernelGGL((k_train_fb<NL, true>), dim3(nb), dim3(256), 0, s, prm, W, rows, idx, row0,
                       n, inv_n, part);
  else
    hipLaunchKernelGGL((k_train_fb<NL, false>), dim3(nb), dim3(256), 0, s, prm, W, rows, idx,
                       row0, n, inv_n, part);
  return hipGetLastError();
}
}  // namespace

#define TCHK(x)                                                      \
  do {                                                               \
    hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) {                                          \
      err = std::string(#x ": ") + hipGetErrorString(e_);            \
      return -1;                                                     \
    }                                                                \
  } while (0)

int MlpTrainer::init(hipStream_t s, int nl, const int* widths, const double* W, const double* b,
                     double lr, double beta1, double beta2, double eps, double wd,
                     std::string& err) {
  if (nl < 1 || nl > PL) {
    err = "training supports 1 to 4 layers";
    return -1;
  }
  if (widths[0] != 5 || widths[nl] != 5) {
    err = "the network maps 5 inputs to 5 vertex weights";
    return -1;
  }
  for (int l = 0; l <= nl; ++l)
    if (widths[l] < 1 || widths[l] > PW) {
      err = "training supports widths up to 8";
      return -1;
    }
  s_ = s;
  nl_ = nl;
  for (int l = 0; l <= nl; ++l) w_[l] = widths[l];
  lr_ = lr;
  b1_ = beta1;
  b2_ = beta2;
  eps_ = eps;
  wd_ = wd;
  t_ = 0;
  std::vector<float> p(NP, 0.0f);
  for (int l = 0; l < nl; ++l)
    for (int j = 0; j < widths[l + 1]; ++j) {
      for (int k = 0; k < widths[l]; ++k)
        p[l * PW * PW + j * PW + k] = (float)W[(size_t)l * 256 + j * 16 + k];
      p[PL * PW * PW + l * PW + j] = (float)b[(size_t)l * 16 + j];
    }
  if (!dparam_) {
    TCHK(hipMalloc(&dparam_, NP * sizeof(float)));
    TCHK(hipMalloc(&dm_, NP * sizeof(float)));
    TCHK(hipMalloc(&dv_, NP * sizeof(float)));
    TCHK(hipMalloc(&dgrad_, NP * sizeof(double)));
    TCHK(hipMalloc(&dloss_, 2 * sizeof(double)));
  }
  TCHK(hipMemcpyAsync(dparam_, p.data(), NP * sizeof(float), hipMemcpyHostToDevice, s_));
  TCHK(hipMemsetAsync(dm_, 0, NP * sizeof(float), s_));
  TCHK(hipMemsetAsync(dv_, 0, NP * sizeof(float), s_));
  TCHK(hipStreamSynchronize(s_));
  return 0;
}

int MlpTrainer::run_batch(const double* drows, const int* didx, int64_t row0, int64_t n, bool bf16,
                          bool update, double* dgrad, std::string& err) {
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (nb > part_cap_) {
    if (dpart_) TCHK(hipFree(dpart_));
    dpart_ = nullptr;
    TCHK(hipMalloc(&dpart_, (size_t)nb * NPL * sizeof(float)));
    part_cap_ = nb;
  }
  Widths W{};
  for (int l = 0; l <= nl_; ++l) W.w[l] = w_[l];
  const double inv_n = 1.0 / (double)n;
  hipError_t e;
  switch (nl_) {
    case 1: e = launch_fb<1>(s_, bf16, nb, dparam_, W, drows, didx, row0, n, inv_n, dpart_); break;
    case 2: e = launch_fb<2>(s_, bf16, nb, dparam_, W, drows, didx, row0, n, inv_n, dpart_); break;
    case 3: e = launch_fb<3>(s_, bf16, nb, dparam_, W, drows, didx, row0, n, inv_n, dpart_); break;
    default: e = launch_fb<4>(s_, bf16, nb, dparam_, W, drows, didx, row0, n, inv_n, dpart_); break;
  }
  TCHK(e);
  float bc1 = 1.0f, bc2 = 1.0f;
  if (update) {
    ++t_;
    bc1 = (float)(1.0 - std::pow(b1_, (double)t_));
    bc2 = (float)(1.0 - std::pow(b2_, (double)t_));
  }
  hipLaunchKernelGGL(k_train_update, dim3(1), dim3(320), 0, s_, dpart_, (int)nb, dparam_, dm_, dv_,
                     dgrad, dloss_, update ? 1 : 0, (float)lr_, (float)b1_, (float)b2_,
                     (float)eps_, (float)wd_, bc1, bc2);
  TCHK(hipGetLastError());
  return 0;
}

int MlpTrainer::step(const double* rows, int64_t n, bool bf16, bool update, double* loss,
                     std::vector<double>* grad, std::string& err) {
  if (!dparam_) {
    err = "trainer not initialised";
    return -1;
  }
  if ((size_t)n > rows_cap_) {
    if (drows_) TCHK(hipFree(drows_));
    drows_ = nullptr;
    TCHK(hipMalloc(&drows_, (size_t)n * ROW * sizeof(double)));
    rows_cap_ = (size_t)n;
  }
  TCHK(hipMemcpyAsync(drows_, rows, (size_t)n * ROW * sizeof(double), hipMemcpyHostToDevice, s_));
  TCHK(hipMemsetAsync(dloss_, 0, 2 * sizeof(double), s_));
  if (run_batch(drows_, nullptr, 0, n, bf16, update, dgrad_, err)) return -1;
  double lh[2];
  TCHK(hipMemcpyAsync(lh, dloss_, 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
  if (grad) {
    grad->assign(NP, 0.0);
    TCHK(hipMemcpyAsync(grad->data(), dgrad_, NP * sizeof(double), hipMemcpyDeviceToHost, s_));
  }
  TCHK(hipStreamSynchronize(s_));
  if (loss) *loss = lh[0] / (double)n;
  return 0;
}

int MlpTrainer::set_data(const double* rows, int64_t n, std::string& err) {
  if (n <= 0 || n > (int64_t)0x7fffffff) {
    err = "data set size must be in 1 .. 2^31-1 rows";
    return -1;
  }
  if (ddata_) TCHK(hipFree(ddata_));
  if (didx_) TCHK(hipFree(didx_));
  ddata_ = nullptr;
  didx_ = nullptr;
  TCHK(hipMalloc(&ddata_, (size_t)n * ROW * sizeof(double)));
  TCHK(hipMalloc(&didx_, (size_t)n * sizeof(int)));
  TCHK(hipMemcpyAsync(ddata_, rows, (size_t)n * ROW * sizeof(double), hipMemcpyHostToDevice, s_));
  TCHK(hipStreamSynchronize(s_));
  ndata_ = n;
  return 0;
}

int MlpTrainer::epoch(int64_t batch, uint64_t seed, bool bf16, double* mean_loss,
                      std::string& err) {
  if (!ddata_) {
    err = "no training data on the device";
    return -1;
  }
  if (batch <= 0) {
    err = "batch size must be positive";
    return -1;
  }
  if (ndata_ > TORCH_RANDPERM_MAX) {
    err = "more training rows than the epoch shuffle takes (214748364)";
    return -1;
  }
  // the row order of the trainer's DataLoader(shuffle=True) whose sampler
  // drew `seed`: torch.randperm under a generator seeded with it
  std::vector<int> order;
  randperm_torch(seed, ndata_, order);
  TCHK(hipMemcpyAsync(didx_, order.data(), (size_t)ndata_ * sizeof(int), hipMemcpyHostToDevice, s_));
  TCHK(hipMemsetAsync(dloss_, 0, 2 * sizeof(double), s_));
  for (int64_t r0 = 0; r0 < ndata_; r0 += batch) {
    const int64_t nb = std::min(batch, ndata_ - r0);
    // the per-batch loss partial is a sum over rows: scale to the batch mean
    if (run_batch(ddata_, didx_, r0, nb, bf16, true, nullptr, err)) return -1;
  }
  double lh[2];
  TCHK(hipMemcpyAsync(lh, dloss_, 2 * sizeof(double), hipMemcpyDeviceToHost, s_));
  TCHK(hipStreamSynchronize(s_));
  // loss_acc holds the sum of every row's loss; report the mean row loss
  if (mean_loss) *mean_loss = lh[0] / (double)ndata_;
  return 0;
}

void MlpTrainer::get_weights(std::vector<double>& W, std::vector<double>& b) {
  std::vector<float> p(NP);
  if (dparam_) {
    (void)hipMemcpyAsync(p.data(), dparam_, NP * sizeof(float), hipMemcpyDeviceToHost, s_);
    (void)hipStreamSynchronize(s_);
  }
  W.assign(PL * PW * PW, 0.0);
  b.assign(PL * PW, 0.0);
  for (int q = 0; q < PL * PW * PW; ++q) W[q] = p[q];
  for (int q = 0; q < PL * PW; ++q) b[q] = p[PL * PW * PW + q];
}

MlpTrainer::~MlpTrainer() {
  for (void* ptr : {(void*)dparam_, (void*)dm_, (void*)dv_, (void*)dpart_, (void*)dloss_,
                    (void*)dgrad_, (void*)drows_, (void*)ddata_, (void*)didx_})
    if (ptr) (void)hipFree(ptr);
}

int make_train_rows(hipStream_t s, const double* u4, const double* target, int64_t n,
                    double* rows, std::string& err) {
  double *du = nullptr, *dt = nullptr, *dr = nullptr;
  auto cleanup = [&]() {
    if (du) (void)hipFree(du);
    if (dt) (void)hipFree(dt);
    if (dr) (void)hipFree(dr);
  };
  hipError_t e = hipMalloc(&du, (size_t)n * 4 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dt, (size_t)n * 2 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&dr, (size_t)n * ROW * sizeof(double));
  if (e == hipSuccess) e = hipMemcpyAsync(du, u4, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(dt, target, (size_t)n * 2 * sizeof(double), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_train_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, du, dt, n, dr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(rows, dr, (size_t)n * ROW * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  cleanup();
  if (e != hipSuccess) {
    err = std::string("make_train_rows: ") + hipGetErrorString(e);
    return -1;
  }
  return 0;
}

}  // namespace tecno
