// netinit.hpp — the seeded initial DSP-WENO network, drawn as the trainer
// draws it (host code).
//
// The trainer seeds PyTorch's CPU generator (torch.manual_seed(seed_no),
// DSP-WENO_network_trainer.py:296; float64 default dtype, :254), builds one
// nn.Linear per layer (:54-56) and then applies nn.init.xavier_uniform_ to
// every weight (:58-60).  PyTorch's generator is MT19937 seeded with the low
// 32 bits of the seed; a float64 uniform_(from, to) takes two 32-bit outputs
// hi, lo per value, x = ((hi << 32 | lo) & (2^53 - 1)) * 2^-53 and returns
// fma(x, to - from, from) (the fused form is what PyTorch's CPU kernel
// computes on x86-64 hosts with FMA).  nn.Linear's own initialisation draws
// the weight (kaiming_uniform_; overwritten later, so only the number of
// draws matters) and then the bias U(+-1/sqrt(fan_in)); xavier_uniform_ draws
// U(+-a), a = sqrt(3) * sqrt(2 / (fan_in + fan_out)), in row-major
// (out x in) order.
#pragma once
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace tecno {

struct TorchStream {
  std::mt19937 mt;
  explicit TorchStream(uint64_t seed) : mt((uint32_t)(seed & 0xffffffffu)) {}
  double uniform(double from, double to) {
    const uint64_t hi = mt(), lo = mt();
    const double x = (double)(((hi << 32) | lo) & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
    return std::fma(x, to - from, from);
  }
};

// widths[0..nl]; W row-major out x in, layers concatenated; b concatenated
inline void init_net_torch(int nl, const int32_t* widths, uint64_t seed, std::vector<double>& W,
                           std::vector<double>& b) {
  TorchStream g(seed);
  W.clear();
  b.clear();
  for (int l = 0; l < nl; ++l) {  // nn.Linear(ni, no) in order
    const int ni = widths[l], no = widths[l + 1];
    for (int q = 0; q < ni * no; ++q) g.uniform(-1.0, 1.0);  // kaiming draws, overwritten
    const double bound = 1.0 / std::sqrt((double)ni);
    for (int q = 0; q < no; ++q) b.push_back(g.uniform(-bound, bound));
  }
  for (int l = 0; l < nl; ++l) {  // xavier_uniform_(weight), gain 1
    const int ni = widths[l], no = widths[l + 1];
    const double a = std::sqrt(3.0) * std::sqrt(2.0 / (double)(ni + no));
    for (int q = 0; q < ni * no; ++q) W.push_back(g.uniform(-a, a));
  }
}

}  // namespace tecno
