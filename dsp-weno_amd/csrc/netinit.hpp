// netinit.hpp — the seeded initial DSP-WENO network and the row order of
// each training epoch, drawn as the trainer draws them (host code).
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
  uint64_t random64() {
    const uint64_t hi = mt(), lo = mt();
    return (hi << 32) | lo;
  }
  double uniform(double from, double to) {
    const double x = (double)(random64() & ((1ull << 53) - 1)) * (1.0 / 9007199254740992.0);
    return std::fma(x, to - from, from);
  }
  // torch.empty((), dtype=torch.int64).random_(): 63 bits of one random64
  uint64_t random_int64() { return random64() & ((1ull << 63) - 1); }
};

// Rows a torch.randperm under a generator seeded with `seed` can take with
// the 32-bit draws below (PyTorch uses them for n < 2^32 / 20).
constexpr int64_t TORCH_RANDPERM_MAX = 214748364;

// torch.randperm(n, generator) with generator.manual_seed(seed), for
// n <= TORCH_RANDPERM_MAX: order[i] <-> order[i + r % (n - i)] for
// i = 0 .. n-2, r a 32-bit draw
inline void randperm_torch(uint64_t seed, int64_t n, std::vector<int>& order) {
  order.resize((size_t)n);
  for (int64_t i = 0; i < n; ++i) order[(size_t)i] = (int)i;
  TorchStream g(seed);
  for (int64_t i = 0; i + 1 < n; ++i) {
    const int64_t z = (int64_t)((uint64_t)g.mt() % (uint64_t)(n - i));
    std::swap(order[(size_t)i], order[(size_t)(i + z)]);
  }
}

// widths[0..nl]; W row-major out x in, layers concatenated; b concatenated;
// g: the generator after torch.manual_seed(seed), left as the trainer leaves
// it after building its network
inline void init_net_torch(int nl, const int32_t* widths, TorchStream& g, std::vector<double>& W,
                           std::vector<double>& b) {
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

inline void init_net_torch(int nl, const int32_t* widths, uint64_t seed, std::vector<double>& W,
                           std::vector<double>& b) {
  TorchStream g(seed);
  init_net_torch(nl, widths, g, W, b);
}

// The seed each epoch's shuffle takes in the trainer (trainer:348, one
// DataLoader(shuffle=True) iterator per epoch, :384-392) started with
// seed_no = seed and this network: every new iterator draws its base seed and
// then its RandomSampler's seed, each a random_() int64 from the global
// generator, and the sampler permutes with torch.randperm under a generator
// seeded with the second.
inline void trainer_epoch_seeds(int nl, const int32_t* widths, uint64_t seed, int n_epochs,
                                uint64_t* out) {
  TorchStream g(seed);
  std::vector<double> W, b;
  init_net_torch(nl, widths, g, W, b);
  for (int e = 0; e < n_epochs; ++e) {
    (void)g.random_int64();  // the iterator's base seed
    out[e] = g.random_int64();
  }
}

}  // namespace tecno
