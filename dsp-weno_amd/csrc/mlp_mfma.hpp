// mlp_mfma.hpp — the DSP-WENO 5-5-5-5-5 network on CDNA4 matrix cores.
//
// One v_mfma_f32_32x32x16_{bf16,f16} per layer evaluates the network for the
// 64 lanes of a wave at once.  Two stencils share each column of the 32x32
// product, placed block-diagonally so that no lane exchanges data:
//   lane n      (n < 32): inputs at k = 0..4,   outputs at rows 0,1,2,3,8
//   lane n + 32          : inputs at k = 8..12,  outputs at rows 4,5,6,7,12
// With the CDNA 32x32x16 operand layout (lane l holds B[8*(l/32)+i][l%32] and
// D rows 8*(s/4) + 4*(l/32) + s%4 at accumulator slot s), every lane finds its
// own stencil's outputs in accumulator slots 0..4, which are directly the
// next layer's inputs.  The bias rides in the weights: each lane's input
// block carries a constant 1 at k = 5 (k = 13) and the matching weight is the
// bias, so the accumulator input C is the inline constant 0 and no register
// block has to be initialised per product.  A (the weights) is packed on the
// host per lane.
#pragma once
#ifndef TECNO_HOST_BUILD
#include <hip/hip_runtime.h>

#include <type_traits>

namespace tecno {
namespace mfma {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <bool H>
struct Net5 {
  using E = std::conditional_t<H, _Float16, __bf16>;
  using V = std::conditional_t<H, f16x8, bf16x8>;
  V A[4];

  // packed: [layer][lane 0..63][8] weights of this lane's row block, bias at 5
  __device__ void load(const unsigned short* packed, int lane) {
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const u16x8 w = *(const u16x8*)(packed + ((size_t)l * 64 + lane) * 8);
      A[l] = __builtin_bit_cast(V, w);
    }
  }
  __device__ static f32x16 mma(V a, V x, f32x16 c) {
    if constexpr (H)
      return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, x, c, 0, 0, 0);
    else
      return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, x, c, 0, 0, 0);
  }
  // two floats -> one dword of two 16-bit values (one packed conversion)
  __device__ static unsigned cvt2(float a, float b) {
    typedef float f32x2 __attribute__((ext_vector_type(2)));
    typedef E e16x2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, e16x2));
  }
  // network input: five features, the constant-1 bias input at 5, zeros
  // (one packed conversion per two values)
  __device__ static V pack(float x0, float x1, float x2, float x3, float x4) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 w = {cvt2(x0, x1), cvt2(x2, x3), cvt2(x4, 1.0f), 0u};
    return __builtin_bit_cast(V, w);
  }
  // ReLU after the conversion: a bf16 / fp16 value is negative exactly when
  // its bit pattern is a negative 16-bit integer, so max(bits, 0) as packed
  // int16 (two per instruction) clears the negative activations, -0 included
  __device__ static unsigned relu2(unsigned w) {
    typedef short s16x2 __attribute__((ext_vector_type(2)));
    const s16x2 b = __builtin_elementwise_max(__builtin_bit_cast(s16x2, w), (s16x2){0, 0});
    return __builtin_bit_cast(unsigned, b);
  }
  // next layer's input: ReLU of five outputs, the constant-1 bias input at 5
  __device__ static V pack_relu(float x0, float x1, float x2, float x3, float x4) {
    typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 w = {relu2(cvt2(x0, x1)), relu2(cvt2(x2, x3)), relu2(cvt2(x4, 0.0f)) |
                                                                   (cvt2(1.0f, 0.0f) << 16),
                     0u};
    return __builtin_bit_cast(V, w);
  }
  // all 64 lanes of the wave must call this together
  __device__ void eval(const float in[5], float out[5]) const {
    V x = pack(in[0], in[1], in[2], in[3], in[4]);
    const f32x16 zero = {};
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const f32x16 d = mma(A[l], x, zero);
      if (l < 3) {
        x = pack_relu(d[0], d[1], d[2], d[3], d[4]);
      } else {
#pragma unroll
        for (int j = 0; j < 5; ++j) out[j] = d[j];
      }
    }
  }
};

}  // namespace mfma
}  // namespace tecno
#endif
