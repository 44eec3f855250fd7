// host_shim.hpp — lets the device numerics (fv_device.hpp, fv_physics.hpp)
// compile as plain host C++ for CPU checks of the kernel arithmetic.
#pragma once
#include <cmath>
#include <cstdint>
#define __device__
#define __host__
#define __forceinline__ inline
using std::fabs;
using std::fma;
using std::log;
using std::exp;
using std::tanh;
using std::sqrt;
inline float logf(float x) { return std::log(x); }
inline float expf(float x) { return std::exp(x); }
inline float tanhf(float x) { return std::tanh(x); }
inline float sqrtf(float x) { return std::sqrt(x); }
inline float fabsf(float x) { return std::fabs(x); }
