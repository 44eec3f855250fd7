// fv_mixed.hpp — the mixed-precision 2D Euler TeCNO face (the FP32 path).
//
// State in HBM, per-cell records and the entropy-conservative flux are FP64
// (the residual is a difference of neighbouring fluxes, so their rounding
// error enters it divided by h); the entropy dissipation -- eigensystem,
// reconstruction of the scaled entropy variables, DSP-WENO network -- is
// about 1e-5 of the flux in smooth flow and runs in FP32 with the hardware
// reciprocal / log2 / exp2 / sqrt.
// Reconstruction works on jumps: with D0 = u1-u0, D1 = u2-u1, D2 = u3-u2 the
// reconstructed jump of SP-WENO (limiter.h:193-252) and of DSP-WENO
// (neural_network.cc:16-77, 320-325) is, exactly in real arithmetic,
//     u+ - u- = 0.5 * ( w1 (D1 - D0) + w~0 (D1 - D2) ),
// w1 = 1/4 - 2 C1, w~0 = 1/4 - 2 C2, so only differences of the scaled
// entropy variables enter.  The differences are taken in FP64 (V is FP64 per
// cell) and projected with the FP32 eigenvectors, so smooth-flow jumps keep
// FP32 relative accuracy instead of losing it to cancellation.
#pragma once
#include "fv_physics.hpp"
#include "mlp_mfma.hpp"

namespace tecno {
namespace fast {

#ifdef TECNO_HOST_BUILD
__device__ __forceinline__ float rcp(float x) { return 1.0f / x; }
__device__ __forceinline__ float log(float x) { return std::log(x); }
__device__ __forceinline__ float exp(float x) { return std::exp(x); }
__device__ __forceinline__ float exp2(float x) { return std::exp2(x); }
__device__ __forceinline__ float sqrt(float x) { return std::sqrt(x); }
#else
__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float log(float x) {
  return __builtin_amdgcn_logf(x) * 0.69314718055994531f;  // v_log_f32 is log2
}
__device__ __forceinline__ float exp(float x) {
  return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
}
__device__ __forceinline__ float exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
#endif
__device__ __forceinline__ float div(float a, float b) { return a * rcp(b); }
// a / b from the one-ulp reciprocal plus one residual correction (Markstein):
// exact whenever the quotient is representable, so cell-jump ratios such as
// theta = Delta0 / Delta1 come out exactly 1 (or -1) for equal (opposite)
// jumps, as the FP64 region tests of the limiters expect.  b == 0 keeps the
// uncorrected a * inf.
__device__ __forceinline__ float quot(float a, float b) {
  const float r = rcp(b);
  const float q = a * r;
  const float qc = fmaf(fmaf(-q, b, a), r, q);
  return (fabsf(qc) <= 3.402823466e38f) ? qc : q;
}
__device__ __forceinline__ float tanh(float x) {
  // 1 - 2/(e^{2x}+1): saturates to +-1 for large |x|; absolute error ~1e-7
  return 1.0f - 2.0f * rcp(exp(2.0f * x) + 1.0f);
}

// logarithmic mean (misc_func.h:14-32); both branches are well conditioned in
// FP32 for positive arguments
__device__ __forceinline__ float logavg(float a, float b) {
  const float xi = b * rcp(a);
  const float f = (xi - 1.0f) * rcp(xi + 1.0f);
  const float u = f * f;
  float F;
  if (u < 1.0e-2f)
    F = 1.0f + u * (1.0f / 3.0f) + (u * u) * (1.0f / 5.0f) + (u * u * u) * (1.0f / 7.0f);
  else
    F = log(xi) * rcp(2.0f * f);
  return 0.5f * (a + b) * rcp(F);
}

// ---------------------------------------------------------------- SP-WENO
// Reconstructed jump of SP-WENO from the cell jumps (u1, u2 only for
// sp_wenoc): sp_weno_rec (limiter.h:193-252) with C1 = Cp(t+, t-) and
// C2 = Cp(t-, t+) (limiter.h:152-190) from one ratio psi = (1 - t-)/(1 - t+):
// Cp's psi for C2 is 1/psi, so both take Cp's first branch together
// (psi < 0), where Cp's a / (a^2 + b^2) / 8 with a = fp(t0, t1) =
// 1 / (1 + psi) and b = fp(t1, t0) = 1 / (1 + 1/psi) (limiter.h:103) is
// (1 + psi) / (1 + psi^2) / 8 for C1 and psi times that for C2.  Otherwise
// (psi >= 0, or a ratio equal to 1) Cp is -3/8 for |t0| < 1 + eps and 1/8
// above; in FP32 1 + 1e-13 rounds to 1, so the floats below 1 + 1e-13 are
// those <= 1 (|t0| = 1 exactly: equal or alternating jumps).
__device__ __forceinline__ float spw_jump(float d0, float d1, float d2, float u1, float u2,
                                          int corr) {
  const float eps = 1.0e-13f;
  if (fabsf(d1) < eps) return d1;  // limiter.h:202-206: (u1, u2)
  const float tp = quot(d0, d1), tm = quot(d2, d1);
  const bool one = fabsf(tp - 1.0f) < eps || fabsf(tm - 1.0f) < eps;
  const float psi = one ? 0.0f : quot(1.0f - tm, 1.0f - tp);
  float C1, C2, c1c = 0.0f, c2c = 0.0f;
  if (psi < 0.0f) {
    if (fabsf(psi + 1.0f) < eps) {  // Cp's second branch (psi = -1)
      C1 = 0.0f;
      C2 = 0.0f;
    } else {
      C1 = 0.125f * (1.0f + psi) * rcp(fmaf(psi, psi, 1.0f));
      C2 = psi * C1;
      if (corr == 1) {
        const float scale = 0.5f * (fabsf(u1) + fabsf(u2));
        const float m = fminf(fabsf(d1 * rcp(scale)), fabsf(d1));
        const float m3 = -0.25f * m * m * m;
        c1c = m3 * rcp(1.0f - tp);
        c2c = m3 * rcp(1.0f - tm);
      }
    }
  } else {
    C1 = fabsf(tp) <= 1.0f ? -0.375f : 0.125f;
    C2 = fabsf(tm) <= 1.0f ? -0.375f : 0.125f;
  }
  if (corr == 1) {
    C1 = fminf(fmaxf(C1 + c1c, -0.375f), 0.125f);
    C2 = fminf(fmaxf(C2 + c2c, -0.375f), 0.125f);
  }
  return 0.5f * ((0.25f - 2.0f * C1) * (d1 - d0) + (0.25f - 2.0f * C2) * (d1 - d2));
}

// ---------------------------------------------------------------- DSP-WENO
// compute_vertices (neural_network.cc:135-260, vertex_mode 2, permutation 2)
__device__ __forceinline__ void vertices(float tm, float tp, float maxd, float xv[5], float yv[5]) {
  const float thr = 1e-8f;
  const float psi_p = quot(1.0f - tm, 1.0f - tp);
  const float psi_m = quot(1.0f, psi_p);
  const float g1 = fminf(maxd, 0.125f);
  const float g2 = -fminf(maxd, 0.375f);
  const float x1 = 0.125f * (1.0f + psi_p) - g1 * psi_p;
  const float y1 = 0.125f * (1.0f + psi_m) - g2 * psi_m;
  const float x2 = 0.125f * (1.0f + psi_p) - g2 * psi_p;
  const float y2 = 0.125f * (1.0f + psi_m) - g1 * psi_m;
  const float e = 0.125f, t = -0.375f, m = -0.125f;
  const float nx = ((psi_m * psi_m) + psi_m) * rcp(8.0f * ((psi_m * psi_m) + 1.0f));
  const float ny = 0.125f * (1.0f + psi_m) - nx * psi_m;
#define FV_SET(a0, a1, a2, a3, a4, b0, b1, b2, b3, b4) \
  { xv[0] = a0; xv[1] = a1; xv[2] = a2; xv[3] = a3; xv[4] = a4; \
    yv[0] = b0; yv[1] = b1; yv[2] = b2; yv[3] = b3; yv[4] = b4; }
  if ((tm - 1.0f) > thr && (tp - 1.0f) > thr) {
    FV_SET(e, e, e, e, e, e, e, e, e, e)
  } else if ((tm - 1.0f) > thr && (tp - 1.0f) < -thr) {
    if (psi_p < -1.0f) {
      if (x1 < g2) FV_SET(nx, nx, nx, nx, nx, ny, ny, ny, ny, ny)
      else {
        const float cx = (2.0f * g2 + x1) * (1.0f / 3.0f), cy = (2.0f * g1 + y2) * (1.0f / 3.0f);
        FV_SET(g2, x1, g2, cx, cx, g1, g1, y2, cy, cy)
      }
    } else {
      if (y2 < g2) FV_SET(g2, g1, g2, g1, 0.0f, g1, g1, g2, g2, 0.0f)
      else FV_SET(g2, g1, g2, g1, x2, g1, g1, g2, y2, g2)
    }
  } else if ((tm - 1.0f) < -thr && (tp - 1.0f) > thr) {
    if (psi_p < -1.0f) {
      if (x1 < g2) FV_SET(g2, g1, g2, g1, 0.0f, g1, g1, g2, g2, 0.0f)
      else FV_SET(g1, g1, g2, x1, g2, g2, g1, g2, g1, y1)
    } else {
      if (y2 < g2) FV_SET(nx, nx, nx, nx, nx, ny, ny, ny, ny, ny)
      else {
        const float cx = (2.0f * g1 + x1) * (1.0f / 3.0f), cy = (2.0f * g2 + y2) * (1.0f / 3.0f);
        FV_SET(g1, x1, g1, cx, cx, g2, g2, y2, cy, cy)
      }
    }
  } else if (fabsf(tm - 1.0f) < thr && (tp - 1.0f) > thr) {
    FV_SET(e, e, e, e, e, t, e, e, t, m)
  } else if (fabsf(tp - 1.0f) < thr && (tm - 1.0f) > thr) {
    FV_SET(e, e, t, t, m, e, e, e, e, e)
  } else {
    FV_SET(e, e, t, t, m, e, t, t, e, m)
  }
#undef FV_SET
}

// features of one component from its jumps and the stencil magnitude
__device__ __forceinline__ void features(float d0, float d1, float d2, float maxabs, float st[5],
                                         float xv[5], float yv[5]) {
  const float inv = rcp(fmaxf(1.0f, maxabs));
  const float tm = quot(d2, d1), tp = quot(d0, d1);
  const float a1 = fabsf(d0) * inv, a2 = fabsf(d1) * inv, a3 = fabsf(d2) * inv;
  st[0] = tanh(tm);
  st[1] = tanh(tp);
  st[2] = a1;
  st[3] = a2;
  st[4] = a3;
  vertices(tm, tp, fmaxf(fmaxf(a1, a2), a3), xv, yv);
}

// features without the vertex lists: the ratios and the largest scaled jump
// are returned for vertex_combine
__device__ __forceinline__ void features_nv(float d0, float d1, float d2, float maxabs,
                                            float st[5], float& tm, float& tp, float& maxd) {
  const float inv = rcp(fmaxf(1.0f, maxabs));
  tm = quot(d2, d1);
  tp = quot(d0, d1);
  const float a1 = fabsf(d0) * inv, a2 = fabsf(d1) * inv, a3 = fabsf(d2) * inv;
  st[0] = tanh(tm);
  st[1] = tanh(tp);
  st[2] = a1;
  st[3] = a2;
  st[4] = a3;
  maxd = fmaxf(fmaxf(a1, a2), a3);
}

// (C1, C2) = sum_i s_i (xv_i, yv_i) / sum_i s_i for the vertex lists of
// compute_vertices (neural_network.cc:135-260, vertex_mode 2, permutation
// 2; the same region tree as `vertices`), with the weighted sum taken inside
// each region: every list uses at most three distinct values per
// coordinate, so a region costs a few products instead of ten vertex
// assignments, and no vertex list is live across the network.
__device__ __forceinline__ void vertex_combine(float tm, float tp, float maxd, const float s[5],
                                               float& C1, float& C2) {
  const float thr = 1e-8f;
  const float S = (s[0] + s[1]) + (s[2] + s[3]) + s[4];
  const float e = 0.125f, t = -0.375f, m = -0.125f;
  float c1, c2;
  if ((tm - 1.0f) > thr && (tp - 1.0f) > thr) {
    c1 = e * S;
    c2 = e * S;
  } else if (((tm - 1.0f) > thr && (tp - 1.0f) < -thr) ||
             ((tm - 1.0f) < -thr && (tp - 1.0f) > thr)) {
    const bool r2 = (tm - 1.0f) > thr;  // theta- > 1, theta+ < 1 (else the mirror case)
    const float psi_p = quot(1.0f - tm, 1.0f - tp);
    const float psi_m = quot(1.0f, psi_p);
    const float g1 = fminf(maxd, 0.125f);
    const float g2 = -fminf(maxd, 0.375f);
    const float x1 = 0.125f * (1.0f + psi_p) - g1 * psi_p;
    const float y2 = 0.125f * (1.0f + psi_m) - g1 * psi_m;
    if (psi_p < -1.0f) {
      if (r2 && x1 < g2) {  // all vertices (nx, ny)
        const float nx = ((psi_m * psi_m) + psi_m) * rcp(8.0f * ((psi_m * psi_m) + 1.0f));
        c1 = nx * S;
        c2 = (0.125f * (1.0f + psi_m) - nx * psi_m) * S;
      } else if (!r2 && x1 < g2) {  // (g2,g1,g2,g1,0), (g1,g1,g2,g2,0)
        c1 = g2 * (s[0] + s[2]) + g1 * (s[1] + s[3]);
        c2 = g1 * (s[0] + s[1]) + g2 * (s[2] + s[3]);
      } else if (r2) {  // (g2,x1,g2,cx,cx), (g1,g1,y2,cy,cy)
        const float cx = (2.0f * g2 + x1) * (1.0f / 3.0f), cy = (2.0f * g1 + y2) * (1.0f / 3.0f);
        c1 = g2 * (s[0] + s[2]) + x1 * s[1] + cx * (s[3] + s[4]);
        c2 = g1 * (s[0] + s[1]) + y2 * s[2] + cy * (s[3] + s[4]);
      } else {  // (g1,g1,g2,x1,g2), (g2,g1,g2,g1,y1)
        const float y1 = 0.125f * (1.0f + psi_m) - g2 * psi_m;
        c1 = g1 * (s[0] + s[1]) + g2 * (s[2] + s[4]) + x1 * s[3];
        c2 = g2 * (s[0] + s[2]) + g1 * (s[1] + s[3]) + y1 * s[4];
      }
    } else {
      if (r2 && y2 < g2) {  // (g2,g1,g2,g1,0), (g1,g1,g2,g2,0)
        c1 = g2 * (s[0] + s[2]) + g1 * (s[1] + s[3]);
        c2 = g1 * (s[0] + s[1]) + g2 * (s[2] + s[3]);
      } else if (!r2 && y2 < g2) {  // all vertices (nx, ny)
        const float nx = ((psi_m * psi_m) + psi_m) * rcp(8.0f * ((psi_m * psi_m) + 1.0f));
        c1 = nx * S;
        c2 = (0.125f * (1.0f + psi_m) - nx * psi_m) * S;
      } else if (r2) {  // (g2,g1,g2,g1,x2), (g1,g1,g2,y2,g2)
        const float x2 = 0.125f * (1.0f + psi_p) - g2 * psi_p;
        c1 = g2 * (s[0] + s[2]) + g1 * (s[1] + s[3]) + x2 * s[4];
        c2 = g1 * (s[0] + s[1]) + g2 * (s[2] + s[4]) + y2 * s[3];
      } else {  // (g1,x1,g1,cx,cx), (g2,g2,y2,cy,cy)
        const float cx = (2.0f * g1 + x1) * (1.0f / 3.0f), cy = (2.0f * g2 + y2) * (1.0f / 3.0f);
        c1 = g1 * (s[0] + s[2]) + x1 * s[1] + cx * (s[3] + s[4]);
        c2 = g2 * (s[0] + s[1]) + y2 * s[2] + cy * (s[3] + s[4]);
      }
    }
  } else if (fabsf(tm - 1.0f) < thr && (tp - 1.0f) > thr) {  // (e,e,e,e,e), (t,e,e,t,m)
    c1 = e * S;
    c2 = t * (s[0] + s[3]) + e * (s[1] + s[2]) + m * s[4];
  } else if (fabsf(tp - 1.0f) < thr && (tm - 1.0f) > thr) {  // (e,e,t,t,m), (e,e,e,e,e)
    c1 = e * (s[0] + s[1]) + t * (s[2] + s[3]) + m * s[4];
    c2 = e * S;
  } else {  // (e,e,t,t,m), (e,t,t,e,m)
    c1 = e * (s[0] + s[1]) + t * (s[2] + s[3]) + m * s[4];
    c2 = e * (s[0] + s[3]) + t * (s[1] + s[2]) + m * s[4];
  }
  const float inv = rcp(S);
  C1 = c1 * inv;
  C2 = c2 * inv;
}

// softmax (zero-initialised maximum, neural_network.cc:276-295) of the
// logits, then the vertex combination -> reconstructed jump
__device__ __forceinline__ float head_combine(const float lg[5], float tm, float tp, float maxd,
                                              float d0, float d1, float d2) {
  float mx = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) mx = fmaxf(mx, lg[i]);
  // exp(l - mx) as 2^(l log2 e - mx log2 e): one fma per weight
  constexpr float L2E = 1.44269504088896341f;
  const float mxl = mx * L2E;
  float s[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) s[i] = exp2(fmaf(lg[i], L2E, -mxl));
  float C1, C2;
  vertex_combine(tm, tp, maxd, s, C1, C2);
  return 0.5f * ((0.25f - 2.0f * C1) * (d1 - d0) + (0.25f - 2.0f * C2) * (d1 - d2));
}

// softmax with zero-initialised maximum + vertex weights -> reconstructed jump
__device__ __forceinline__ float head(const float lg[5], const float xv[5], const float yv[5],
                                      float d0, float d1, float d2) {
  float mx = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) mx = fmaxf(mx, lg[i]);
  float s[5], sum = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    s[i] = exp(lg[i] - mx);
    sum += s[i];
  }
  const float inv = rcp(sum);
  float C1 = 0.0f, C2 = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    C1 += s[i] * xv[i];
    C2 += s[i] * yv[i];
  }
  C1 *= inv;
  C2 *= inv;
  return 0.5f * ((0.25f - 2.0f * C1) * (d1 - d0) + (0.25f - 2.0f * C2) * (d1 - d2));
}

// Sign property up to rounding.  The softmax weights put (C1, C2) in the
// convex hull of vertices that lie in, or on the boundary of, the region
// where the reconstructed jump has the sign of d1 (neural_network.cc:135-260),
// so in exact arithmetic j * d1 >= 0; on a boundary vertex the exact jump is
// 0.  A wrong-signed j no larger than the rounding bound of its two terms
// (1e-5 of |d1 - d0| + |d1 - d2|, about 80 FP32 ulp) is that zero rounded and
// becomes 0; a larger one is a genuine failure and is left for the SP-WENO
// fallback.
__device__ __forceinline__ float sign_round(float j, float d0, float d1, float d2) {
  const float tol = 1.0e-5f * (fabsf(d1 - d0) + fabsf(d1 - d2));
  return (j * d1 < 0.0f && fabsf(j) <= tol) ? 0.0f : j;
}

// softmax + vertex combination -> (C1, C2)
__device__ __forceinline__ void head_c(const float lg[5], const float xv[5], const float yv[5],
                                       float& C1, float& C2) {
  float mx = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) mx = fmaxf(mx, lg[i]);
  float s[5], sum = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    s[i] = exp(lg[i] - mx);
    sum += s[i];
  }
  const float inv = rcp(sum);
  C1 = 0.0f;
  C2 = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    C1 += s[i] * xv[i];
    C2 += s[i] * yv[i];
  }
  C1 *= inv;
  C2 *= inv;
}

// 5-5-5-5-5 network body, FP32 SIMT
__device__ __forceinline__ void body5(const NetView<float>& net, float st[5]) {
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const float* W = net.W + l * TECNO_NET_MAXW * TECNO_NET_MAXW;
    const float* b = net.b + l * TECNO_NET_MAXW;
    float o[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      float dot = b[j];
#pragma unroll
      for (int k = 0; k < 5; ++k) dot = fmaf(W[j * TECNO_NET_MAXW + k], st[k], dot);
      o[j] = dot;
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) st[j] = (l != 3) ? fmaxf(0.0f, o[j]) : o[j];
  }
}

}  // namespace fast

// no instruction is moved across this point by the machine scheduler
__device__ __forceinline__ void sched_fence() {
#ifndef TECNO_HOST_BUILD
  __builtin_amdgcn_sched_barrier(0);
#endif
}

// x[i] for a lane-varying i without indexing a register array
__device__ __forceinline__ float sel4(int i, float a, float b, float c, float d) {
  return i == 0 ? a : i == 1 ? b : i == 2 ? c : d;
}

// reconstruction arithmetic of the FP32 jump-form faces on the FP64 state
// (3D faces, and the 2D TeCNO variants other than the fused kepes_tecno4_roe
// stage): SP-WENO, or DSP-WENO with the network as FP32 SIMT or one
// v_mfma_f32_32x32x16 per layer
enum { M3_NONE = 0, M3_SPW = 1, M3_DSP_SIMT = 2, M3_DSP_BF16 = 3, M3_DSP_FP16 = 4 };


// Reconstructed jumps <<Z_i>> of NC characteristic components from their
// three FP32 cell jumps d0, d1, d2 and centre values z1: SP-WENO's jump
// (limiter.h:193-252) or DSP-WENO's (neural_network.cc:16-77, 320-325,
// zero where |d1| < 1e-15 as reconstruct.cc:152).  The matrix-core network
// needs every lane of the wave (all lanes call this together).  With
// sign_fallback a network jump whose sign disagrees with the cell jump is
// replaced by SP-WENO's; nfb counts the replacements.
template <int MIX, int NC>
__device__ __forceinline__ void jumps_fp32(const Phys<double>& P, const float* d0,
                                           const float* d1, const float* d2, const float* z1,
                                           float* jmp, unsigned& nfb) {
#pragma unroll
  for (int i = 0; i < NC; ++i) jmp[i] = 0.0f;
  if constexpr (MIX == M3_SPW) {
#pragma unroll
    for (int i = 0; i < NC; ++i)
      jmp[i] = fast::spw_jump(d0[i], d1[i], d2[i], z1[i], z1[i] + d1[i], P.corr);
  } else {
#ifndef TECNO_HOST_BUILD
    [[maybe_unused]] mfma::Net5<MIX == M3_DSP_FP16> net;
    if constexpr (MIX != M3_DSP_SIMT) net.load(P.mfma_w[MIX == M3_DSP_FP16 ? 1 : 0], threadIdx.x & 63);
#endif
    // The loop stays rolled (the network's registers); the components move
    // through slot 0 of rotating copies of the jumps, and each result enters
    // the last slot of jr, which the remaining rotations carry to its place.
    // (Indexing the arrays by the loop counter instead puts them in scratch
    // memory.)
    float a0[NC], a1[NC], a2[NC], az[NC], jr[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      a0[k] = d0[k];
      a1[k] = d1[k];
      a2[k] = d2[k];
      az[k] = z1[k];
      jr[k] = 0.0f;
    }
#pragma unroll 1
    for (int i = 0; i < NC; ++i) {
      const float e1z = a1[0];
      const bool skip = !(fabsf(e1z) >= 1e-15f);
      float j = 0.0f;
#ifndef TECNO_HOST_BUILD
      if (!__all(skip))  // else a zero centre jump on every lane: jump 0
#endif
      {
        // a skipped lane is evaluated with a centre jump of 1: its ratios stay
        // finite, so the partner lane of its matrix column sees finite
        // products; its jump is set to 0 below
        const float e1 = skip ? 1.0f : e1z;
        const float e0 = a0[0], e2 = a2[0], y1 = az[0];
        const float y2 = y1 + e1, y0 = y1 - e0, y3 = y2 + e2;
        const float maxabs = fmaxf(fmaxf(fabsf(y0), fabsf(y1)), fmaxf(fabsf(y2), fabsf(y3)));
        float st[5], tm, tp, maxd;
        fast::features_nv(e0, e1, e2, maxabs, st, tm, tp, maxd);
#ifndef TECNO_HOST_BUILD
        if constexpr (MIX != M3_DSP_SIMT) {
          float lg[5];
          net.eval(st, lg);
          j = fast::sign_round(fast::head_combine(lg, tm, tp, maxd, e0, e1, e2), e0, e1, e2);
        } else
#endif
        {
          fast::body5(P.net32, st);
          j = fast::sign_round(fast::head_combine(st, tm, tp, maxd, e0, e1, e2), e0, e1, e2);
        }
        if (skip) j = 0.0f;
        // sign property: SP-WENO replaces a network jump whose sign disagrees
        // with the cell jump, here, so that no copy of the stencil jumps has
        // to outlive the loop (a wave branches into it only when one of its
        // lanes needs it)
        const bool fail = P.sign_fallback && j * e1 < 0.0f;
#ifndef TECNO_HOST_BUILD
        if (__any(fail))
#endif
        {
          if (fail) {
            j = fast::spw_jump(e0, e1, e2, y1, y1 + e1, 0);
            ++nfb;
          }
        }
      }
#pragma unroll
      for (int k = 0; k + 1 < NC; ++k) {
        a0[k] = a0[k + 1];
        a1[k] = a1[k + 1];
        a2[k] = a2[k + 1];
        az[k] = az[k + 1];
        jr[k] = jr[k + 1];
      }
      jr[NC - 1] = j;
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) jmp[k] = jr[k];
  }
}

// ---------------------------------------------------------------- FP64 helpers
// (rcp64, log64 and logavg_fast are in fv_device.hpp)

// logarithmic mean with the logarithms of the arguments known per cell
// (misc_func.h:14-32): with f = (b-a)/(b+a) and u = f^2, the reference takes
// (a+b) / (2 F(u)), F = 1 + u/3 + u^2/5 + u^3/7, for u < 1e-2 and otherwise
// (a+b) f / ln(b/a), which equals (b-a) / (ln b - ln a).  The branch test
// u < 1e-2 is made as (b-a)^2 < 1e-2 (b+a)^2, so each branch needs one
// reciprocal of its own denominator; both are evaluated and one is selected
// (no divergent branch in a wave).
// The second reciprocal serves the mean's inverse as well, when asked for:
// 1/mean is 2F/(a+b) = 2 F r in the series branch and (ln b - ln a)/(b - a)
// in the other, whose reciprocal of (b - a) is the one taken there.
__device__ __forceinline__ double logavg_ln(double a, double b, double la, double lb,
                                            double* inv = nullptr) {
  const double dab = b - a, sab = b + a;
  const bool series = dab * dab < 1.0e-2 * (sab * sab);
  const double r = rcp64(series ? sab : lb - la);
  const double f = dab * r;
  const double u = f * f;
  const double F = 1.0 + u * (1.0 / 3.0) + (u * u) * (1.0 / 5.0) + (u * u * u) * (1.0 / 7.0);
  if (!inv) return series ? 0.5 * sab * rcp64(F) : dab * r;
  const double q = rcp64(series ? F : dab);
  *inv = series ? 2.0 * F * r : (lb - la) * q;
  return series ? 0.5 * sab * q : dab * r;
}

// ---------------------------------------------------------------- the face
enum { MM_SIMT = 0, MM_BF16 = 1, MM_FP16 = 2 };

// record layout in LDS, double [7][NB]: rho, ux, uy, beta = rho/(2p),
// ln rho, ln beta, Vf (first entropy variable)
enum { R_RHO = 0, R_UX, R_UY, R_BETA, R_LRHO, R_LBETA, R_VF, R_N };

// Scalar problems on an FP64 state with FP32 reconstruction: state,
// update and the entropy-conservative part of the lin_/bur_ TeCNO fluxes as
// ScalarPh<double>, the SP-WENO jump <<u>> of the dissipation term in FP32
// from the cell jumps taken in FP64 (spw_jump).  Its rounding scales with
// the jumps, so it stays below the discretisation error on fine meshes,
// where an FP32 state's own rounding does not.
template <int G>
struct ScalarMixedPh : ScalarPh<double, G> {
  using Base = ScalarPh<double, G>;
  __device__ static void face(const Phys<double>& P, const void* ldsv, int NB, int base,
                              int stride, int ax, double F[1], unsigned& nfb) {
    const double* lds = (const double*)ldsv;
    double s[2 * G];
#pragma unroll
    for (int l = 0; l < 2 * G; ++l) s[l] = lds[base + l * stride];
    double f = Base::ec_flux(P, s, ax);
    if (Base::tecno(P)) {
      const double D = Base::diss_coef(P, s, ax);
      const float jump = fast::spw_jump(float(s[G - 1] - s[G - 2]), float(s[G] - s[G - 1]),
                                        float(s[G + 1] - s[G]), float(s[G - 1]), float(s[G]),
                                        P.corr);
      f -= 0.5 * D * double(jump);
    }
    F[0] = f * P.area[ax];
  }
};

template <int RV, int MM>
struct Euler2Mixed {
  using TS = double;
  // waves per SIMD the register budget is cut to: four for the network paths
  // (the 16x15 tile needs 38 KB of LDS, so four workgroups fit a CU; the
  // matrix-core and transcendental chains need the extra waves for latency
  // hiding, which outweighs a few spilled registers -- measured on the KH
  // 2048^2 stage: 4 waves 0.448 ms, 3 waves 0.488 ms), the compiler's choice
  // otherwise
#ifdef TECNO_MIXED_MIN_WAVES
  static constexpr int MIN_WAVES = TECNO_MIXED_MIN_WAVES;
#else
  static constexpr int MIN_WAVES = RV == RV_DSP ? 4 : 1;
#endif
  using TF = double;
  using TP = float;
  static constexpr int G = 2;
  static constexpr int NV = 4;
  static constexpr size_t rec_bytes(int NB) { return (size_t)R_N * NB * 8; }

  __device__ static void load(const Phys<float>& P, const double* U, int64_t N, int64_t idx,
                              int flipx, int flipy, void* lds, int NB, int c) {
    const double rho = U[idx], mx = U[N + idx], my = U[2 * N + idx], E = U[3 * N + idx];
    const double inv = rcp64(rho);
    double ux = mx * inv, uy = my * inv;
    const double p = P.gm1_d * (E - 0.5 * (mx * mx + my * my) * inv);
    if (flipx) ux = -ux;
    if (flipy) uy = -uy;
    const double beta = 0.5 * rho * rcp64(p);
    const double lr = log64(rho), lp = log64(p);
    const double s = lp - P.gamma_d * lr;  // Entropy (fv.cc:557-560)
    double* ld = (double*)lds;
    ld[R_RHO * NB + c] = rho;
    ld[R_UX * NB + c] = ux;
    ld[R_UY * NB + c] = uy;
    ld[R_BETA * NB + c] = beta;
    ld[R_LRHO * NB + c] = lr;
    ld[R_LBETA * NB + c] = lr - lp - 0.69314718055994531;  // ln(rho/(2p))
    ld[R_VF * NB + c] = (P.gamma_d - s) * P.inv_gm1_d - (ux * ux + uy * uy) * beta;
  }

  // kepec_flux (kepec.cc:10-48) per unit area in the face frame, FP64
  __device__ static void kepec(const Phys<float>& P, double rl, double nl, double tl, double bl,
                               double lrl, double lbl, double rr, double nr, double tr, double br,
                               double lrr, double lbr, double F[4], double* aux = nullptr) {
    double ibeta;
    const double rho = logavg_ln(rl, rr, lrl, lrr);
    const double beta = logavg_ln(bl, br, lbl, lbr, &ibeta);
    const double vel2 = 0.5 * ((nl * nl + tl * tl) + (nr * nr + tr * tr));
    const double p = 0.5 * (rl + rr) * rcp64(bl + br);
    const double vn = 0.5 * (nl + nr), vt = 0.5 * (tl + tr);
    const double mass = rho * vn;
    F[0] = mass;
    F[1] = p + vn * mass;
    F[2] = vt * mass;
    F[3] = 0.5 * (ibeta * P.inv_gm1_d - vel2) * mass + (F[1] * vn + F[2] * vt);
    if (aux) {
      aux[0] = rho;
      aux[1] = beta;
      aux[2] = p;
      aux[3] = vn;
      aux[4] = vt;
      aux[5] = ibeta;
    }
  }

  // two-point flux of the distance-2 pair (ca, cb) in the frame of axis ax
  __device__ static void pair_flux(const Phys<float>& P, const double* ld, int NB, int ca, int cb,
                                   int ax, double f[4]) {
    const int rn = ax ? R_UY : R_UX, rt = ax ? R_UX : R_UY;
    kepec(P, ld[R_RHO * NB + ca], ld[rn * NB + ca], ld[rt * NB + ca], ld[R_BETA * NB + ca],
          ld[R_LRHO * NB + ca], ld[R_LBETA * NB + ca], ld[R_RHO * NB + cb], ld[rn * NB + cb],
          ld[rt * NB + cb], ld[R_BETA * NB + cb], ld[R_LRHO * NB + cb], ld[R_LBETA * NB + cb], f);
  }

  // the pair fluxes of a tile, double [4][NP] (one per pair, shared by the
  // two faces whose stencils hold the pair)
  static constexpr bool PAIRS = true;
  __device__ static void pair(const Phys<float>& P, const void* ldsv, int NB, int ca, int cb,
                              int ax, double* out, int NP, int q) {
    double f[4];
    pair_flux(P, (const double*)ldsv, NB, ca, cb, ax, f);
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v * NP + q] = f[v];
  }

  // a face whose pair fluxes are in LDS: pairs q (cells 0,2) and q+ps (1,3)
  __device__ static void face(const Phys<float>& P, const void* ldsv, int NB, int base, int stride,
                              int ax, double F[4], unsigned& nfb, const double* pairs, int NP,
                              int q, int ps) {
    face_core(P, ldsv, NB, base, stride, ax, F, nfb, [&](int pr, double f0[4]) {
#pragma unroll
      for (int v = 0; v < 4; ++v) f0[v] = pairs[v * NP + q + pr * ps];
    });
  }

  // a lone face (batched seam): the pair fluxes are evaluated here
  __device__ static void face(const Phys<float>& P, const void* ldsv, int NB, int base, int stride,
                              int ax, double F[4], unsigned& nfb) {
    face_core(P, ldsv, NB, base, stride, ax, F, nfb, [&](int pr, double f0[4]) {
      pair_flux(P, (const double*)ldsv, NB, base + pr * stride, base + (pr + 2) * stride, ax, f0);
    });
  }

  template <class PairSrc>
  __device__ static void face_core(const Phys<float>& P, const void* ldsv, int NB, int base,
                                   int stride, int ax, double F[4], unsigned& nfb,
                                   PairSrc&& pair_src) {
    const double* ld = (const double*)ldsv;
    const int rn = ax ? R_UY : R_UX, rt = ax ? R_UX : R_UY;
    // Order of work keeps the register footprint small: the eigensystem and
    // the projected jumps first, then the FP32 reconstruction with only float
    // copies of the eigenvectors live, then the FP64 EC4 flux from the
    // records re-read out of LDS.
    // ---- central two-point flux F(1,2) and its means (kepec.cc:10-48)
    double Fc[4], aux[6];
    {
      const int c1 = base + stride, c2 = base + 2 * stride;
      kepec(P, ld[R_RHO * NB + c1], ld[rn * NB + c1], ld[rt * NB + c1], ld[R_BETA * NB + c1],
            ld[R_LRHO * NB + c1], ld[R_LBETA * NB + c1], ld[R_RHO * NB + c2], ld[rn * NB + c2],
            ld[rt * NB + c2], ld[R_BETA * NB + c2], ld[R_LRHO * NB + c2], ld[R_LBETA * NB + c2], Fc,
            aux);
    }
    // ---- eigensystem at the face (kepes_tecno4_roe.cc:39-113) and the
    // projection of the entropy-variable jumps, FP64: a jump of a scaled
    // variable can be a near-total cancellation of its terms (e.g. the shear
    // wave at a symmetric extremum), and the zero-jump tests (1e-13 / 1e-15)
    // are absolute
    float dZ[4][3], Z1[4];
    float vnf, af, s0, s1, Rn0f, Rt0f, RE0f, Rn1f, Rt1f, RE1f, Rt2f, RE2f, Rn4f, RE4f;
    {
      const double rhoh = aux[0], ph = aux[2], vn = aux[3], vt = aux[4], ibetah = aux[5];
      // c1 and S2 in FP32 (one instruction each instead of a scaled FP64
      // sequence): they scale whole rows of R~^T, so their rounding scales a
      // projected jump without disturbing the cancellation inside it.  The
      // sound speed a sits inside the acoustic rows and stays FP64.
      const double a = sqrt(0.5 * P.gamma_d * ibetah);
      const double vsq = vn * vn + vt * vt;
      const double H = a * a * P.inv_gm1_d + 0.5 * vsq;
      const double c1 = (double)fast::sqrt((float)(rhoh * P.inv_gamma_d));
      const double S0 = P.s05_d * c1, S1 = P.sgm1_d * c1, S2 = (double)fast::sqrt((float)ph);
      const double Rn0 = (vn - a) * S0, Rt0 = vt * S0, RE0 = (H - a * vn) * S0;
      const double Rn1 = vn * S1, Rt1 = vt * S1, RE1 = (0.5 * vsq) * S1;
      const double Rt2 = -S2, RE2 = -vt * S2;
      const double Rn4 = (vn + a) * S0, Rt4 = Rt0, RE4 = (H + a * vn) * S0;
      double Vf[4], Vn[4], Vt[4], Vl[4];
#pragma unroll
      for (int l = 0; l < 4; ++l) {
        const int cc = base + l * stride;
        const double be = ld[R_BETA * NB + cc];
        Vf[l] = ld[R_VF * NB + cc];
        Vn[l] = 2.0 * be * ld[rn * NB + cc];
        Vt[l] = 2.0 * be * ld[rt * NB + cc];
        Vl[l] = -2.0 * be;
      }
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double df = Vf[j + 1] - Vf[j], dn = Vn[j + 1] - Vn[j];
        const double dt = Vt[j + 1] - Vt[j], dl = Vl[j + 1] - Vl[j];
        dZ[0][j] = (float)(S0 * df + Rn0 * dn + Rt0 * dt + RE0 * dl);
        dZ[1][j] = (float)(S1 * df + Rn1 * dn + Rt1 * dt + RE1 * dl);
        dZ[2][j] = (float)(Rt2 * dt + RE2 * dl);
        dZ[3][j] = (float)(S0 * df + Rn4 * dn + Rt4 * dt + RE4 * dl);
      }
      Z1[0] = (float)(S0 * Vf[1] + Rn0 * Vn[1] + Rt0 * Vt[1] + RE0 * Vl[1]);
      Z1[1] = (float)(S1 * Vf[1] + Rn1 * Vn[1] + Rt1 * Vt[1] + RE1 * Vl[1]);
      Z1[2] = (float)(Rt2 * Vt[1] + RE2 * Vl[1]);
      Z1[3] = (float)(S0 * Vf[1] + Rn4 * Vn[1] + Rt4 * Vt[1] + RE4 * Vl[1]);
      vnf = (float)vn;
      af = (float)a;
      s0 = (float)S0;
      s1 = (float)S1;
      Rn0f = (float)Rn0;
      Rt0f = (float)Rt0;
      RE0f = (float)RE0;
      Rn1f = (float)Rn1;
      Rt1f = (float)Rt1;
      RE1f = (float)RE1;
      Rt2f = (float)Rt2;
      RE2f = (float)RE2;
      Rn4f = (float)Rn4;
      RE4f = (float)RE4;
    }
    // DSP: keep the scheduler from hoisting the later LDS re-reads (and their
    // registers) across the network evaluation
    if constexpr (RV == RV_DSP) sched_fence();
    float jmp[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    unsigned bad = 0;  // components whose network jump fails the sign property
#ifndef TECNO_HOST_BUILD
    if constexpr (RV == RV_DSP && MM != MM_SIMT) {
      // network on the matrix cores: every lane of the wave takes part in
      // each of the 4 component evaluations
      mfma::Net5<MM == MM_FP16> net;
      net.load(P.mfma_w[MM == MM_FP16 ? 1 : 0], threadIdx.x & 63);
      // components two at a time: two independent feature / network / head
      // chains interleave (127 VGPRs at the four waves per SIMD that LDS
      // allows; KH 2048^2 stage 0.38 -> 0.37 ms), where all four at once
      // spill (0.46 ms)
#pragma unroll 1
      for (int p = 0; p < 2; ++p) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        // component i = 2p + h: one select per value between the pair's
        // components (p is the same on every lane)
        const int i = 2 * p + h;
        const float d1z = p ? dZ[2 + h][1] : dZ[h][1];
        const bool skip = !(fabsf(d1z) >= 1e-15f);
        // a component whose centre jump is zero on every lane of the wave
        // needs no network: its reconstructed jump is 0 (reconstruct.cc:152)
        // and jmp[i] already holds it.  Uniform regions of a flow (a shear
        // layer's free streams) have whole waves of such faces.
        if (__all(skip)) continue;
        // a skipped stencil (zero centre jump) is evaluated with a centre jump
        // of 1, so its ratios stay finite and the partner lane of its matrix
        // column sees finite products; its jump is set to 0 below.  (The
        // features are finite for any finite state: tanh of an overflowing
        // ratio saturates to +-1.)
        const float d1 = skip ? 1.0f : d1z;
        const float d0 = p ? dZ[2 + h][0] : dZ[h][0];
        const float d2 = p ? dZ[2 + h][2] : dZ[h][2];
        const float z1 = p ? Z1[2 + h] : Z1[h];
        const float z2 = z1 + d1, z0 = z1 - d0, z3 = z2 + d2;
        const float maxabs = fmaxf(fmaxf(fabsf(z0), fabsf(z1)), fmaxf(fabsf(z2), fabsf(z3)));
        float st[5], lg[5], tm, tp, maxd;
        fast::features_nv(d0, d1, d2, maxabs, st, tm, tp, maxd);
        net.eval(st, lg);
        float j = fast::sign_round(fast::head_combine(lg, tm, tp, maxd, d0, d1, d2), d0, d1, d2);
        if (skip) j = 0.0f;
        bad |= (j * d1 < 0.0f) ? (1u << i) : 0u;
        jmp[h] = p ? jmp[h] : j;
        jmp[2 + h] = p ? j : jmp[2 + h];
      }
      }
    } else
#endif
    if constexpr (RV == RV_DSP) {
      // FP32 network on the vector units, components one after the other
#pragma unroll 1
      for (int i = 0; i < 4; ++i) {
        const float d1 = sel4(i, dZ[0][1], dZ[1][1], dZ[2][1], dZ[3][1]);
#ifndef TECNO_HOST_BUILD
        if (__all(!(fabsf(d1) >= 1e-15f))) continue;  // zero jump on every lane
#endif
        const float d0 = sel4(i, dZ[0][0], dZ[1][0], dZ[2][0], dZ[3][0]);
        const float d2 = sel4(i, dZ[0][2], dZ[1][2], dZ[2][2], dZ[3][2]);
        const float z1 = sel4(i, Z1[0], Z1[1], Z1[2], Z1[3]);
        const float z2 = z1 + d1, z0 = z1 - d0, z3 = z2 + d2;
        const float maxabs = fmaxf(fmaxf(fabsf(z0), fabsf(z1)), fmaxf(fabsf(z2), fabsf(z3)));
        float st[5], tm, tp, maxd;
        fast::features_nv(d0, d1, d2, maxabs, st, tm, tp, maxd);
        fast::body5(P.net, st);
        float j = fast::sign_round(fast::head_combine(st, tm, tp, maxd, d0, d1, d2), d0, d1, d2);
        if (fabsf(d1) < 1e-15f) j = 0.0f;  // reconstruct_sp_weno_dl zero-jump rule
        bad |= (j * d1 < 0.0f) ? (1u << i) : 0u;
#pragma unroll
        for (int k = 0; k < 4; ++k) jmp[k] = (k == i) ? j : jmp[k];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float d0 = dZ[i][0], d1 = dZ[i][1], d2 = dZ[i][2];
        const float z1 = Z1[i], z2 = z1 + d1;
        if (RV == RV_SPWENO || (RV == RV_GENERIC && P.recon != REC_DSP)) {
          jmp[i] = fast::spw_jump(d0, d1, d2, z1, z2, P.corr);
        } else {
          const float z0 = z1 - d0, z3 = z2 + d2;
          const float maxabs = fmaxf(fmaxf(fabsf(z0), fabsf(z1)), fmaxf(fabsf(z2), fabsf(z3)));
          float st[5], tm, tp, maxd;
          fast::features_nv(d0, d1, d2, maxabs, st, tm, tp, maxd);
          fast::body5(P.net, st);
          float j = fast::sign_round(fast::head_combine(st, tm, tp, maxd, d0, d1, d2), d0, d1, d2);
          if (fabsf(d1) < 1e-15f) j = 0.0f;  // reconstruct_sp_weno_dl zero-jump rule
          bad |= (j * d1 < 0.0f) ? (1u << i) : 0u;
          jmp[i] = j;
        }
      }
    }
    // sign property: SP-WENO replaces the network's jump where its sign
    // disagrees with the cell jump.  One component per pass, so a wave pays
    // for the largest count of one lane rather than for every component that
    // any of its lanes needs.
    if (!P.sign_fallback) bad = 0;
    while (bad) {
      const int i = __builtin_ctz(bad);
      bad &= bad - 1;
      const float d0 = sel4(i, dZ[0][0], dZ[1][0], dZ[2][0], dZ[3][0]);
      const float d1 = sel4(i, dZ[0][1], dZ[1][1], dZ[2][1], dZ[3][1]);
      const float d2 = sel4(i, dZ[0][2], dZ[1][2], dZ[2][2], dZ[3][2]);
      const float z1 = sel4(i, Z1[0], Z1[1], Z1[2], Z1[3]);
      const float j = fast::spw_jump(d0, d1, d2, z1, z1 + d1, 0);
#pragma unroll
      for (int k = 0; k < 4; ++k) jmp[k] = (k == i) ? j : jmp[k];
      ++nfb;
    }
    // ---- dissipation R~ Lambda dZ (FP32)
    const float L1 = fabsf(vnf), L0 = fabsf(vnf - af), L4 = fabsf(vnf + af);
    const float Dr = s0 * L0 * jmp[0] + s1 * L1 * jmp[1] + s0 * L4 * jmp[3];
    const float Dn = Rn0f * L0 * jmp[0] + Rn1f * L1 * jmp[1] + Rn4f * L4 * jmp[3];
    const float Dt = Rt0f * L0 * jmp[0] + Rt1f * L1 * jmp[1] + Rt2f * L1 * jmp[2] +
                     Rt0f * L4 * jmp[3];
    const float DE = RE0f * L0 * jmp[0] + RE1f * L1 * jmp[1] + RE2f * L1 * jmp[2] +
                     RE4f * L4 * jmp[3];
    if constexpr (RV == RV_DSP) sched_fence();
    // ---- entropy-conservative flux, 4th order (kepec4.cc:10-38), FP64
    double Fe[4], f0[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) Fe[v] = (4.0 / 3.0) * Fc[v];
#pragma unroll
    for (int pr = 0; pr < 2; ++pr) {
      pair_src(pr, f0);  // F(0,2), then F(1,3)
#pragma unroll
      for (int v = 0; v < 4; ++v) Fe[v] -= (1.0 / 6.0) * f0[v];
    }
    // ---- TeCNO flux
    const double area = P.area_d[ax];
    const double Fn = (Fe[1] - 0.5 * (double)Dn) * area;
    const double Ft = (Fe[2] - 0.5 * (double)Dt) * area;
    F[0] = (Fe[0] - 0.5 * (double)Dr) * area;
    F[1] = ax ? Ft : Fn;
    F[2] = ax ? Fn : Ft;
    F[3] = (Fe[3] - 0.5 * (double)DE) * area;
  }

  __device__ static double speed(const Phys<float>& P, const double* U, int64_t N, int64_t idx) {
    const double rho = U[idx], mx = U[N + idx], my = U[2 * N + idx], E = U[3 * N + idx];
    const double inv = 1.0 / rho;
    const double ux = mx * inv, uy = my * inv;
    const double p = P.gm1_d * (E - 0.5 * (mx * mx + my * my) * inv);
    return sqrt(ux * ux + uy * uy) + sqrt(P.gamma_d * p * inv);
  }
  __device__ static bool bad(const double* U, int64_t N, int64_t idx) {
    const double rho = U[idx], mx = U[N + idx], my = U[2 * N + idx], E = U[3 * N + idx];
    const double e = E - 0.5 * (mx * mx + my * my) / rho;
    return !(rho > 0.0) || !(e > 0.0) || !(e < INFINITY);
  }
};

}  // namespace tecno
