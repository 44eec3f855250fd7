// fv_device.hpp — device numerics of the TeCNO step (templated on the
// arithmetic type T = double or float).
//
// Each function restates one reference routine; the operation order is the
// reference's wherever rounding depends on it, so that the FP64 build
// (compiled with -ffp-contract=off) reproduces the reference's arithmetic.
// Terms that the reference adds as exact zeros in two dimensions (w = 0) are
// left out: x + (+-0) == x, so this changes no bit of a nonzero result.
#pragma once
#ifdef TECNO_HOST_BUILD
#include "host_shim.hpp"
#else
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace tecno {

// ---------------------------------------------------------------- math shims
__device__ __forceinline__ double m_log(double x) { return log(x); }
__device__ __forceinline__ float m_log(float x) { return logf(x); }
__device__ __forceinline__ double m_exp(double x) { return exp(x); }
__device__ __forceinline__ float m_exp(float x) { return expf(x); }
__device__ __forceinline__ double m_tanh(double x) { return tanh(x); }
__device__ __forceinline__ float m_tanh(float x) { return tanhf(x); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float m_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double m_abs(double x) { return fabs(x); }
__device__ __forceinline__ float m_abs(float x) { return fabsf(x); }
// std::min / std::max semantics (return the first argument on ties / NaN)
template <typename T> __device__ __forceinline__ T m_min(T a, T b) { return (b < a) ? b : a; }
template <typename T> __device__ __forceinline__ T m_max(T a, T b) { return (a < b) ? b : a; }
// x^3 with the product error compensated (close to a correctly rounded pow(x,3))
__device__ __forceinline__ double m_cube(double x) {
  double p = x * x;
  double e = fma(x, x, -p);
  return fma(p, x, e * x);
}
__device__ __forceinline__ float m_cube(float x) { return x * x * x; }

// ---------------------------------------------------------------- logavg
// enflo/src/misc_func.h:14-32
template <typename T>
__device__ __forceinline__ T logavg(T a, T b) {
  T xi = b / a;
  T f = (xi - T(1)) / (xi + T(1));
  T u = f * f;
  T F;
  if (u < T(1.0e-2)) {
    T u2 = u * u;
    T u3 = u2 * u;
    F = T(1) + u / T(3) + u2 / T(5) + u3 / T(7);
  } else {
    F = m_log(xi) / T(2) / f;
  }
  return T(0.5) * (a + b) / F;
}

// ---------------------------------------------------------------- FP64 helpers of the FP32 paths
#ifdef TECNO_HOST_BUILD
__device__ __forceinline__ double rcp64(double x) { return 1.0 / x; }
#else
// v_rcp_f64 refined by two Newton steps (about 1 ulp)
__device__ __forceinline__ double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  r = fma(fma(-x, r, 1.0), r, r);
  return r;
}
#endif

// Natural logarithm of a double for the per-cell records of the FP32 paths.
// x = m 2^e with m in [1/sqrt2, sqrt2) (frexp, then one doubling), and
// ln m = 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.1716:
//   2 atanh(s) = 2s + 2s z (1/3 + z/5 + ... + z^8/19),  z = s^2 <= 0.0295;
// the first omitted term, 2s z^10/21, is below 2^-55 of the sum.  s is the
// residual-corrected quotient of the refined reciprocal, ln 2 is split into
// a 42-bit head (e ln2_hi is exact for any exponent) and its tail.  About 35
// instructions against about 75 for the library routine; within 2 ulp of the
// C library's log (99.7% of arguments within 1 ulp; the host build of this
// function is checked against std::log).
// x <= 0 or NaN gives NaN, as the library routine does.
__device__ __forceinline__ double log64(double x) {
#ifdef TECNO_HOST_BUILD
  int e0;
  const double m0 = std::frexp(x, &e0);
#else
  const double m0 = __builtin_amdgcn_frexp_mant(x);
  const int e0 = __builtin_amdgcn_frexp_exp(x);
#endif
  const bool lo = m0 < 0.70710678118654752;
  const double m = lo ? m0 + m0 : m0;
  const double e = (double)(lo ? e0 - 1 : e0);
  const double f = m - 1.0, d = m + 1.0;
  const double r = rcp64(d);
  double s = f * r;
  s = fma(fma(-s, d, f), r, s);
  const double z = s * s, t = s + s;
  double q = 1.0 / 19.0;
  q = fma(q, z, 1.0 / 17.0);
  q = fma(q, z, 1.0 / 15.0);
  q = fma(q, z, 1.0 / 13.0);
  q = fma(q, z, 1.0 / 11.0);
  q = fma(q, z, 1.0 / 9.0);
  q = fma(q, z, 1.0 / 7.0);
  q = fma(q, z, 1.0 / 5.0);
  q = fma(q, z, 1.0 / 3.0);
  const double lnm = fma(t * z, q, t);
  const double res = fma(e, 0x1.62e42fefa3800p-1, fma(e, 0x1.ef35793c76730p-45, lnm));
  return x > 0.0 ? res : __builtin_nan("");
}

// logarithmic mean (misc_func.h:14-32) for the FP64 flux of the FP32
// reconstruction paths, which are measured against the FP64 path within a
// tolerance: f = (b - a)/(b + a) (the reference's (xi - 1)/(xi + 1), xi =
// b/a) from one refined reciprocal, the series' divisions by 3, 5, 7 as
// products, and in the logarithm branch (a + b)/(2F) = (b - a)/ln(b/a) with
// log64.  Within a few ulp of the reference form; 2 reciprocals where that
// takes 7 divisions.
__device__ __forceinline__ double logavg_fast(double a, double b) {
  const double f = (b - a) * rcp64(b + a);
  const double u = f * f;
  if (u < 1.0e-2) {
    const double F = 1.0 + u * (1.0 / 3.0) + (u * u) * (1.0 / 5.0) + (u * u * u) * (1.0 / 7.0);
    return 0.5 * (a + b) * rcp64(F);
  }
  return (b - a) * rcp64(log64(b * rcp64(a)));
}

// FAST: the FP32 reconstruction paths' FP64 flux, measured against the FP64
// path within a tolerance, takes the refined reciprocal for divisions and
// logavg_fast for the logarithmic means (a few ulp from the reference form,
// which the FP64 path keeps)
template <bool FAST, typename T>
__device__ __forceinline__ T lavg(T a, T b) {
  if constexpr (FAST) return logavg_fast(a, b);
  else return logavg(a, b);
}
template <bool FAST, typename T>
__device__ __forceinline__ T qdiv(T a, T b) {
  if constexpr (FAST) return a * rcp64(b);
  else return a / b;
}

// ---------------------------------------------------------------- SP-WENO
// enflo/src/limiter.h:103-252.  eps = 1e-13 (limiter.h:14).
template <typename T> struct SpwConst { static constexpr double eps = 1.0e-13; };

template <typename T>
__device__ __forceinline__ T spw_fp(T t0, T t1) {
  const T eps = T(SpwConst<T>::eps);
  T psi = (m_abs(t0 - T(1)) < eps) ? T(0) : (T(1) - t1) / (T(1) - t0);
  return (m_abs(t0 - T(1)) > eps && m_abs(psi + T(1)) > eps) ? T(1) / (T(1) + psi) : T(1);
}

template <typename T>
__device__ __forceinline__ void spw_cp(T t0, T t1, T del1, T scale, bool want_corr, T& C,
                                       T& corr) {
  const T eps = T(SpwConst<T>::eps);
  T psi;
  if (m_abs(t0 - T(1)) < eps || m_abs(t1 - T(1)) < eps)
    psi = T(0);
  else
    psi = (T(1) - t1) / (T(1) - t0);
  corr = T(0);
  if (m_abs(psi + T(1)) > eps && m_abs(t0 - T(1)) > eps && psi < T(0)) {
    T a = spw_fp(t0, t1), b = spw_fp(t1, t0);
    C = (a / (a * a + b * b)) / T(8);
    if (want_corr)
      corr += T(-0.25) * m_cube(m_min(m_abs(del1 / scale), m_abs(del1))) / (T(1) - t0);
  } else if (m_abs(t0 - T(1)) > eps && m_abs(psi + T(1)) < eps) {
    C = T(0);
  } else if (m_abs(t0 - T(1)) < eps ||
             (psi > -eps && (sizeof(T) == 4 ? m_abs(t0) <= T(1) : m_abs(t0) < T(1) + eps))) {
    // |t0| < 1 + eps; in FP32 1 + 1e-13 rounds to 1, so the floats below
    // 1 + 1e-13 are those <= 1 (|t0| = 1 exactly: equal or alternating jumps)
    C = T(-3.0 / 8.0);
  } else if (psi > -eps && m_abs(t0) > T(1)) {
    C = T(1.0 / 8.0);
  }
}

// FP32 form of the two Cp calls of sp_weno_rec from one ratio
// psi = (1 - t-)/(1 - t+): C2's psi is 1/psi, so both weights take Cp's first
// branch together (psi < 0), where a / (a^2 + b^2) / 8 with a = fp(t0, t1),
// b = fp(t1, t0) is (1 + psi) / (1 + psi^2) / 8 for C1 and psi times that
// for C2 (limiter.h:103, 152-190); otherwise Cp is -3/8 for |t0| <= 1 (the
// floats below 1 + 1e-13) and 1/8 above.  Four divisions instead of ten.
template <typename T>
__device__ __forceinline__ void spw_c12_ratio(T tp, T tm, T del1, T scale, bool want_corr, T& C1,
                                              T& C2, T& C1c, T& C2c) {
  const T eps = T(SpwConst<T>::eps);
  const bool one = m_abs(tp - T(1)) < eps || m_abs(tm - T(1)) < eps;
  const T psi = one ? T(0) : (T(1) - tm) / (T(1) - tp);
  C1c = T(0);
  C2c = T(0);
  if (psi < T(0)) {
    if (m_abs(psi + T(1)) < eps) {  // Cp's second branch (psi = -1)
      C1 = T(0);
      C2 = T(0);
    } else {
      C1 = T(0.125) * (T(1) + psi) / (psi * psi + T(1));
      C2 = psi * C1;
      if (want_corr) {
        const T m3 = T(-0.25) * m_cube(m_min(m_abs(del1 / scale), m_abs(del1)));
        C1c = m3 / (T(1) - tp);
        C2c = m3 / (T(1) - tm);
      }
    }
  } else {
    C1 = m_abs(tp) <= T(1) ? T(-3.0 / 8.0) : T(1.0 / 8.0);
    C2 = m_abs(tm) <= T(1) ? T(-3.0 / 8.0) : T(1.0 / 8.0);
  }
}

// sp_weno_rec (limiter.h:193-252); corr 0 sp_weno, 1 sp_wenoc, 2 corr2 (== 0).
// FP64 keeps the reference's two Cp calls (its operation order); FP32 takes
// spw_c12_ratio.
template <typename T>
__device__ __forceinline__ void sp_weno(T s0, T s1, T s2, T s3, int corr, T& recl, T& recr) {
  const T eps = T(SpwConst<T>::eps);
  T del0 = s1 - s0, del1 = s2 - s1, del2 = s3 - s2;
  if (m_abs(del1) < eps) {
    recl = s1;
    recr = s2;
    return;
  }
  T tp = del0 / del1;
  T tm = del2 / del1;
  T scale = T(0.5) * (m_abs(s1) + m_abs(s2));
  T C1 = 0, C2 = 0, C1c = 0, C2c = 0;
  bool wc = (corr == 1);
  if constexpr (sizeof(T) == 4) {
    spw_c12_ratio(tp, tm, del1, scale, wc, C1, C2, C1c, C2c);
  } else {
    spw_cp(tp, tm, del1, scale, wc, C1, C1c);
    spw_cp(tm, tp, del1, scale, wc, C2, C2c);
  }
  if (corr > 0) {
    if (corr == 1) {
      C1 = C1 + C1c;
      C2 = C2 + C2c;
    }
    C1 = m_min(m_max(C1, T(-3.0 / 8.0)), T(1.0 / 8.0));
    C2 = m_min(m_max(C2, T(-3.0 / 8.0)), T(1.0 / 8.0));
  }
  T w0 = T(3.0 / 4.0) + T(2) * C1;
  T w0t = T(1.0 / 4.0) - T(2) * C2;
  recl = w0 * T(0.5) * (s1 + s2) + (T(1) - w0) * (T(-0.5) * s0 + T(1.5) * s1);
  recr = (T(1) - w0t) * T(0.5) * (s1 + s2) + w0t * (T(-0.5) * s3 + T(1.5) * s2);
}

// ---------------------------------------------------------------- ENO
// pointwise ENO k = 2, 3: limiter.h:8-12, 63-97; reconstruct_eno_pw
// (enflo_scalar/src_scalar/reconstruct.cc:80-104).  st has 2*ng values.
template <typename T>
__device__ __forceinline__ T eno_pw(int k, const T* s) {
  T dd1[4], dd2[3];
  int r = 0;
  const int sc = k - 1;
  for (int j = 0; j < 2 * k - 2; ++j) dd1[j] = s[j + 1] - s[j];
  if (k == 3)
    for (int j = 0; j < 3; ++j) dd2[j] = dd1[j + 1] - dd1[j];
  // i = 1
  if (m_abs(dd1[sc - r]) > m_abs(dd1[sc - r - 1])) r++;
  if (k == 3 && m_abs(dd2[sc - r]) > m_abs(dd2[sc - r - 1])) r++;
  T res = T(0);
  const int row = sc - r;
  if (k == 2) {
    const T c0 = row == 0 ? T(-0.5) : T(0.5);
    const T c1 = row == 0 ? T(1.5) : T(0.5);
    res += c0 * s[row];
    res += c1 * s[row + 1];
  } else {
    T c0, c1, c2;
    if (row == 0) { c0 = T(0.375); c1 = T(-1.25); c2 = T(1.875); }
    else if (row == 1) { c0 = T(-0.125); c1 = T(0.75); c2 = T(0.375); }
    else { c0 = T(0.375); c1 = T(0.75); c2 = T(-0.125); }
    res += c0 * s[row];
    res += c1 * s[row + 1];
    res += c2 * s[row + 2];
  }
  return res;
}

template <typename T>
__device__ __forceinline__ void eno_rec(int k, const T* st, int ng, T& l, T& r) {
  T comp[5];
  const int s = 2 * ng;
  for (int j = 0; j < 2 * k - 1; ++j) comp[j] = st[s / 2 - k + j];
  l = eno_pw(k, comp);
  for (int j = 0; j < 2 * k - 1; ++j) comp[2 * k - 2 - j] = st[s / 2 - k + 1 + j];
  r = eno_pw(k, comp);
}

// ---------------------------------------------------------------- DSP-WENO
// Neural_Network::forward (enflo/src/neural_network.cc:16-77) with
// compute_vertices (135-260, vertex_mode 2, permutation 2), softmax with a
// zero-initialised max (276-295) and compute_recon on the unscaled input.
#define TECNO_NET_MAXL 8
#define TECNO_NET_MAXW 16

template <typename T>
struct NetView {
  int nl;
  int w[TECNO_NET_MAXL + 1];
  const T* W;  // [l][TECNO_NET_MAXW * TECNO_NET_MAXW], row = output
  const T* b;  // [l][TECNO_NET_MAXW]
};

template <typename T>
__device__ __forceinline__ void nn_vertices(T tm, T tp, T maxd, T xv[5], T yv[5]) {
  const T thr = T(1e-8);
  T psi_p = (T(1) - tm) / (T(1) - tp);
  T psi_m = T(1) / psi_p;
  T g1 = m_min(maxd, T(1.0 / 8.0));
  T g2 = -m_min(maxd, T(3.0 / 8.0));
  T x1 = T(1.0 / 8.0) * (T(1) + psi_p) - g1 * psi_p;
  T y1 = T(1.0 / 8.0) * (T(1) + psi_m) - g2 * psi_m;
  T x2 = T(1.0 / 8.0) * (T(1) + psi_p) - g2 * psi_p;
  T y2 = T(1.0 / 8.0) * (T(1) + psi_m) - g1 * psi_m;
  const T e = T(1.0 / 8.0), t = T(-3.0 / 8.0), m = T(-1.0 / 8.0);
#define NV_SET(a0, a1, a2, a3, a4, b0, b1, b2, b3, b4) \
  { xv[0] = a0; xv[1] = a1; xv[2] = a2; xv[3] = a3; xv[4] = a4; \
    yv[0] = b0; yv[1] = b1; yv[2] = b2; yv[3] = b3; yv[4] = b4; }
  if ((tm - T(1)) > thr && (tp - T(1)) > thr) {
    NV_SET(e, e, e, e, e, e, e, e, e, e);
  } else if ((tm - T(1)) > thr && (tp - T(1)) < -thr) {
    if (psi_p < T(-1)) {
      if (x1 < g2) {
        T nx = ((psi_m * psi_m) + psi_m) / (T(8) * ((psi_m * psi_m) + T(1)));
        T ny = T(1.0 / 8.0) * (T(1) + psi_m) - nx * psi_m;
        NV_SET(nx, nx, nx, nx, nx, ny, ny, ny, ny, ny);
      } else {
        T cx = (T(2) * g2 + x1) / T(3);
        T cy = (T(2) * g1 + y2) / T(3);
        NV_SET(g2, x1, g2, cx, cx, g1, g1, y2, cy, cy);
      }
    } else {
      if (y2 < g2) NV_SET(g2, g1, g2, g1, T(0), g1, g1, g2, g2, T(0))
      else NV_SET(g2, g1, g2, g1, x2, g1, g1, g2, y2, g2)
    }
  } else if ((tm - T(1)) < -thr && (tp - T(1)) > thr) {
    if (psi_p < T(-1)) {
      if (x1 < g2) NV_SET(g2, g1, g2, g1, T(0), g1, g1, g2, g2, T(0))
      else NV_SET(g1, g1, g2, x1, g2, g2, g1, g2, g1, y1)
    } else {
      if (y2 < g2) {
        T nx = ((psi_m * psi_m) + psi_m) / (T(8) * ((psi_m * psi_m) + T(1)));
        T ny = T(1.0 / 8.0) * (T(1) + psi_m) - nx * psi_m;
        NV_SET(nx, nx, nx, nx, nx, ny, ny, ny, ny, ny);
      } else {
        T cx = (T(2) * g1 + x1) / T(3);
        T cy = (T(2) * g2 + y2) / T(3);
        NV_SET(g1, x1, g1, cx, cx, g2, g2, y2, cy, cy);
      }
    }
  } else if (m_abs(tm - T(1)) < thr && (tp - T(1)) > thr) {
    NV_SET(e, e, e, e, e, t, e, e, t, m);
  } else if (m_abs(tp - T(1)) < thr && (tm - T(1)) > thr) {
    NV_SET(e, e, t, t, m, e, e, e, e, e);
  } else {
    NV_SET(e, e, t, t, m, e, t, t, e, m);
  }
#undef NV_SET
}

// features of a raw stencil: scale (neural_network.cc:300-313) and
// (tanh theta_m, tanh theta_p, |d0|, |d1|, |d2|) (neural_network.cc:22-38)
template <typename T>
__device__ __forceinline__ void nn_features(T u0, T u1, T u2, T u3, T st[5], T xv[5], T yv[5]) {
  T mv = T(1);
  mv = m_max(mv, m_abs(u0));
  mv = m_max(mv, m_abs(u1));
  mv = m_max(mv, m_abs(u2));
  mv = m_max(mv, m_abs(u3));
  T c0 = u0 / mv, c1 = u1 / mv, c2 = u2 / mv, c3 = u3 / mv;
  T tm = (c3 - c2) / (c2 - c1);
  T tp = (c1 - c0) / (c2 - c1);
  T d1 = m_abs(c1 - c0), d2 = m_abs(c2 - c1), d3 = m_abs(c3 - c2);
  T md = m_max(m_max(d1, d2), d3);
  st[0] = m_tanh(tm);
  st[1] = m_tanh(tp);
  st[2] = d1;
  st[3] = d2;
  st[4] = d3;
  nn_vertices(tm, tp, md, xv, yv);
}

// softmax (zero-initialised max) + vertex combination + compute_recon
template <typename T>
__device__ __forceinline__ void nn_head(const T lg[5], const T xv[5], const T yv[5], T u0, T u1,
                                        T u2, T u3, T& recl, T& recr) {
  T mx = T(0);
  for (int i = 0; i < 5; ++i) mx = m_max(mx, lg[i]);
  T s[5];
  T sum = T(0);
  for (int i = 0; i < 5; ++i) {
    s[i] = m_exp(lg[i] - mx);
    sum += s[i];
  }
  for (int i = 0; i < 5; ++i) s[i] /= sum;
  T C1 = T(0), C2 = T(0);
  for (int i = 0; i < 5; ++i) {
    C1 += s[i] * xv[i];
    C2 += s[i] * yv[i];
  }
  T w1 = T(1.0 / 4.0) - T(2) * C1;
  T tw0 = T(1.0 / 4.0) - T(2) * C2;
  recl = (T(1) - w1) * T(0.5) * (u1 + u2) + w1 * (T(-0.5) * u0 + T(1.5) * u1);
  recr = (T(1) - tw0) * T(0.5) * (u1 + u2) + tw0 * (T(-0.5) * u3 + T(1.5) * u2);
}

// SIMT network body: layers with ReLU on all but the last (neural_network.cc:40-60)
template <typename T, typename TW>
__device__ __forceinline__ void nn_body(const NetView<TW>& net, T st[TECNO_NET_MAXW]) {
  for (int l = 0; l < net.nl; ++l) {
    const int no = net.w[l + 1], ni = net.w[l];
    T tmp[TECNO_NET_MAXW];
    for (int j = 0; j < no; ++j) {
      T dot = T(0);
      const TW* Wr = net.W + (size_t)l * TECNO_NET_MAXW * TECNO_NET_MAXW + j * TECNO_NET_MAXW;
      for (int k = 0; k < ni; ++k) dot += T(Wr[k]) * st[k];
      tmp[j] = dot;
    }
    const TW* bl = net.b + (size_t)l * TECNO_NET_MAXW;
    for (int j = 0; j < no; ++j) st[j] = tmp[j] + T(bl[j]);
    if (l != net.nl - 1) {
      const int nr = no > 5 ? no : 5;
      for (int j = 0; j < nr; ++j) st[j] = m_max(T(0), st[j]);
    }
  }
}

// Fast path for the shipped 5-5-5-5-5 architecture (unrolled, no local arrays)
template <typename T, typename TW>
__device__ __forceinline__ void nn_body5(const NetView<TW>& net, T st[5]) {
#pragma unroll
  for (int l = 0; l < 4; ++l) {
    const TW* W = net.W + (size_t)l * TECNO_NET_MAXW * TECNO_NET_MAXW;
    const TW* b = net.b + (size_t)l * TECNO_NET_MAXW;
    T o[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      T dot = T(0);
#pragma unroll
      for (int k = 0; k < 5; ++k) dot += T(W[j * TECNO_NET_MAXW + k]) * st[k];
      o[j] = dot + T(b[j]);
    }
#pragma unroll
    for (int j = 0; j < 5; ++j) st[j] = (l != 3) ? m_max(T(0), o[j]) : o[j];
  }
}

template <typename T, typename TW>
__device__ __forceinline__ void nn_forward(const NetView<TW>& net, bool arch5, T u0, T u1, T u2,
                                           T u3, T& recl, T& recr) {
  T st[TECNO_NET_MAXW];
  T xv[5], yv[5];
  nn_features(u0, u1, u2, u3, st, xv, yv);
  if (arch5) {
    nn_body5(net, st);
  } else {
    for (int i = 5; i < TECNO_NET_MAXW; ++i) st[i] = T(0);
    nn_body(net, st);
  }
  nn_head(st, xv, yv, u0, u1, u2, u3, recl, recr);
}

// reconstruct_sp_weno_dl: zero jump (< 1e-15) gives (0, 0)
// (enflo/src/reconstruct.cc:138-157)
template <typename T, typename TW>
__device__ __forceinline__ void dsp_rec(const NetView<TW>& net, bool arch5, T u0, T u1, T u2,
                                        T u3, T& l, T& r) {
  if (m_abs(u2 - u1) < T(1e-15)) {
    l = T(0);
    r = T(0);
  } else {
    nn_forward(net, arch5, u0, u1, u2, u3, l, r);
  }
}

// ---------------------------------------------------------------- params
enum { REC_SPWENO = 0, REC_ENO = 1, REC_DSP = 2, REC_FIRST = 3, REC_SECOND = 4, REC_MINMOD = 5,
       REC_MINABS = 6 };

// limited_slope (reconstruct.cc:162-186) with minmod / minabs (limiter.h:20-50)
template <typename T>
__device__ __forceinline__ T limited_slope(T a, T b, bool use_minabs) {
  T m;
  if (use_minabs) {
    m = (m_abs(a) > m_abs(b)) ? b : a;
  } else {
    m = (a * b > T(0)) ? m_min(m_abs(a), m_abs(b)) * ((a < T(0)) ? T(-1) : T(1)) : T(0);
  }
  return T(0.5) * m;
}

// the two-cell-slope family, l = u_i + s_l, r = u_{i+1} - s_r:
//   first  (reconstruct.cc:13-24)  s = 0
//   second (reconstruct.cc:30-50)  unlimited MUSCL, kappa = 1/3; the
//          reference's KKK macro makes the weights the double constants
//          1 - 1/3 and 1 + 1/3
//   minmod / minabs (reconstruct.cc:55-80)  s = limited_slope
// One body for the four keeps the stage kernels that switch on the scheme
// at run time small.
template <typename T>
__device__ __forceinline__ void muscl(const T* st, int s, int kind, T& l, T& r) {
  const T u1 = st[s / 2 - 1], u2 = st[s / 2];
  T sl = T(0), sr = T(0);
  if (kind != REC_FIRST) {
    const T d = u2 - u1;
    const T dl = u1 - st[s / 2 - 2];
    const T dr = st[s / 2 + 1] - u2;
    if (kind == REC_SECOND) {
      sl = (dl * T(1.0 - 1.0 / 3.0) + d * T(1.0 + 1.0 / 3.0)) * T(0.25);
      sr = (dr * T(1.0 - 1.0 / 3.0) + d * T(1.0 + 1.0 / 3.0)) * T(0.25);
    } else {
      sl = limited_slope(dl, d, kind == REC_MINABS);
      sr = limited_slope(dr, d, kind == REC_MINABS);
    }
  }
  l = u1 + sl;
  r = u2 - sr;
}

template <typename T>
struct Phys {
  int system, model, flux, recon, corr, eno_k;
  int ec1_fix, fix_zmom, sign_fallback, arch5, mlp_mode, fix_edges, fix_visc;
  T gamma, R, gm1, inv_gm1;
  T vel[3];      // linear advection velocity
  T area[3];     // face areas (param.in spacings)
  T h[3];
  T hmin, cfl, vol;
  T s05, sgm1, gpgm;  // sqrt(0.5), sqrt(gamma-1), (gamma+1)/(gamma-1)
  T inv_gamma;
  double gamma_d, gm1_d, inv_gm1_d, R_d;
  double area_d[3];
  double s05_d, sgm1_d, inv_gamma_d;
  const unsigned short* mfma_w[2];  // per-lane packed weights: [0] bf16, [1] fp16
  double vol_d;       // dx*dy*dz (param.in spacings)
  double hmin_d, cfl_d;
  double vel_norm;    // |vel| (dt of linear advection)
  // Navier-Stokes (model 11 ns_kep, 12 ns_es): Parameter::Viscosity and Cp
  int mu_model;       // 0 constant, 1 sutherland, 2 power
  double mu_ref, T_ref, T_0, omega, prandtl, Cp;
  NetView<T> net;
  NetView<float> net32;  // FP32 weights, for the FP32 network of the FP64-state 3D face
};

// reconstruct one component on a stencil of 2*ng values.  RS selects the
// schemes compiled in: 0 all, 1 all but second/minmod/minabs, 2 only those
// three (kernels whose register budget is tight instantiate 1 and 2 apart).
template <int RS = 0, typename T>
__device__ __forceinline__ void recon_any(const Phys<T>& P, const T* st, int ng, T& l, T& r) {
  const int s = 2 * ng;
  if (RS == 2 || (RS == 0 && P.recon >= REC_SECOND)) {
    muscl(st, s, P.recon, l, r);
    return;
  }
  switch (P.recon) {
    case REC_ENO:
      eno_rec(P.eno_k, st, ng, l, r);
      break;
    case REC_DSP:
      dsp_rec(P.net, P.arch5 != 0, st[s / 2 - 2], st[s / 2 - 1], st[s / 2], st[s / 2 + 1], l, r);
      break;
    case REC_FIRST:
      l = st[s / 2 - 1];
      r = st[s / 2];
      break;
    default:
      sp_weno(st[s / 2 - 2], st[s / 2 - 1], st[s / 2], st[s / 2 + 1], P.corr, l, r);
      break;
  }
}

}  // namespace tecno
