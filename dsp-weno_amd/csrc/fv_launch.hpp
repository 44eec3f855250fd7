// fv_launch.hpp — host launchers and the small kernels around the stage
// kernel: time-step control, state conversion, audit sums, batched seams.
// Included by one translation unit per arithmetic type (kern_fp64.hip,
// kern_fp32.hip), which instantiate Launch<T>.
#pragma once
#include "fv_kernels.hpp"
#include "fv_mixed.hpp"
#include "fv_euler3.hpp"

namespace tecno {

struct LaunchInfo {
  int system;     // 0 scalar, 1 euler
  int dims;       // 1, 2, 3
  int G;          // halo width (2 or 3)
  int fast_flux;  // FV_* id for the 2D Euler kernel
  int fast_rec;   // RV_* id
  int state_fp64; // state held in HBM as double (1) or float (0)
  int mlp_mode;   // MM_* of the mixed-precision DSP-WENO face
  int scalar_mixed;  // scalar, FP64 state, FP32 SP-WENO jump (ScalarMixedPh)
  int mix3;          // M3_* reconstruction of the 3D face (M3_NONE: FP64)
  int mix2;          // M3_* of the 2D generic TeCNO face on the FP64 state (0: FP64)
  GridDev grid;
  hipStream_t stream;
  Ctl* ctl;
  void* R2;       // classical RK4 residual accumulator (state type, NV x N) or null
};

// ---------------------------------------------------------------- control
hipError_t launch_dt(const LaunchInfo& L, double hmin, double cfl, int const_speed, double speed);
hipError_t launch_finish(const LaunchInfo& L);

#ifdef TECNO_DEFINE_CONTROL
// dt of the next step (FV::compute_dt, enflo/src/fv.cc:366-383, and the
// output-stamp clamp of FV::run, fv.cc:1262-1266)
__global__ void k_dt(Ctl* c, double hmin, double cfl, int const_speed, double speed) {
  if (!(c->time < c->final_time && c->iter < c->max_iter)) {
    c->active = 0;
    c->dt = 0.0;
    return;
  }
  c->active = 1;
  double smax = const_speed ? speed : __longlong_as_double((long long)c->smax_bits);
  double dt = 1.0e20;
  double q = hmin / smax;
  dt = (q < dt) ? q : dt;
  dt *= cfl;
  if (c->time + dt > c->t_save) dt = c->t_save - c->time;
  c->dt = dt;
  c->smax_bits = 0ull;
}

__global__ void k_finish(Ctl* c) {
  if (!c->active) return;
  c->time += c->dt;
  c->iter += 1;
  if (c->time == c->t_save) c->t_save += c->t_int;
}

hipError_t launch_dt(const LaunchInfo& L, double hmin, double cfl, int const_speed, double speed) {
  hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, L.stream, L.ctl, hmin, cfl, const_speed, speed);
  return hipGetLastError();
}
hipError_t launch_finish(const LaunchInfo& L) {
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(1), 0, L.stream, L.ctl);
  return hipGetLastError();
}
#endif

// ---------------------------------------------------------------- conversions
// prim_to_con (enflo/src/fv.cc:405-416); scalar: copy
template <typename T>
__global__ void k_prim_to_con(int system, int dims, int64_t N, const double* prim, T* U,
                              double gamma) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (system == 0) {
    U[i] = T(prim[i]);
    return;
  }
  const int nv = dims == 3 ? 5 : 4;  // primitive: rho, u, v, [w,] p
  double rho = prim[i], p = prim[(nv - 1) * N + i];
  double u[3] = {prim[N + i], prim[2 * N + i], dims == 3 ? prim[3 * N + i] : 0.0};
  U[i] = T(rho);
  for (int d = 0; d < nv - 2; ++d) U[(1 + d) * N + i] = T(u[d] * rho);
  double q = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  U[(nv - 1) * N + i] = T(0.5 * rho * q + p / (gamma - 1.0));
}

// con_to_prim (enflo/src/fv.cc:388-400)
template <typename T>
__global__ void k_con_to_prim(int system, int dims, int64_t N, const T* U, double* prim,
                              double gamma) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  if (system == 0) {
    prim[i] = double(U[i]);
    return;
  }
  const int nv = dims == 3 ? 5 : 4;
  T rho = U[i];
  T inv = T(1) / rho;
  T m2 = T(0);
  for (int d = 0; d < nv - 2; ++d) {
    T m = U[(1 + d) * N + i];
    prim[(1 + d) * N + i] = double(m * inv);
    m2 += m * m;
  }
  prim[i] = double(rho);
  T p = T(gamma - 1.0) * (U[(nv - 1) * N + i] - T(0.5) * m2 / rho);
  prim[(nv - 1) * N + i] = double(p);
}

// max wave speed of a state (first step's dt)
template <class PH>
__global__ void k_speed(Phys<typename PH::TP> P, int64_t N, const typename PH::TS* U, Ctl* c) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x)
    s = fmax(s, (double)PH::speed(P, U, N, i));
  for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomic_max_pos(&c->smax_bits, m);
  }
}

// FP64 audit sums (compute_global, enflo/src/compute_global.cc:10-89):
// out[0] sum -rho*s/(gamma-1) (scalar: sum u^2/2), out[1] sum rho|u|^2,
// out[2..6] sums of the conserved variables, out[7] max speed, out[8] #bad.
template <typename T>
__global__ void k_audit(int system, int dims, int64_t N, const T* U, double gamma,
                        double* part) {
  __shared__ double red[4][9];
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    if (system == 0) {
      double u = (double)U[i];
      acc[0] += 0.5 * u * u;
      acc[2] += u;
      acc[7] = fmax(acc[7], fabs(u));
      if (!(u == u)) acc[8] += 1;
    } else {
      const int nv = dims == 3 ? 5 : 4;
      double rho = U[i], m[3] = {U[N + i], U[2 * N + i], dims == 3 ? (double)U[3 * N + i] : 0.0};
      double E = U[(nv - 1) * N + i];
      double inv = 1.0 / rho;
      double u[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
      double q = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
      double p = (gamma - 1.0) * (E - 0.5 * (m[0] * m[0] + m[1] * m[1] + m[2] * m[2]) / rho);
      double s = log(p) - gamma * log(rho);
      acc[0] += -rho * s / (gamma - 1.0);
      acc[1] += rho * q;
      acc[2] += rho;
      acc[3] += m[0];
      acc[4] += m[1];
      acc[5] += m[2];
      acc[6] += E;
      acc[7] = fmax(acc[7], sqrt(q) + sqrt(gamma * p / rho));
      if (!(rho > 0.0) || !(p > 0.0) || !(p < INFINITY)) acc[8] += 1;
    }
  }
  for (int v = 0; v < 9; ++v) {
    double x = acc[v];
    for (int off = 32; off > 0; off >>= 1) {
      double y = __shfl_xor(x, off);
      x = (v == 7) ? fmax(x, y) : x + y;
    }
    acc[v] = x;
  }
  if ((threadIdx.x & 63) == 0)
    for (int v = 0; v < 9; ++v) red[threadIdx.x >> 6][v] = acc[v];
  __syncthreads();
  if (threadIdx.x < 9) {
    int v = threadIdx.x;
    double x = red[0][v];
    for (int w = 1; w < 4; ++w) x = (v == 7) ? fmax(x, red[w][v]) : x + red[w][v];
    part[(int64_t)blockIdx.x * 9 + v] = x;
  }
}

// L1 error (compute_error, enflo/src/compute_errors.cc:9-34): per-block sums
// of |exact - prim| for each primitive variable
template <typename T>
__global__ void k_l1(int system, int dims, int64_t N, const T* U, const double* exact,
                     double gamma, double* part) {
  __shared__ double red[4][5];
  const int nv = system == 0 ? 1 : (dims == 3 ? 5 : 4);
  double acc[5] = {0, 0, 0, 0, 0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    double pr[5];
    if (system == 0) {
      pr[0] = U[i];
    } else {
      double rho = U[i];
      double inv = 1.0 / rho, m2 = 0.0;
      for (int d = 0; d < nv - 2; ++d) {
        double m = U[(1 + d) * N + i];
        pr[1 + d] = m * inv;
        m2 += m * m;
      }
      pr[0] = rho;
      pr[nv - 1] = (gamma - 1.0) * (U[(nv - 1) * N + i] - 0.5 * m2 / rho);
    }
    for (int v = 0; v < nv; ++v) acc[v] += fabs(exact[v * N + i] - pr[v]);
  }
  for (int v = 0; v < 5; ++v) {
    double x = acc[v];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    acc[v] = x;
  }
  if ((threadIdx.x & 63) == 0)
    for (int v = 0; v < 5; ++v) red[threadIdx.x >> 6][v] = acc[v];
  __syncthreads();
  if (threadIdx.x < 5) {
    int v = threadIdx.x;
    part[(int64_t)blockIdx.x * 5 + v] = red[0][v] + red[1][v] + red[2][v] + red[3][v];
  }
}

// ---------------------------------------------------------------- seams
template <typename T>
__global__ void k_recon_batch(Phys<T> P, const double* st, int ng, int64_t n, double* l,
                              double* r) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T s[6];
  for (int k = 0; k < 2 * ng; ++k) s[k] = T(st[i * 2 * ng + k]);
  T a, b;
  recon_any(P, s, ng, a, b);
  l[i] = a;
  r[i] = b;
}

// Neural_Network::forward on raw stencils.  FP64: the reference arithmetic.
// FP32: jump form -- cell differences in FP64, features, network, softmax and
// vertex weights in FP32, the reconstructed values assembled in FP64 from
// l = u1 + (1-w1) D1/2 + w1 D0/2 and r = u2 - (1-w~0) D1/2 - w~0 D2/2.
template <typename T>
__global__ void k_mlp_batch(Phys<T> P, const double* u4, int64_t n, double* l, double* r) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if constexpr (sizeof(T) == 4) {
    const double u0 = u4[4 * i], u1 = u4[4 * i + 1], u2 = u4[4 * i + 2], u3 = u4[4 * i + 3];
    const float d0 = (float)(u1 - u0), d1 = (float)(u2 - u1), d2 = (float)(u3 - u2);
    const float mx = (float)fmax(fmax(fabs(u0), fabs(u1)), fmax(fabs(u2), fabs(u3)));
    float st[5], xv[5], yv[5], C1, C2;
    fast::features(d0, d1, d2, mx, st, xv, yv);
    fast::body5(P.net, st);
    fast::head_c(st, xv, yv, C1, C2);
    const double w1 = 0.25 - 2.0 * (double)C1, tw0 = 0.25 - 2.0 * (double)C2;
    l[i] = u1 + 0.5 * (1.0 - w1) * (u2 - u1) + 0.5 * w1 * (u1 - u0);
    r[i] = u2 - 0.5 * (1.0 - tw0) * (u2 - u1) - 0.5 * tw0 * (u3 - u2);
  } else {
    T a, b;
    nn_forward(P.net, P.arch5 != 0, T(u4[4 * i]), T(u4[4 * i + 1]), T(u4[4 * i + 2]),
               T(u4[4 * i + 3]), a, b);
    l[i] = a;
    r[i] = b;
  }
}

// FV::num_flux on independent faces.  Scalar: P = n x 2ng.  2D Euler: P, V
// = n x 2ng x 5 (w and Vz must be zero), F = n x 5.  The float instantiation
// evaluates the mixed-precision face (FP32 flux, FP64 entropy variables).
template <typename T, int G>
__global__ void k_flux_batch(Phys<T> P, int system, const double* Pin, const double* Vin,
                             int ng, int64_t n, int ax, double* F) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int S = 2 * G;
  if (system == 0) {
    T loc[S];
    for (int k = 0; k < S; ++k) loc[k] = T(Pin[i * 2 * ng + k]);
    T f[1];
    unsigned nfb = 0;
    ScalarPh<T, G>::face(P, loc, S, 0, 1, ax, f, nfb);
    F[i] = f[0];
    return;
  }
  double f[4];
  if constexpr (sizeof(T) == 4 && G == 2) {
    using PH = Euler2Mixed<RV_GENERIC, MM_SIMT>;
    double ld[R_N * S];
    for (int k = 0; k < S; ++k) {
      const double* q = Pin + (i * S + k) * 5;
      const double* v = Vin + (i * S + k) * 5;
      const double beta = 0.5 * q[0] / q[4];
      ld[R_RHO * S + k] = q[0];
      ld[R_UX * S + k] = q[1];
      ld[R_UY * S + k] = q[2];
      ld[R_BETA * S + k] = beta;
      ld[R_LRHO * S + k] = log(q[0]);
      ld[R_LBETA * S + k] = log(beta);
      ld[R_VF * S + k] = v[0];
    }
    unsigned nfb = 0;
    PH::face(P, ld, S, 0, 1, ax, f, nfb);
  } else {
    T loc[E_NREC * S];
    for (int k = 0; k < S; ++k) {
      const double* q = Pin + (i * S + k) * 5;
      const double* v = Vin + (i * S + k) * 5;
      T rho = q[0], p = q[4];
      loc[E_RHO * S + k] = rho;
      loc[E_UX * S + k] = T(q[1]);
      loc[E_UY * S + k] = T(q[2]);
      loc[E_P * S + k] = p;
      loc[E_BK * S + k] = T(0.5) / (P.R * (p / (P.R * rho)));
      loc[E_VF * S + k] = T(v[0]);
      loc[E_VX * S + k] = T(v[1]);
      loc[E_VY * S + k] = T(v[2]);
      loc[E_VL * S + k] = T(v[4]);
    }
    T ft[4];
    unsigned nfb = 0;
    Euler2Ph<T, G, FV_GENERIC, RV_GENERIC>::face(P, loc, S, 0, 1, ax, ft, nfb);
    for (int v = 0; v < 4; ++v) f[v] = ft[v];
  }
  F[5 * i + 0] = f[0];
  F[5 * i + 1] = f[1];
  F[5 * i + 2] = f[2];
  F[5 * i + 3] = 0.0;
  F[5 * i + 4] = f[3];
}

// ---------------------------------------------------------------- launchers
template <typename T>
struct Launch {
  static hipError_t stage(const LaunchInfo& L, const Phys<T>& P, const void* U, const void* U0,
                          void* Uout, const StageCoef& k);
  static hipError_t speed(const LaunchInfo& L, const Phys<T>& P, const void* U);
  static hipError_t recon_batch(const LaunchInfo& L, const Phys<T>& P, const double* st, int ng,
                                int64_t n, double* l, double* r);
  static hipError_t mlp_batch(const LaunchInfo& L, const Phys<T>& P, const double* u4, int64_t n,
                              double* l, double* r);
  static hipError_t flux_batch(const LaunchInfo& L, const Phys<T>& P, const double* Pin,
                               const double* Vin, int ng, int64_t n, int ax, double* F);
};

// kernels on the state, instantiated for the state type
template <typename TS>
struct StateOps {
  static hipError_t prim_to_con(const LaunchInfo& L, const double* prim, void* U, double gamma) {
    int nblk = (int)((L.grid.N + 255) / 256);
    hipLaunchKernelGGL((k_prim_to_con<TS>), dim3(nblk), dim3(256), 0, L.stream, L.system, L.dims,
                       L.grid.N, prim, (TS*)U, gamma);
    return hipGetLastError();
  }
  static hipError_t con_to_prim(const LaunchInfo& L, const void* U, double* prim, double gamma) {
    int nblk = (int)((L.grid.N + 255) / 256);
    hipLaunchKernelGGL((k_con_to_prim<TS>), dim3(nblk), dim3(256), 0, L.stream, L.system, L.dims,
                       L.grid.N, (const TS*)U, prim, gamma);
    return hipGetLastError();
  }
  static hipError_t audit(const LaunchInfo& L, const void* U, double gamma, double* part,
                          int nblk) {
    hipLaunchKernelGGL((k_audit<TS>), dim3(nblk), dim3(256), 0, L.stream, L.system, L.dims,
                       L.grid.N, (const TS*)U, gamma, part);
    return hipGetLastError();
  }
  static hipError_t l1(const LaunchInfo& L, const void* U, const double* exact, double gamma,
                       double* part, int nblk) {
    hipLaunchKernelGGL((k_l1<TS>), dim3(nblk), dim3(256), 0, L.stream, L.system, L.dims, L.grid.N,
                       (const TS*)U, exact, gamma, part);
    return hipGetLastError();
  }
};

#define TECNO_TILE2_X 32
#define TECNO_TILE2_Y 15
#define TECNO_TILE1_X 255
// tile of the mixed-precision 2D Euler stage (records + pair fluxes in LDS)
// tile of the FP64 2D Euler stage with kepes_tecno4_roe (records + pair
// fluxes in LDS: 44.7 KB at 16x15, three workgroups per CU)
#ifndef TECNO_TILE2D_X
#define TECNO_TILE2D_X 16
#endif
#ifndef TECNO_TILE2D_Y
#define TECNO_TILE2D_Y 15
#endif
#ifndef TECNO_TILE2M_X
#define TECNO_TILE2M_X 16
#endif
#ifndef TECNO_TILE2M_Y
#define TECNO_TILE2M_Y 15
#endif

template <class PH, int DIMS, int TX, int TY, int G>
static hipError_t launch_stage_t(const LaunchInfo& L, const Phys<typename PH::TP>& P,
                                 const void* U, const void* U0, void* Uo, const StageCoef& k) {
  using TS = typename PH::TS;
  StageArgs<TS> A{(const TS*)U, (const TS*)U0, (TS*)Uo, L.ctl, k.a, k.b, k.c, k.last, (TS*)L.R2, k.w, k.r2};
  dim3 grid((L.grid.n[0] + TX - 1) / TX, DIMS >= 2 ? (L.grid.n[1] + TY - 1) / TY : 1, 1);
  hipLaunchKernelGGL((stage_kernel<PH, DIMS, TX, TY, G>), grid, dim3(256), 0, L.stream, A, P,
                     L.grid);
  return hipGetLastError();
}

template <class PH, int G, bool VISC>
static hipError_t launch_stage3_t(const LaunchInfo& L, const Phys<typename PH::TP>& P,
                                  const void* U, const void* U0, void* Uo, const StageCoef& k) {
  using TS = typename PH::TS;
  constexpr int TX = 8, TY = 8, TZ = 2;
  StageArgs<TS> A{(const TS*)U, (const TS*)U0, (TS*)Uo, L.ctl, k.a, k.b, k.c, k.last, (TS*)L.R2, k.w, k.r2};
  dim3 grid((L.grid.n[0] + TX - 1) / TX, (L.grid.n[1] + TY - 1) / TY, (L.grid.n[2] + TZ - 1) / TZ);
  hipLaunchKernelGGL((stage3_kernel<PH, TX, TY, TZ, G, VISC>), grid, dim3(256), 0, L.stream, A, P,
                     L.grid);
  return hipGetLastError();
}

template <class PH>
static hipError_t launch_speed_t(const LaunchInfo& L, const Phys<typename PH::TP>& P,
                                 const void* U) {
  int nblk = (int)((L.grid.N + 255) / 256);
  if (nblk > 4096) nblk = 4096;
  hipLaunchKernelGGL((k_speed<PH>), dim3(nblk), dim3(256), 0, L.stream, P, L.grid.N,
                     (const typename PH::TS*)U, L.ctl);
  return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::stage(const LaunchInfo& L, const Phys<T>& P, const void* U, const void* U0,
                            void* Uo, const StageCoef& k) {
  constexpr int X2 = TECNO_TILE2_X, Y2 = TECNO_TILE2_Y, X1 = TECNO_TILE1_X;
  if (L.system == 0) {
    if constexpr (sizeof(T) == 8) {
      if (L.scalar_mixed) {
        if (L.G != 2) return hipErrorNotSupported;
        if (L.dims == 1) return launch_stage_t<ScalarMixedPh<2>, 1, X1, 1, 2>(L, P, U, U0, Uo, k);
        if (L.dims == 2) return launch_stage_t<ScalarMixedPh<2>, 2, X2, Y2, 2>(L, P, U, U0, Uo, k);
        return hipErrorNotSupported;
      }
    }
    // the SP-WENO family and ENO run the kernel with that reconstruction only
    const bool spw = P.recon == REC_SPWENO, eno = P.recon == REC_ENO;
    if (L.dims == 1)
      return L.G == 3 ? (eno ? launch_stage_t<ScalarPh<T, 3, 2>, 1, X1, 1, 3>(L, P, U, U0, Uo, k)
                             : launch_stage_t<ScalarPh<T, 3>, 1, X1, 1, 3>(L, P, U, U0, Uo, k))
             : spw    ? launch_stage_t<ScalarPh<T, 2, 1>, 1, X1, 1, 2>(L, P, U, U0, Uo, k)
             : eno    ? launch_stage_t<ScalarPh<T, 2, 2>, 1, X1, 1, 2>(L, P, U, U0, Uo, k)
                      : launch_stage_t<ScalarPh<T, 2>, 1, X1, 1, 2>(L, P, U, U0, Uo, k);
    if (L.dims == 2)
      return L.G == 3 ? (eno ? launch_stage_t<ScalarPh<T, 3, 2>, 2, X2, Y2, 3>(L, P, U, U0, Uo, k)
                             : launch_stage_t<ScalarPh<T, 3>, 2, X2, Y2, 3>(L, P, U, U0, Uo, k))
             : spw    ? launch_stage_t<ScalarPh<T, 2, 1>, 2, X2, Y2, 2>(L, P, U, U0, Uo, k)
             : eno    ? launch_stage_t<ScalarPh<T, 2, 2>, 2, X2, Y2, 2>(L, P, U, U0, Uo, k)
                      : launch_stage_t<ScalarPh<T, 2>, 2, X2, Y2, 2>(L, P, U, U0, Uo, k);
    return hipErrorNotSupported;
  }
  if (L.dims == 3) {
    // 3D Euler / Navier-Stokes: FP64 reference arithmetic only
    if constexpr (sizeof(T) == 8) {
      const bool visc = P.model == 11 || P.model == 12;
      if (L.G != 2) return hipErrorNotSupported;
      switch (L.mix3) {
        case M3_SPW:
          if (visc) return launch_stage3_t<Euler3Ph<T, 2, 0, M3_SPW>, 2, true>(L, P, U, U0, Uo, k);
          return launch_stage3_t<Euler3Ph<T, 2, 0, M3_SPW>, 2, false>(L, P, U, U0, Uo, k);
        case M3_DSP_SIMT:
          if (visc) return launch_stage3_t<Euler3Ph<T, 2, 0, M3_DSP_SIMT>, 2, true>(L, P, U, U0, Uo, k);
          return launch_stage3_t<Euler3Ph<T, 2, 0, M3_DSP_SIMT>, 2, false>(L, P, U, U0, Uo, k);
        case M3_DSP_BF16:
          if (visc) return launch_stage3_t<Euler3Ph<T, 2, 0, M3_DSP_BF16>, 2, true>(L, P, U, U0, Uo, k);
          return launch_stage3_t<Euler3Ph<T, 2, 0, M3_DSP_BF16>, 2, false>(L, P, U, U0, Uo, k);
        case M3_DSP_FP16:
          if (visc) return launch_stage3_t<Euler3Ph<T, 2, 0, M3_DSP_FP16>, 2, true>(L, P, U, U0, Uo, k);
          return launch_stage3_t<Euler3Ph<T, 2, 0, M3_DSP_FP16>, 2, false>(L, P, U, U0, Uo, k);
        default:
          break;
      }
      if (P.flux >= 125) {
        if (visc) return launch_stage3_t<Euler3Ph<T, 2, 1>, 2, true>(L, P, U, U0, Uo, k);
        return launch_stage3_t<Euler3Ph<T, 2, 1>, 2, false>(L, P, U, U0, Uo, k);
      }
      if (P.recon >= REC_SECOND) {
        if (visc) return launch_stage3_t<Euler3Ph<T, 2, 2>, 2, true>(L, P, U, U0, Uo, k);
        return launch_stage3_t<Euler3Ph<T, 2, 2>, 2, false>(L, P, U, U0, Uo, k);
      }
      if (visc) return launch_stage3_t<Euler3Ph<T, 2>, 2, true>(L, P, U, U0, Uo, k);
      return launch_stage3_t<Euler3Ph<T, 2>, 2, false>(L, P, U, U0, Uo, k);
    }
    return hipErrorNotSupported;
  }
  if (L.dims != 2) return hipErrorNotSupported;
  if constexpr (sizeof(T) == 4) {
    // mixed precision: kepes_tecno4_roe with sp_weno / sp_wenoc or sp_weno_dl
    if (L.fast_flux != FV_KEPES4_ROE || L.G != 2 || L.fast_rec == RV_GENERIC)
      return hipErrorNotSupported;
    constexpr int XM = TECNO_TILE2M_X, YM = TECNO_TILE2M_Y;
    if (L.fast_rec == RV_SPWENO)
      return launch_stage_t<Euler2Mixed<RV_SPWENO, MM_SIMT>, 2, XM, YM, 2>(L, P, U, U0, Uo, k);
    if (L.mlp_mode == MM_BF16)
      return launch_stage_t<Euler2Mixed<RV_DSP, MM_BF16>, 2, XM, YM, 2>(L, P, U, U0, Uo, k);
    if (L.mlp_mode == MM_FP16)
      return launch_stage_t<Euler2Mixed<RV_DSP, MM_FP16>, 2, XM, YM, 2>(L, P, U, U0, Uo, k);
    return launch_stage_t<Euler2Mixed<RV_DSP, MM_SIMT>, 2, XM, YM, 2>(L, P, U, U0, Uo, k);
  } else {
    switch (L.mix2) {
      case M3_SPW:
        return launch_stage_t<Euler2Ph<T, 2, FV_GENERIC, RV_GENERIC, M3_SPW>, 2, X2, Y2, 2>(
            L, P, U, U0, Uo, k);
      case M3_DSP_SIMT:
        return launch_stage_t<Euler2Ph<T, 2, FV_GENERIC, RV_GENERIC, M3_DSP_SIMT>, 2, X2, Y2, 2>(
            L, P, U, U0, Uo, k);
      case M3_DSP_BF16:
        return launch_stage_t<Euler2Ph<T, 2, FV_GENERIC, RV_GENERIC, M3_DSP_BF16>, 2, X2, Y2, 2>(
            L, P, U, U0, Uo, k);
      case M3_DSP_FP16:
        return launch_stage_t<Euler2Ph<T, 2, FV_GENERIC, RV_GENERIC, M3_DSP_FP16>, 2, X2, Y2, 2>(
            L, P, U, U0, Uo, k);
      default:
        break;
    }
    // ENO k = 3 takes 3 ghost layers; ENO alone in the face as below
    if (L.G == 3 && P.recon == REC_ENO)
      return launch_stage_t<Euler2Ph<T, 3, FV_GENERIC, RV_ENO>, 2, X2, Y2, 3>(L, P, U, U0, Uo, k);
    if (L.G == 3)
      return launch_stage_t<Euler2Ph<T, 3, FV_GENERIC, RV_GENERIC>, 2, X2, Y2, 3>(L, P, U, U0, Uo,
                                                                                 k);
    constexpr int XD = TECNO_TILE2D_X, YD = TECNO_TILE2D_Y;
    if (L.fast_flux == FV_KEPES4_ROE && L.fast_rec == RV_SPWENO)
      return launch_stage_t<Euler2Ph<T, 2, FV_KEPES4_ROE, RV_SPWENO>, 2, XD, YD, 2>(L, P, U, U0,
                                                                                    Uo, k);
    if (L.fast_flux == FV_KEPES4_ROE && L.fast_rec == RV_DSP)
      return launch_stage_t<Euler2Ph<T, 2, FV_KEPES4_ROE, RV_DSP>, 2, XD, YD, 2>(L, P, U, U0, Uo,
                                                                                 k);
    // the other TeCNO fluxes with the SP-WENO family: the generic face built
    // with that reconstruction only (no ENO, network or MUSCL code, far fewer
    // registers), same arithmetic as the general kernel's SP-WENO branch
    if (L.fast_rec == RV_SPWENO)
      return launch_stage_t<Euler2Ph<T, 2, FV_GENERIC, RV_SPWENO>, 2, X2, Y2, 2>(L, P, U, U0, Uo,
                                                                                k);
    if (L.fast_rec == RV_DSP)
      return launch_stage_t<Euler2Ph<T, 2, FV_GENERIC, RV_DSP>, 2, X2, Y2, 2>(L, P, U, U0, Uo, k);
    if (P.recon == REC_ENO)
      return launch_stage_t<Euler2Ph<T, 2, FV_GENERIC, RV_ENO>, 2, X2, Y2, 2>(L, P, U, U0, Uo, k);
    return launch_stage_t<Euler2Ph<T, 2, FV_GENERIC, RV_GENERIC>, 2, X2, Y2, 2>(L, P, U, U0, Uo,
                                                                               k);
  }
}

template <typename T>
hipError_t Launch<T>::speed(const LaunchInfo& L, const Phys<T>& P, const void* U) {
  if (L.system == 0) return launch_speed_t<ScalarPh<T, 2>>(L, P, U);
  if (L.dims == 3) {
    if constexpr (sizeof(T) == 8) return launch_speed_t<Euler3Ph<T, 2>>(L, P, U);
    return hipErrorNotSupported;
  }
  if constexpr (sizeof(T) == 4)
    return launch_speed_t<Euler2Mixed<RV_SPWENO, MM_SIMT>>(L, P, U);
  else
    return launch_speed_t<Euler2Ph<T, 2, FV_GENERIC, RV_GENERIC>>(L, P, U);
}

template <typename T>
hipError_t Launch<T>::recon_batch(const LaunchInfo& L, const Phys<T>& P, const double* st, int ng,
                                  int64_t n, double* l, double* r) {
  hipLaunchKernelGGL((k_recon_batch<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     L.stream, P, st, ng, n, l, r);
  return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::mlp_batch(const LaunchInfo& L, const Phys<T>& P, const double* u4, int64_t n,
                                double* l, double* r) {
  hipLaunchKernelGGL((k_mlp_batch<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, L.stream,
                     P, u4, n, l, r);
  return hipGetLastError();
}

template <typename T>
hipError_t Launch<T>::flux_batch(const LaunchInfo& L, const Phys<T>& P, const double* Pin,
                                 const double* Vin, int ng, int64_t n, int ax, double* F) {
  dim3 gr((unsigned)((n + 255) / 256));
  if (ng == 3)
    hipLaunchKernelGGL((k_flux_batch<T, 3>), gr, dim3(256), 0, L.stream, P, L.system, Pin, Vin,
                       ng, n, ax, F);
  else
    hipLaunchKernelGGL((k_flux_batch<T, 2>), gr, dim3(256), 0, L.stream, P, L.system, Pin, Vin,
                       ng, n, ax, F);
  return hipGetLastError();
}

}  // namespace tecno
