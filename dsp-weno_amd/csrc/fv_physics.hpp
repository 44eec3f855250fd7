// fv_physics.hpp — per-cell records and per-face fluxes of the scalar and 2D
// Euler TeCNO schemes (device functions used by the stage kernels and the
// batched seams).
#pragma once
#include "fv_device.hpp"

namespace tecno {

// ================================================================ physics
// -------- scalar (enflo_scalar): record = u
// RF 1: the face reconstructs with the SP-WENO family only (sp_weno,
// sp_wenoc, sp_weno_corr2: P.corr), for problems that use it.  Without the
// other reconstructions compiled in -- the FP64 network above all -- the FP64
// 2D kernels need 88-92 VGPRs instead of 161-166 and run five waves per SIMD
// instead of three, which their chains of FP64 divisions need.  RF 2: ENO
// (k = 2, 3) only, in the same way.
template <typename T, int G, int RF = 0>
struct ScalarPh {
  using TS = T;  // state in HBM
  static constexpr int MIN_WAVES = 1;
  using TF = T;  // face flux
  using TP = T;  // parameters
  static constexpr int NV = 1;
  static constexpr size_t rec_bytes(int NB) { return (size_t)NB * sizeof(T); }
  __device__ static void load(const Phys<T>& P, const T* U, int64_t N, int64_t idx, int flipx,
                              int flipy, void* lds, int NB, int c) {
    T u = U[idx];
    ((T*)lds)[c] = (flipx || flipy) ? -u : u;  // scalar wall: ghost = -mirror
  }
  // entropy-conservative part of the face flux: enflo_scalar num_flux.cc:12-57
  // and the lin_/bur_ flux files
  __device__ static T ec_flux(const Phys<T>& P, const T* s, int ax) {
    const T ul = s[G - 1], ur = s[G];
    const T vel = P.vel[ax];
    T f = T(0), f0;
    switch (P.flux) {
      case 0: case 2:  // lin_ec, lin_tecno
        f = T(0.5) * vel * (ul + ur);
        break;
      case 1: case 3:  // lin_ec4, lin_tecno4: flux0*4.0/3.0 == (flux0*4)/3
        f0 = T(0.5) * vel * (s[G - 1] + s[G]);
        f = T(0) + f0 * T(4) / T(3);
        f0 = T(0.5) * vel * (s[G - 2] + s[G]);
        f += f0 * T(-1.0 / 6.0);
        f0 = T(0.5) * vel * (s[G - 1] + s[G + 1]);
        f += f0 * T(-1.0 / 6.0);
        break;
      case 4: case 6:  // bur_ec, bur_tecno
        f = (ul * ul + ur * ur + ul * ur) / T(6);
        break;
      default: {       // bur_ec4, bur_tecno4
        T a = s[G - 1], b = s[G];
        f0 = (a * a + b * b + a * b) / T(6);
        f = T(0) + f0 * T(4.0 / 3.0);
        a = s[G - 2]; b = s[G];
        f0 = (a * a + b * b + a * b) / T(6);
        f += f0 * T(-1.0 / 6.0);
        a = s[G - 1]; b = s[G + 1];
        f0 = (a * a + b * b + a * b) / T(6);
        f += f0 * T(-1.0 / 6.0);
        break;
      }
    }
    return f;
  }
  // TeCNO fluxes: the dissipation coefficient of f - D/2 <<u>>
  __device__ static bool tecno(const Phys<T>& P) {
    return P.flux == 2 || P.flux == 3 || P.flux == 6 || P.flux == 7;
  }
  __device__ static T diss_coef(const Phys<T>& P, const T* s, int ax) {
    return (P.flux <= 3) ? m_abs(T(P.vel[ax])) : T(0.5) * (m_abs(s[G - 1]) + m_abs(s[G]));
  }
  __device__ static void face(const Phys<T>& P, const void* ldsv, int NB, int base, int stride,
                              int ax, T F[1], unsigned& nfb) {
    const T* lds = (const T*)ldsv;
    T s[2 * G];
#pragma unroll
    for (int l = 0; l < 2 * G; ++l) s[l] = lds[base + l * stride];
    T f = ec_flux(P, s, ax);
    if (tecno(P)) {
      const T D = diss_coef(P, s, ax);
      T uij, uji;
      if constexpr (RF == 1)
        sp_weno(s[G - 2], s[G - 1], s[G], s[G + 1], P.corr, uij, uji);
      else if constexpr (RF == 2)
        eno_rec(P.eno_k, s, G, uij, uji);
      else
        recon_any(P, s, G, uij, uji);
      f -= T(0.5) * D * (uji - uij);
    }
    F[0] = f * P.area[ax];
  }
  __device__ static T speed(const Phys<T>& P, const T* U, int64_t N, int64_t idx) {
    return (P.model == 0) ? T(P.vel_norm) : m_abs(U[idx]);
  }
  __device__ static bool bad(const T* U, int64_t N, int64_t idx) {
    T u = U[idx];
    return !(u == u) || m_abs(u) == T(INFINITY);
  }
};

// -------- 2D Euler (enflo with w = 0): conserved (rho, mx, my, E); record
// (rho, ux, uy, p, beta_k, Vf, Vx, Vy, Vl), beta_k = 0.5/(R*(p/(R*rho))) as
// kepec_flux computes it (enflo/src/kepec.cc:21-25).
enum { E_RHO = 0, E_UX, E_UY, E_P, E_BK, E_VF, E_VX, E_VY, E_VL, E_NREC };

// FAST: the FP32 reconstruction path's records (refined reciprocal, log64,
// beta_k = beta), a few ulp from the reference's form that the FP64 path keeps
template <bool FAST = false, typename T>
__device__ __forceinline__ void euler2_record(const Phys<T>& P, T rho, T mx, T my, T E, int flipx,
                                              int flipy, T* rec) {
  // con_to_prim: enflo/src/fv.cc:388-400
  T inv = qdiv<FAST>(T(1), rho);
  T ux = mx * inv, uy = my * inv;
  T p = P.gm1 * (E - (FAST ? T(0.5) * (mx * mx + my * my) * inv
                           : T(0.5) * (mx * mx + my * my) / rho));
  if (flipx) ux = -ux;  // wall ghost: mirrored normal velocity (fv.cc:134-140)
  if (flipy) uy = -uy;
  // prim_to_ent: fv.cc:421-431, Entropy 557-560
  T beta = FAST ? T(0.5) * rho * qdiv<FAST>(T(1), p) : rho / (T(2) * p);
  T s;
  if constexpr (FAST) s = log64(p) - P.gamma * log64(rho);
  else s = m_log(p) - P.gamma * m_log(rho);
  rec[E_RHO] = rho;
  rec[E_UX] = ux;
  rec[E_UY] = uy;
  rec[E_P] = p;
  rec[E_BK] = FAST ? beta : T(0.5) / (P.R * (p / (P.R * rho)));
  rec[E_VF] = (P.gamma - s) / P.gm1 - (ux * ux + uy * uy) * beta;
  rec[E_VX] = ux * (T(2) * beta);
  rec[E_VY] = uy * (T(2) * beta);
  rec[E_VL] = T(-2) * beta;
}

// kepec_flux in the face frame (n = normal, t = tangential): kepec.cc:10-48
template <bool FAST = false, typename T>
__device__ __forceinline__ void kepec2(const Phys<T>& P, T rl, T nl, T tl, T bl, T rr, T nr, T tr,
                                       T br, T area, T F[4], T* rho_o = nullptr,
                                       T* beta_o = nullptr) {
  T rho = lavg<FAST>(rl, rr);
  T vel2 = T(0.5) * ((nl * nl + tl * tl) + (nr * nr + tr * tr));
  T beta = lavg<FAST>(bl, br);
  T p = qdiv<FAST>(T(0.5) * (rl + rr), bl + br);
  T vn = nl * T(0.5) + nr * T(0.5);
  T vt = tl * T(0.5) + tr * T(0.5);
  T mass = rho * vn;
  T mn = p + vn * mass;
  T mt = vt * mass;
  T en = T(0.5) * (qdiv<FAST>(T(1), P.gm1 * beta) - vel2) * mass + (mn * vn + mt * vt);
  F[0] = mass * area;
  F[1] = mn * area;
  F[2] = mt * area;
  F[3] = en * area;
  if (rho_o) { *rho_o = rho; *beta_o = beta; }
}

// Ismail-Roe EC flux (roe_ec.cc:10-69) in the face frame
template <bool FAST = false, typename T>
__device__ __forceinline__ void roeec2(const Phys<T>& P, T rl, T nl, T tl, T pl, T rr, T nr, T tr,
                                       T pr, T area, T F[4]) {
  T bl = m_sqrt(qdiv<FAST>(rl, pl)), br = m_sqrt(qdiv<FAST>(rr, pr));
  T z1b = T(0.5) * (bl + br);
  T zn = nl * (T(0.5) * bl) + nr * (T(0.5) * br);
  T zt = tl * (T(0.5) * bl) + tr * (T(0.5) * br);
  T z5b = T(0.5) * (bl * pl + br * pr);
  T z1l = lavg<FAST>(bl, br);
  T z5l = lavg<FAST>(bl * pl, br * pr);
  T mass = z5l * zn;
  T mn = qdiv<FAST>(z5b, z1b) + zn * qdiv<FAST>(mass, z1b);
  T mt = zt * qdiv<FAST>(mass, z1b);
  T en = qdiv<FAST>(T(0.5) * (qdiv<FAST>(P.gpgm * mass, z1l) + (mn * zn + mt * zt)), z1b);
  F[0] = mass * area;
  F[1] = mn * area;
  F[2] = mt * area;
  F[3] = en * area;
}

// KEP flux (kep.cc:10-44), Enthalpy (fv.cc:565-569)
template <bool FAST = false, typename T>
__device__ __forceinline__ void kep2(const Phys<T>& P, T rl, T nl, T tl, T pl, T rr, T nr, T tr,
                                     T pr, T area, T F[4]) {
  T rho = T(0.5) * (rl + rr);
  T p = T(0.5) * (pl + pr);
  T vn = nl * T(0.5) + nr * T(0.5);
  T vt = tl * T(0.5) + tr * T(0.5);
  T al = m_sqrt(qdiv<FAST>(P.gamma * pl, rl)), ar = m_sqrt(qdiv<FAST>(P.gamma * pr, rr));
  T Hl = qdiv<FAST>(al * al, P.gm1) + T(0.5) * (nl * nl + tl * tl);
  T Hr = qdiv<FAST>(ar * ar, P.gm1) + T(0.5) * (nr * nr + tr * tr);
  T H = T(0.5) * (Hl + Hr);
  T mass = rho * vn;
  T mn = p + vn * mass;
  T mt = vt * mass;
  T en = rho * H * vn;
  F[0] = mass * area;
  F[1] = mn * area;
  F[2] = mt * area;
  F[3] = en * area;
}

// KEPS flux (keps.cc:10-48) in the face frame: arithmetic density, mean of
// the two betas
template <typename T>
__device__ __forceinline__ void keps2(const Phys<T>& P, T rl, T nl, T tl, T pl, T rr, T nr, T tr,
                                      T pr, T area, T F[4]) {
  const T rho = T(0.5) * (rl + rr);
  const T templ = pl / (P.R * rl), tempr = pr / (P.R * rr);
  const T vel2 = T(0.5) * ((nl * nl + tl * tl) + (nr * nr + tr * tr));
  const T betal = T(0.5) / (P.R * templ), betar = T(0.5) / (P.R * tempr);
  const T beta = T(0.5) * (betal + betar);
  const T p = T(0.5) * rho / beta;
  const T vn = nl * T(0.5) + nr * T(0.5);
  const T vt = tl * T(0.5) + tr * T(0.5);
  const T mass = rho * vn;
  const T mn = p + vn * mass;
  const T mt = vt * mass;
  const T en = T(0.5) * (T(1) / (P.gm1 * beta) - vel2) * mass + (mn * vn + mt * vt);
  F[0] = mass * area;
  F[1] = mn * area;
  F[2] = mt * area;
  F[3] = en * area;
}

// Roe's flux without entropy fix (roe.cc:10-86) in the face frame: the
// upwind cell's flux plus the one wave that crosses the face
template <typename T>
__device__ __forceinline__ void roe2(const Phys<T>& P, T rl, T nl, T tl, T pl, T rr, T nr, T tr,
                                     T pr, T area, T F[4]) {
  const T hl = P.gamma * pl / rl / P.gm1 + T(0.5) * (nl * nl + tl * tl);
  const T hr = P.gamma * pr / rr / P.gm1 + T(0.5) * (nr * nr + tr * tr);
  const T sl = m_sqrt(rl), sr = m_sqrt(rr);
  const T fl = sl / (sl + sr), fr = T(1) - fl;
  const T rho = sl * sr;
  const T vn = nl * fl + nr * fr, vt = tl * fl + tr * fr;
  const T h = hl * fl + hr * fr;
  const T a = m_sqrt(P.gm1 * (h - T(0.5) * (vn * vn + vt * vt)));
  const T dp = pr - pl, dV = nr - nl;
  T mass, mn, mt, en;
  if (vn >= T(0)) {
    const T coeff = T(0.5) * (dp - rho * a * dV) / (a * a);
    const T factor = m_min(vn - a, T(0)) * coeff;
    mass = rl * nl;
    mn = pl + nl * (rl * nl);
    mt = tl * (rl * nl);
    en = hl * mass;
    mass += factor;
    mn += vn * factor + (-a * factor);
    mt += vt * factor;
    en += (h - a * vn) * factor;
  } else {
    const T coeff = T(0.5) * (dp + rho * a * dV) / (a * a);
    const T factor = m_max(vn + a, T(0)) * coeff;
    mass = rr * nr;
    mn = pr + nr * (rr * nr);
    mt = tr * (rr * nr);
    en = hr * mass;
    mass -= factor;
    mn += vn * (-factor) + (-a * factor);
    mt += vt * (-factor);
    en -= (h + a * vn) * factor;
  }
  F[0] = mass * area;
  F[1] = mn * area;
  F[2] = mt * area;
  F[3] = en * area;
}

enum { FV_KEPES4_ROE = 0, FV_GENERIC = 1 };
enum { RV_SPWENO = 0, RV_DSP = 1, RV_GENERIC = 2, RV_ENO = 3 };

// FP32 jumps of NC characteristic components (defined in fv_mixed.hpp)
template <int MIX, int NC>
__device__ __forceinline__ void jumps_fp32(const Phys<double>& P, const float* d0,
                                           const float* d1, const float* d2, const float* z1,
                                           float* jmp, unsigned& nfb);

// MIX (M3_* of fv_mixed.hpp) != 0: the TeCNO face with its four
// characteristic jumps reconstructed in FP32 jump form (jumps_fp32) on the
// FP64 state -- the FP32 path of the TeCNO variants other than
// kepes_tecno4_roe, whose fused stage is Euler2Mixed
template <typename T, int G, int FV, int RV, int MIX = 0>
struct Euler2Ph {
  static_assert(MIX == 0 || (sizeof(T) == 8 && G == 2 && FV == FV_GENERIC && RV == RV_GENERIC),
                "the FP32 jumps run on the FP64 state of the generic face");
  using TS = T;
  // waves per SIMD the register budget is cut to: three for DSP-WENO (its
  // network chain needs the latency hiding; KH 2048^2 stage 1.29 ms vs 1.43
  // ms at the compiler's choice), the compiler's choice otherwise
#ifdef TECNO_EULER2_MIN_WAVES
  static constexpr int MIN_WAVES = TECNO_EULER2_MIN_WAVES;
#else
  static constexpr int MIN_WAVES = (RV == RV_DSP && FV == FV_KEPES4_ROE) ? 3 : 1;
#endif
  using TF = T;
  using TP = T;
  static constexpr int NV = 4, NREC = E_NREC;
  // the FP32 reconstruction path's FP64 records and fluxes take the
  // refined reciprocal, log64 and logavg_fast
  static constexpr bool FAST = MIX != 0;
  static constexpr size_t rec_bytes(int NB) { return (size_t)NREC * NB * sizeof(T); }
  // kepes_tecno4_roe: the distance-2 KEPEC fluxes F(0,2), F(1,3) of a face's
  // stencil are evaluated once per cell pair (with the face area, exactly as
  // the face would) and shared by the two faces that use the pair
  static constexpr bool PAIRS = FV == FV_KEPES4_ROE;
  __device__ static void pair(const Phys<T>& P, const void* ldsv, int NB, int ca, int cb, int ax,
                              T* out, int NP, int q) {
    const T* lds = (const T*)ldsv;
    const int rn = ax ? E_UY : E_UX, rt = ax ? E_UX : E_UY;
    T f0[4];
    kepec2<FAST>(P, lds[E_RHO * NB + ca], lds[rn * NB + ca], lds[rt * NB + ca], lds[E_BK * NB + ca],
           lds[E_RHO * NB + cb], lds[rn * NB + cb], lds[rt * NB + cb], lds[E_BK * NB + cb],
           P.area[ax], f0);
#pragma unroll
    for (int v = 0; v < 4; ++v) out[v * NP + q] = f0[v];
  }
  __device__ static void load(const Phys<T>& P, const T* U, int64_t N, int64_t idx, int flipx,
                              int flipy, void* lds, int NB, int c) {
    T rec[NREC];
    euler2_record<FAST>(P, U[idx], U[N + idx], U[2 * N + idx], U[3 * N + idx], flipx, flipy, rec);
#pragma unroll
    for (int r = 0; r < NREC; ++r) ((T*)lds)[r * NB + c] = rec[r];
  }

  // TeCNO face flux (kepes_tecno4_roe.cc:10-176 and variants), evaluated in
  // the face frame (normal n, tangential t) and returned as (rho, x, y, E).  The shear wave uses the x-face sign convention on both axes; its
  // scaled variable and eigenvector entries both change sign on y faces, so
  // the dissipation is unchanged bit for bit.  Sums over the x and y
  // components keep the reference's x-before-y order via `ax`.
  // pairs: the shared pair fluxes (pairs q and q+ps are this face's F(0,2)
  // and F(1,3)); nullptr evaluates them here
  __device__ static void face(const Phys<T>& P, const void* ldsv, int NB, int base, int stride,
                              int ax, T F[4], unsigned& nfb, const T* pairs = nullptr,
                              int NP = 0, int q = 0, int ps = 1) {
    const T* lds = (const T*)ldsv;
    constexpr int S = 2 * G;
    const int rn = ax ? E_UY : E_UX, rt = ax ? E_UX : E_UY;
    const int vn_ = ax ? E_VY : E_VX, vt_ = ax ? E_VX : E_VY;
    T rho[S], un[S], ut[S], bk[S];
#pragma unroll
    for (int l = 0; l < S; ++l) {
      const int c = base + l * stride;
      rho[l] = lds[E_RHO * NB + c];
      un[l] = lds[rn * NB + c];
      ut[l] = lds[rt * NB + c];
      bk[l] = lds[E_BK * NB + c];
    }
    const T area = P.area[ax];
    const int L = G - 1, R = G;  // centre pair
    const int fl = (FV == FV_KEPES4_ROE) ? 101 : P.flux;
    if (FV == FV_GENERIC && fl >= 125) {
      // fluxes without dissipation on reconstructed jumps: roe, kep4, keps,
      // keps4 (num_flux.cc:12-72); the 4th-order forms combine three
      // two-point fluxes as kep4.cc / keps4.cc
      T Fe[4];
      auto pl_ = [&](int l) { return lds[E_P * NB + base + l * stride]; };
      if (fl == 125) {
        roe2(P, rho[L], un[L], ut[L], pl_(L), rho[R], un[R], ut[R], pl_(R), area, Fe);
      } else if (fl == 127) {
        keps2(P, rho[L], un[L], ut[L], pl_(L), rho[R], un[R], ut[R], pl_(R), area, Fe);
      } else {
        auto two = [&](int a, int b, T f[4]) {
          if (fl == 126)
            kep2(P, rho[a], un[a], ut[a], pl_(a), rho[b], un[b], ut[b], pl_(b), area, f);
          else
            keps2(P, rho[a], un[a], ut[a], pl_(a), rho[b], un[b], ut[b], pl_(b), area, f);
        };
        T f0[4];
        two(L, R, f0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Fe[v] = T(0) + f0[v] * T(4.0 / 3.0);
        two(L - 1, R, f0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Fe[v] += f0[v] * T(-1.0 / 6.0);
        two(L, R + 1, f0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Fe[v] += f0[v] * T(-1.0 / 6.0);
      }
      F[0] = Fe[0];
      F[1] = ax ? Fe[2] : Fe[1];  // face frame (n, t) -> (x, y)
      F[2] = ax ? Fe[1] : Fe[2];
      F[3] = Fe[3];
      return;
    }
    T Fe[4], rhoh, betah;
    const bool pure_ec = fl >= 120;
    // ---- entropy-conservative part
    if (fl == 101 || fl == 103 || fl == 121) {  // kepec4
      T f0[4];
      kepec2<FAST>(P, rho[L], un[L], ut[L], bk[L], rho[R], un[R], ut[R], bk[R], area, f0, &rhoh, &betah);
#pragma unroll
      for (int v = 0; v < 4; ++v) Fe[v] = T(0) + f0[v] * T(4.0 / 3.0);
      if (G == 2 && pairs) {
#pragma unroll
        for (int v = 0; v < 4; ++v) f0[v] = pairs[v * NP + q];
      } else {
        kepec2<FAST>(P, rho[L - 1], un[L - 1], ut[L - 1], bk[L - 1], rho[R], un[R], ut[R], bk[R], area,
               f0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) Fe[v] += f0[v] * T(-1.0 / 6.0);
      if (G == 2 && pairs) {
#pragma unroll
        for (int v = 0; v < 4; ++v) f0[v] = pairs[v * NP + q + ps];
      } else {
        kepec2<FAST>(P, rho[L], un[L], ut[L], bk[L], rho[R + 1], un[R + 1], ut[R + 1], bk[R + 1], area,
               f0);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) Fe[v] += f0[v] * T(-1.0 / 6.0);
    } else {
      // centre-pair logavg values are needed by every TeCNO variant
      rhoh = lavg<FAST>(rho[L], rho[R]);
      betah = lavg<FAST>(bk[L], bk[R]);
      T pl = lds[E_P * NB + base + L * stride], pr = lds[E_P * NB + base + R * stride];
      if (fl == 100 || fl == 102 || fl == 120) {
        kepec2<FAST>(P, rho[L], un[L], ut[L], bk[L], rho[R], un[R], ut[R], bk[R], area, Fe);
      } else if (fl == 104 || fl == 122) {
        roeec2<FAST>(P, rho[L], un[L], ut[L], pl, rho[R], un[R], ut[R], pr, area, Fe);
      } else if (fl == 105 || fl == 123) {
        T f0[4];
        T p0 = lds[E_P * NB + base + (L - 1) * stride], p3 = lds[E_P * NB + base + (R + 1) * stride];
        roeec2<FAST>(P, rho[L], un[L], ut[L], pl, rho[R], un[R], ut[R], pr, area, f0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Fe[v] = T(0) + f0[v] * T(4.0 / 3.0);
        roeec2<FAST>(P, rho[L - 1], un[L - 1], ut[L - 1], p0, rho[R], un[R], ut[R], pr, area, f0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Fe[v] += f0[v] * T(-1.0 / 6.0);
        roeec2<FAST>(P, rho[L], un[L], ut[L], pl, rho[R + 1], un[R + 1], ut[R + 1], p3, area, f0);
#pragma unroll
        for (int v = 0; v < 4; ++v) Fe[v] += f0[v] * T(-1.0 / 6.0);
      } else {  // 106, 124: kep
        kep2<FAST>(P, rho[L], un[L], ut[L], pl, rho[R], un[R], ut[R], pr, area, Fe);
      }
    }
    if (pure_ec) {
      F[0] = Fe[0];
      F[1] = ax ? Fe[2] : Fe[1];  // face frame (n, t) -> (x, y)
      F[2] = ax ? Fe[1] : Fe[2];
      F[3] = Fe[3];
      return;
    }
    const T inv_area = qdiv<FAST>(T(1), area);
#pragma unroll
    for (int v = 0; v < 4; ++v) Fe[v] *= inv_area;

    // ---- entropy dissipation: R~ Lambda (Z+ - Z-)
    const T p = qdiv<FAST>(T(0.5) * (rho[L] + rho[R]), bk[L] + bk[R]);
    const T vn = un[L] * T(0.5) + un[R] * T(0.5);
    const T vt = ut[L] * T(0.5) + ut[R] * T(0.5);
    const T a = m_sqrt(qdiv<FAST>(T(0.5) * P.gamma, betah));
    const T vsq = vn * vn + vt * vt;
    const T H = qdiv<FAST>(a * a, P.gm1) + T(0.5) * vsq;
    const T c1 = m_sqrt(qdiv<FAST>(rhoh, P.gamma));
    const T psq = m_sqrt(p);
    const T S0 = P.s05 * c1, S1 = P.sgm1 * c1, S2 = psq;
    // scaled eigenvectors (rows rho, n, t, E; waves 0, 1, 2, 4)
    const T Rr0 = S0, Rn0 = (vn - a) * S0, Rt0 = vt * S0, RE0 = (H - a * vn) * S0;
    const T Rr1 = S1, Rn1 = vn * S1, Rt1 = vt * S1, RE1 = (T(0.5) * vsq) * S1;
    const T Rt2 = -S2, RE2 = (-vt) * S2;
    const T Rr4 = S0, Rn4 = (vn + a) * S0, Rt4 = vt * S0, RE4 = (H + a * vn) * S0;
    // eigenvalues
    T L0, L1, L4;
    const bool rus = (FV == FV_GENERIC) && (fl == 102 || fl == 103);
    if (rus) {
      L0 = L1 = L4 = m_abs(vn) + m_abs(a);
    } else {
      L1 = m_abs(vn);
      L0 = m_abs(vn - a);
      L4 = m_abs(vn + a);
      if (FV == FV_GENERIC && P.ec1_fix && fl != 106) {
        const T al = m_sqrt(qdiv<FAST>(T(0.5) * P.gamma, bk[L])),
                ar = m_sqrt(qdiv<FAST>(T(0.5) * P.gamma, bk[R]));
        L0 = L0 + T(1.0 / 6.0) * m_abs((un[L] - al) - (un[R] - ar));
        L4 = L4 + T(1.0 / 6.0) * m_abs(un[L] - un[R]);
      }
    }
    const T L2 = L1;
    // scaled entropy variables of the stencil
    T Z0[S], Z1[S], Z2[S], Z4[S];
#pragma unroll
    for (int l = 0; l < S; ++l) {
      const int c = base + l * stride;
      const T f = lds[E_VF * NB + c], Vn = lds[vn_ * NB + c], Vt = lds[vt_ * NB + c],
              lv = lds[E_VL * NB + c];
      T a0 = Rn0 * Vn, b0 = Rt0 * Vt;
      T a1 = Rn1 * Vn, b1 = Rt1 * Vt;
      T a4 = Rn4 * Vn, b4 = Rt4 * Vt;
      // x component first (reference order): n first on x faces, t first on y faces
      const T x0 = ax ? b0 : a0, y0 = ax ? a0 : b0;
      const T x1 = ax ? b1 : a1, y1 = ax ? a1 : b1;
      const T x4 = ax ? b4 : a4, y4 = ax ? a4 : b4;
      Z0[l] = ((Rr0 * f + x0) + y0) + RE0 * lv;
      Z1[l] = ((Rr1 * f + x1) + y1) + RE1 * lv;
      Z2[l] = Rt2 * Vt + RE2 * lv;
      Z4[l] = ((Rr4 * f + x4) + y4) + RE4 * lv;
    }
    T dZ0, dZ1, dZ2, dZ4;
    if constexpr (MIX != 0) {
      // projected jumps of neighbouring cells (FP64 differences), FP32 jumps
      float d0[4], d1[4], d2[4], z1[4], jmp[4];
      const T* Zc[4] = {Z0, Z1, Z2, Z4};
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        d0[i] = (float)(Zc[i][L] - Zc[i][L - 1]);
        d1[i] = (float)(Zc[i][R] - Zc[i][L]);
        d2[i] = (float)(Zc[i][R + 1] - Zc[i][R]);
        z1[i] = (float)Zc[i][L];
      }
      jumps_fp32<MIX, 4>(P, d0, d1, d2, z1, jmp, nfb);
      dZ0 = jmp[0];
      dZ1 = jmp[1];
      dZ2 = jmp[2];
      dZ4 = jmp[3];
    } else {
      T l_, r_;
      if (RV == RV_SPWENO) {
        sp_weno(Z0[L - 1], Z0[L], Z0[R], Z0[R + 1], P.corr, l_, r_); dZ0 = r_ - l_;
        sp_weno(Z1[L - 1], Z1[L], Z1[R], Z1[R + 1], P.corr, l_, r_); dZ1 = r_ - l_;
        sp_weno(Z2[L - 1], Z2[L], Z2[R], Z2[R + 1], P.corr, l_, r_); dZ2 = r_ - l_;
        sp_weno(Z4[L - 1], Z4[L], Z4[R], Z4[R + 1], P.corr, l_, r_); dZ4 = r_ - l_;
      } else if (RV == RV_DSP) {
        const bool a5 = P.arch5 != 0;
        dsp_rec(P.net, a5, Z0[L - 1], Z0[L], Z0[R], Z0[R + 1], l_, r_); dZ0 = r_ - l_;
        dsp_rec(P.net, a5, Z1[L - 1], Z1[L], Z1[R], Z1[R + 1], l_, r_); dZ1 = r_ - l_;
        dsp_rec(P.net, a5, Z2[L - 1], Z2[L], Z2[R], Z2[R + 1], l_, r_); dZ2 = r_ - l_;
        dsp_rec(P.net, a5, Z4[L - 1], Z4[L], Z4[R], Z4[R + 1], l_, r_); dZ4 = r_ - l_;
      } else if (RV == RV_ENO) {
        eno_rec(P.eno_k, Z0, G, l_, r_); dZ0 = r_ - l_;
        eno_rec(P.eno_k, Z1, G, l_, r_); dZ1 = r_ - l_;
        eno_rec(P.eno_k, Z2, G, l_, r_); dZ2 = r_ - l_;
        eno_rec(P.eno_k, Z4, G, l_, r_); dZ4 = r_ - l_;
      } else {
        recon_any(P, Z0, G, l_, r_); dZ0 = r_ - l_;
        recon_any(P, Z1, G, l_, r_); dZ1 = r_ - l_;
        recon_any(P, Z2, G, l_, r_); dZ2 = r_ - l_;
        recon_any(P, Z4, G, l_, r_); dZ4 = r_ - l_;
      }
    }
    const T Dr = (Rr0 * L0 * dZ0 + Rr1 * L1 * dZ1) + Rr4 * L4 * dZ4;
    const T Dn = (Rn0 * L0 * dZ0 + Rn1 * L1 * dZ1) + Rn4 * L4 * dZ4;
    const T Dt = ((Rt0 * L0 * dZ0 + Rt1 * L1 * dZ1) + Rt2 * L2 * dZ2) + Rt4 * L4 * dZ4;
    const T DE = ((RE0 * L0 * dZ0 + RE1 * L1 * dZ1) + RE2 * L2 * dZ2) + RE4 * L4 * dZ4;
    const T Fn = (Fe[1] - T(0.5) * Dn) * area;
    const T Ft = (Fe[2] - T(0.5) * Dt) * area;
    F[0] = (Fe[0] - T(0.5) * Dr) * area;
    F[1] = ax ? Ft : Fn;  // face frame (n, t) -> (x, y)
    F[2] = ax ? Fn : Ft;
    F[3] = (Fe[3] - T(0.5) * DE) * area;
  }

  __device__ static T speed(const Phys<T>& P, const T* U, int64_t N, int64_t idx) {
    T rho = U[idx], mx = U[N + idx], my = U[2 * N + idx], E = U[3 * N + idx];
    T inv = T(1) / rho;
    T ux = mx * inv, uy = my * inv;
    T p = P.gm1 * (E - T(0.5) * (mx * mx + my * my) / rho);
    return m_sqrt(ux * ux + uy * uy) + m_sqrt(P.gamma * p / rho);
  }
  __device__ static bool bad(const T* U, int64_t N, int64_t idx) {
    T rho = U[idx], mx = U[N + idx], my = U[2 * N + idx], E = U[3 * N + idx];
    T p = E - T(0.5) * (mx * mx + my * my) / rho;
    return !(rho > T(0)) || !(p > T(0)) || !(p < T(INFINITY));
  }
};

}  // namespace tecno
