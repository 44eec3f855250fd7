// fv_euler3.hpp — 3D Euler / Navier-Stokes stage in FP64 (reference
// arithmetic): the TeCNO and EC fluxes in the global frame with the unit axis
// normal (kepes_tecno4_roe.cc:10-176 and variants, kepec.cc, roe_ec.cc,
// kep.cc) and the node-centred viscous fluxes of compute_visc_residual
// (fv.cc:786-1157, vflux_ES.cc:15-151, vflux_KE.cc:15-96).
//
// A workgroup owns an 8x8x2 tile and loads the whole (TX+2G)x(TY+2G)x(TZ+2G)
// box of cell records into LDS (edges and corners included: the viscous node
// operators read the 8 cells around each node).  Face fluxes are evaluated
// one per lane per round; the (TX+1)(TY+1)(TZ+1) viscous node fluxes of the
// tile, one per lane; both go to LDS over the dead records; each cell then
// gathers its 6 faces and 8 corner nodes, applies the SSPRK3 update and
// reduces the CFL speed.
//
// With MIX != M3_NONE (the FP32 path, Euler3Ph<double, 2, 0, MIX>) the face
// reconstructs the five characteristic jumps in FP32 jump form, as the 2D
// mixed face does (fv_mixed.hpp); state, records, entropy-conservative flux,
// eigensystem, dissipation product and viscous terms stay FP64.
#pragma once
#include "fv_kernels.hpp"
#include "fv_mixed.hpp"

namespace tecno {

// record layout, double [E3_NREC][NB]: primitive, beta as kepec_flux computes
// it, entropy variables (fv.cc:421-431).  The FP32 reconstruction path keeps
// the first E3_NREC_MIX fields only (its beta record is beta itself) and
// forms V1..V3 and V5 from u and beta where they are read (e3_vent), so that
// its records fit three workgroups per CU.
enum { E3_RHO = 0, E3_U, E3_V, E3_W, E3_P, E3_BK, E3_VF, E3_V1, E3_V2, E3_V3, E3_VL, E3_NREC };
constexpr int E3_NREC_MIX = E3_VF + 1;

// entropy variable r (0..4) of record c: V1..V3 = u (2 beta) and V5 = -2 beta
// are the products the full record stores, so both layouts give the same bits
template <int MIX, typename T>
__device__ __forceinline__ T e3_vent(const T* ld, int NB, int c, int r) {
  if constexpr (MIX == 0) {
    return ld[(E3_VF + r) * NB + c];
  } else {
    if (r == 0) return ld[E3_VF * NB + c];
    const T b2 = T(2) * ld[E3_BK * NB + c];
    if (r == 4) return -b2;
    return ld[(E3_U + r - 1) * NB + c] * b2;
  }
}

#ifndef TECNO_STAGE3_WAVES
#define TECNO_STAGE3_WAVES 2
#endif
#ifndef TECNO_STAGE3_MIX_WAVES
#define TECNO_STAGE3_MIX_WAVES 3
#endif

// 3D face: the five characteristic jumps in FP32 jump form.  The
// differences of the FP64 entropy variables of neighbouring cells are
// projected on the scaled eigenvectors Rm in FP64 (a projected jump can be a
// near-total cancellation of its terms); jumps_fp32 (fv_mixed.hpp) forms
// SP-WENO's or DSP-WENO's jump in FP32.
template <int MIX>
__device__ __forceinline__ void jumps3_fp32(const Phys<double>& P, const double* ld, int NB,
                                            int base, int stride, const double Rm[5][5],
                                            double dZ[5], unsigned& nfb) {
  float d0[5], d1[5], d2[5], z1[5];
  {
    double V[4][5];
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int cc = base + l * stride;
#pragma unroll
      for (int r = 0; r < 5; ++r) V[l][r] = e3_vent<MIX>(ld, NB, cc, r);
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, zz = 0.0;
#pragma unroll
      for (int r = 0; r < 5; ++r) {
        a0 += Rm[r][i] * (V[1][r] - V[0][r]);
        a1 += Rm[r][i] * (V[2][r] - V[1][r]);
        a2 += Rm[r][i] * (V[3][r] - V[2][r]);
        zz += Rm[r][i] * V[1][r];
      }
      d0[i] = (float)a0;
      d1[i] = (float)a1;
      d2[i] = (float)a2;
      z1[i] = (float)zz;
    }
  }
  float jmp[5];
  jumps_fp32<MIX, 5>(P, d0, d1, d2, z1, jmp, nfb);
#pragma unroll
  for (int i = 0; i < 5; ++i) dZ[i] = (double)jmp[i];
}


template <typename T>
struct Pr3 {
  T rho, u[3], p, bk;
};

// Parameter::Viscosity parameter.cc:646-663
template <typename T>
__device__ __forceinline__ double visc_mu(const Phys<T>& P, double Tm) {
  if (P.mu_model == 1) return P.mu_ref * pow(Tm / P.T_ref, 1.5) * (P.T_ref + P.T_0) / (Tm + P.T_0);
  if (P.mu_model == 2) return P.mu_ref * pow(Tm / P.T_ref, P.omega);
  return P.mu_ref;
}

template <typename T>
__device__ __forceinline__ T vdot3(const T a[3], const T b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// kepec_flux kepec.cc:10-48
template <bool FAST = false, typename T>
__device__ __forceinline__ void kepec3(const Phys<T>& P, const Pr3<T>& l, const Pr3<T>& r,
                                       const T n[3], T area, T F[5]) {
  const T rho = lavg<FAST>(l.rho, r.rho);
  const T vel2 = T(0.5) * (vdot3(l.u, l.u) + vdot3(r.u, r.u));
  const T beta = lavg<FAST>(l.bk, r.bk);
  const T p = qdiv<FAST>(T(0.5) * (l.rho + r.rho), l.bk + r.bk);
  T vel[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) vel[d] = l.u[d] * T(0.5) + r.u[d] * T(0.5);
  const T vn = vdot3(vel, n);
  const T mass = rho * vn;
  T mom[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) mom[d] = n[d] * p + vel[d] * mass;
  const T en = T(0.5) * (qdiv<FAST>(T(1), (P.gamma - T(1)) * beta) - vel2) * mass + vdot3(mom, vel);
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

// roe_ec_flux roe_ec.cc:10-69
template <bool FAST = false, typename T>
__device__ __forceinline__ void roeec3(const Phys<T>& P, const Pr3<T>& l, const Pr3<T>& r,
                                       const T n[3], T area, T F[5]) {
  const T bl = m_sqrt(qdiv<FAST>(l.rho, l.p)), br = m_sqrt(qdiv<FAST>(r.rho, r.p));
  const T z1b = T(0.5) * (bl + br);
  T zvb[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) zvb[d] = l.u[d] * (T(0.5) * bl) + r.u[d] * (T(0.5) * br);
  const T z5b = T(0.5) * (bl * l.p + br * r.p);
  const T zvn = vdot3(zvb, n);
  const T z1l = lavg<FAST>(bl, br);
  const T z5l = lavg<FAST>(bl * l.p, br * r.p);
  const T mass = z5l * zvn;
  T mom[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) mom[d] = n[d] * qdiv<FAST>(z5b, z1b) + zvb[d] * qdiv<FAST>(mass, z1b);
  const T en = qdiv<FAST>(T(0.5) * (qdiv<FAST>((P.gamma + T(1)) / (P.gamma - T(1)) * mass, z1l) +
                                    vdot3(mom, zvb)),
                          z1b);
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

// kep_flux kep.cc:10-44 with Enthalpy fv.cc:565-569
template <bool FAST = false, typename T>
__device__ __forceinline__ void kep3(const Phys<T>& P, const Pr3<T>& l, const Pr3<T>& r,
                                     const T n[3], T area, T F[5]) {
  const T rho = T(0.5) * (l.rho + r.rho);
  const T p = T(0.5) * (l.p + r.p);
  T vel[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) vel[d] = l.u[d] * T(0.5) + r.u[d] * T(0.5);
  const T al = m_sqrt(qdiv<FAST>(P.gamma * l.p, l.rho)), ar = m_sqrt(qdiv<FAST>(P.gamma * r.p, r.rho));
  const T Hl = qdiv<FAST>(al * al, P.gamma - T(1)) + T(0.5) * vdot3(l.u, l.u);
  const T Hr = qdiv<FAST>(ar * ar, P.gamma - T(1)) + T(0.5) * vdot3(r.u, r.u);
  const T H = T(0.5) * (Hl + Hr);
  const T vn = vdot3(vel, n);
  const T mass = rho * vn;
  T mom[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) mom[d] = n[d] * p + vel[d] * mass;
  const T en = rho * H * vn;
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

// keps_flux keps.cc:10-48
template <typename T>
__device__ __forceinline__ void keps3(const Phys<T>& P, const Pr3<T>& l, const Pr3<T>& r,
                                      const T n[3], T area, T F[5]) {
  const T rho = T(0.5) * (l.rho + r.rho);
  const T templ = l.p / (P.R * l.rho), tempr = r.p / (P.R * r.rho);
  const T vel2 = T(0.5) * (vdot3(l.u, l.u) + vdot3(r.u, r.u));
  const T betal = T(0.5) / (P.R * templ), betar = T(0.5) / (P.R * tempr);
  const T beta = T(0.5) * (betal + betar);
  const T p = T(0.5) * rho / beta;
  T vel[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) vel[d] = l.u[d] * T(0.5) + r.u[d] * T(0.5);
  const T mass = rho * vdot3(vel, n);
  T mom[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) mom[d] = n[d] * p + vel[d] * mass;
  const T en = T(0.5) * (T(1) / ((P.gamma - T(1)) * beta) - vel2) * mass + vdot3(mom, vel);
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

// roe_flux roe.cc:10-86 (no entropy fix)
template <typename T>
__device__ __forceinline__ void roe3(const Phys<T>& P, const Pr3<T>& l, const Pr3<T>& r,
                                     const T n[3], T area, T F[5]) {
  const T hl = P.gamma * l.p / l.rho / (P.gamma - T(1)) + T(0.5) * vdot3(l.u, l.u);
  const T hr = P.gamma * r.p / r.rho / (P.gamma - T(1)) + T(0.5) * vdot3(r.u, r.u);
  const T sl = m_sqrt(l.rho), sr = m_sqrt(r.rho);
  const T fl = sl / (sl + sr), fr = T(1) - fl;
  const T rho = sl * sr;
  T vel[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) vel[d] = l.u[d] * fl + r.u[d] * fr;
  const T h = hl * fl + hr * fr;
  const T vn = vdot3(vel, n);
  const T a = m_sqrt((P.gamma - T(1)) * (h - T(0.5) * vdot3(vel, vel)));
  const T dp = r.p - l.p;
  const T vln = vdot3(l.u, n), vrn = vdot3(r.u, n);
  const T dV = vrn - vln;
  T mass, mom[3], en;
  if (vn >= T(0)) {
    const T coeff = T(0.5) * (dp - rho * a * dV) / (a * a);
    const T factor = m_min(vn - a, T(0)) * coeff;
    mass = l.rho * vln;
#pragma unroll
    for (int d = 0; d < 3; ++d) mom[d] = n[d] * l.p + l.u[d] * (l.rho * vln);
    en = hl * mass;
    mass += factor;
#pragma unroll
    for (int d = 0; d < 3; ++d) mom[d] += vel[d] * factor + n[d] * (-a * factor);
    en += (h - a * vn) * factor;
  } else {
    const T coeff = T(0.5) * (dp + rho * a * dV) / (a * a);
    const T factor = m_max(vn + a, T(0)) * coeff;
    mass = r.rho * vrn;
#pragma unroll
    for (int d = 0; d < 3; ++d) mom[d] = n[d] * r.p + r.u[d] * (r.rho * vrn);
    en = hr * mass;
    mass -= factor;
#pragma unroll
    for (int d = 0; d < 3; ++d) mom[d] += vel[d] * (-factor) + n[d] * (-a * factor);
    en -= (h + a * vn) * factor;
  }
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

// FAM 0: the TeCNO and entropy-conservative fluxes (100-106, 120-124) with
//        sp_weno(c), eno, sp_weno_dl or first;
// FAM 1: roe, kep4, keps, keps4 (125-128);
// FAM 2: as FAM 0 with second, minmod or minabs.
// Separate instantiations keep the face of the common case inlined and its
// register pressure unchanged by the rarely used options.
template <typename T, int G, int FAM = 0, int MIX = M3_NONE>
struct Euler3Ph {
  static_assert(MIX == M3_NONE || (sizeof(T) == 8 && G == 2 && FAM == 0),
                "the FP32 reconstruction runs on the FP64 state with the TeCNO family");
  using TS = T;
  using TF = T;
  using TP = T;
  static constexpr int NV = 5, NREC = MIX != M3_NONE ? E3_NREC_MIX : E3_NREC;
  static constexpr bool FR_SELECT = MIX != M3_NONE;
  // The FP64 kernel of FAM 0 runs at the 256-VGPR limit of two waves per
  // SIMD: with the XCD-aware tile order and the stage-dependent state reads
  // its compiler output spills 1.4 KB per lane (TGV 128^3 stage twice as
  // slow), so it keeps the launch order and reads U and U^n in every stage.
  static constexpr bool REG_SLACK = !(FAM == 0 && MIX == M3_NONE);
  // waves per SIMD the register budget is cut to: the FP32 path's seven
  // record fields (48 KB of LDS per 8x8x2 tile) let three workgroups share a
  // CU, so its registers are cut to three waves per SIMD
  static constexpr int MIN_WAVES3 = MIX == M3_NONE ? TECNO_STAGE3_WAVES : TECNO_STAGE3_MIX_WAVES;

  __device__ static Pr3<T> prim(const T* ld, int NB, int c) {
    Pr3<T> q;
    q.rho = ld[E3_RHO * NB + c];
    q.u[0] = ld[E3_U * NB + c];
    q.u[1] = ld[E3_V * NB + c];
    q.u[2] = ld[E3_W * NB + c];
    q.p = ld[E3_P * NB + c];
    q.bk = ld[E3_BK * NB + c];
    return q;
  }

  // con_to_prim fv.cc:388-400, wall mirror of the normal velocity, then
  // prim_to_ent fv.cc:421-431 (Entropy 557-560)
  __device__ static void load(const Phys<T>& P, const T* U, int64_t N, int64_t idx, int flip_axis,
                              T* ld, int NB, int c) {
    const T rho = U[idx], E = U[4 * N + idx];
    // the FP32 reconstruction path (MIX) is measured against the FP64 path
    // within a tolerance, so its records take the refined reciprocal and
    // log64 (2 ulp) where the FP64 path keeps the reference's divisions and
    // the library logarithm
    const T inv = MIX != M3_NONE ? rcp64(rho) : T(1) / rho;
    T u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = U[(1 + d) * N + idx] * inv;
    const T m2 = U[N + idx] * U[N + idx] + U[2 * N + idx] * U[2 * N + idx] +
                 U[3 * N + idx] * U[3 * N + idx];
    const T p = (P.gamma - T(1)) * (E - T(0.5) * (MIX != M3_NONE ? m2 * inv : m2 / rho));
#pragma unroll
    for (int d = 0; d < 3; ++d)
      if (flip_axis & (1 << d)) u[d] = -u[d];
    const T beta = MIX != M3_NONE ? T(0.5) * rho * rcp64(p) : rho / (T(2) * p);
    const T s = MIX != M3_NONE ? log64(p) - P.gamma * log64(rho)
                               : m_log(p) - P.gamma * m_log(rho);
    ld[E3_RHO * NB + c] = rho;
    ld[E3_U * NB + c] = u[0];
    ld[E3_V * NB + c] = u[1];
    ld[E3_W * NB + c] = u[2];
    ld[E3_P * NB + c] = p;
    // 0.5 / (R T) with T = p / (R rho): beta itself on the FP32 path
    ld[E3_BK * NB + c] = MIX != M3_NONE ? beta : T(0.5) / (P.R * (p / (P.R * rho)));
    ld[E3_VF * NB + c] = (P.gamma - s) / (P.gamma - T(1)) - vdot3(u, u) * beta;
    if constexpr (MIX == M3_NONE) {
      ld[E3_V1 * NB + c] = u[0] * (T(2) * beta);
      ld[E3_V2 * NB + c] = u[1] * (T(2) * beta);
      ld[E3_V3 * NB + c] = u[2] * (T(2) * beta);
      ld[E3_VL * NB + c] = T(-2) * beta;
    }
  }

  // FV::num_flux for one face from the 2G cells base + l*stride
  __device__ static void face(const Phys<T>& P, const T* ld, int NB, int base, int stride, int ax,
                              T F[5]) {
    unsigned nfb = 0;
    face(P, ld, NB, base, stride, ax, F, nfb);
  }
  __device__ static void face(const Phys<T>& P, const T* ld, int NB, int base, int stride, int ax,
                              T F[5], unsigned& nfb) {
    constexpr int S = 2 * G;
    const int L = G - 1, R = G;
    // unit normal of axis ax, without indexing by the lane's axis (a register
    // array indexed that way would live in scratch memory)
    const T n[3] = {ax == 0 ? T(1) : T(0), ax == 1 ? T(1) : T(0), ax == 2 ? T(1) : T(0)};
    const T area = P.area[ax];
    Pr3<T> q[S];
#pragma unroll
    for (int l = 0; l < S; ++l) q[l] = prim(ld, NB, base + l * stride);
    const int fl = P.flux;
    constexpr bool FAST = MIX != M3_NONE;
    auto ec = [&](int kind, const Pr3<T>& a, const Pr3<T>& b, T out[5]) {
      if (kind == 0)
        kepec3<FAST>(P, a, b, n, area, out);
      else if (kind == 1)
        roeec3<FAST>(P, a, b, n, area, out);
      else if (kind == 2 || FAM == 0)
        kep3<FAST>(P, a, b, n, area, out);
      else
        keps3(P, a, b, n, area, out);
    };
    auto ec4 = [&](int kind, T out[5]) {  // kepec4.cc:10-38, roe_ec4.cc:10-34
      T f0[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] = T(0);
      ec(kind, q[L], q[R], f0);
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] += f0[v] * T(4.0 / 3.0);
      ec(kind, q[L - 1], q[R], f0);
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] += f0[v] * T(-1.0 / 6.0);
      ec(kind, q[L], q[R + 1], f0);
#pragma unroll
      for (int v = 0; v < 5; ++v) out[v] += f0[v] * T(-1.0 / 6.0);
    };
    // pure entropy-conservative fluxes
    if (fl == 120) { ec(0, q[L], q[R], F); return; }
    if (fl == 121) { ec4(0, F); return; }
    if (fl == 122) { ec(1, q[L], q[R], F); return; }
    if (fl == 123) { ec4(1, F); return; }
    if (fl == 124) { ec(2, q[L], q[R], F); return; }
    if constexpr (FAM == 1) {
      if (fl == 125) roe3(P, q[L], q[R], n, area, F);
      else if (fl == 126) ec4(2, F);
      else if (fl == 127) ec(3, q[L], q[R], F);
      else ec4(3, F);  // 128 keps4
      return;
    }
    // TeCNO: entropy-conservative part
    if (fl == 101 || fl == 103) ec4(0, F);
    else if (fl == 100 || fl == 102) ec(0, q[L], q[R], F);
    else if (fl == 105) ec4(1, F);
    else if (fl == 104) ec(1, q[L], q[R], F);
    else ec(2, q[L], q[R], F);  // 106 kep_tecno_roe
    const T inv_area = qdiv<FAST>(T(1), area);
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] *= inv_area;
    // means of the centre pair, eigensystem (kepes_tecno4_roe.cc:39-113)
    const Pr3<T>& pl = q[L];
    const Pr3<T>& pr = q[R];
    const T rho = lavg<FAST>(pl.rho, pr.rho);
    const T beta = lavg<FAST>(pl.bk, pr.bk);
    const T p = qdiv<FAST>(T(0.5) * (pl.rho + pr.rho), pl.bk + pr.bk);
    T vel[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) vel[d] = pl.u[d] * T(0.5) + pr.u[d] * T(0.5);
    const T vn = vdot3(vel, n);
    const T a = m_sqrt(qdiv<FAST>(T(0.5) * P.gamma, beta));
    const T H = qdiv<FAST>(a * a, P.gamma - T(1)) + T(0.5) * vdot3(vel, vel);
    const T v1 = vel[0] * n[1] - vel[1] * n[0];
    const T v2 = vel[2] * n[0] - vel[0] * n[2];
    T Rm[5][5] = {{T(1), T(1), T(0), T(0), T(1)},
                  {vel[0] - a * n[0], vel[0], n[1], -n[2], vel[0] + a * n[0]},
                  {vel[1] - a * n[1], vel[1], -n[0], T(0), vel[1] + a * n[1]},
                  {vel[2] - a * n[2], vel[2], T(0), n[0], vel[2] + a * n[2]},
                  {H - a * vn, T(0.5) * vdot3(vel, vel), v1, v2, H + a * vn}};
    const T const1 = m_sqrt(qdiv<FAST>(rho, P.gamma));
    const T psq = m_sqrt(p);
    const T Sc[5] = {m_sqrt(T(0.5)) * const1, m_sqrt(P.gamma - T(1)) * const1, psq, psq,
                     m_sqrt(T(0.5)) * const1};
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
      for (int j = 0; j < 5; ++j) Rm[j][i] *= Sc[i];
    T Lam[5];
    if (fl == 102 || fl == 103) {
      const T lr = m_abs(vn) + m_abs(a);
#pragma unroll
      for (int i = 0; i < 5; ++i) Lam[i] = lr;
    } else {
      const T l2 = m_abs(vn);
      T ea = T(0);
      if (P.ec1_fix && fl != 106) ea = T(1.0 / 6.0);
      const T vnl = vdot3(pl.u, n), vnr = vdot3(pr.u, n);
      const T al = m_sqrt(qdiv<FAST>(T(0.5) * P.gamma, pl.bk)),
              ar = m_sqrt(qdiv<FAST>(T(0.5) * P.gamma, pr.bk));
      Lam[0] = (T(1) + T(0)) * m_abs(vn - a) + ea * m_abs((vnl - al) - (vnr - ar));
      Lam[1] = l2;
      Lam[2] = l2;
      Lam[3] = l2;
      Lam[4] = (T(1) + T(0)) * m_abs(vn + a) + ea * m_abs(vnl - vnr);
    }
    // Z = R~^T V on the stencil, reconstruct, Diff = R~ Lambda dZ
    T dZ[5];
    if constexpr (MIX != M3_NONE) {
      jumps3_fp32<MIX>(P, ld, NB, base, stride, Rm, dZ, nfb);
    } else {
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      T st[S];
#pragma unroll
      for (int k = 0; k < S; ++k) {
        const int cc = base + k * stride;
        st[k] = T(0) + (Rm[0][i] * ld[E3_VF * NB + cc] + Rm[1][i] * ld[E3_V1 * NB + cc] +
                        Rm[2][i] * ld[E3_V2 * NB + cc] + Rm[3][i] * ld[E3_V3 * NB + cc] +
                        Rm[4][i] * ld[E3_VL * NB + cc]);
      }
      T zl, zr;
      recon_any<FAM == 2 ? 2 : 1>(P, st, G, zl, zr);
      dZ[i] = zr - zl;
    }
    }
    T Diff[5];
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      Diff[i] = T(0);
#pragma unroll
      for (int j = 0; j < 5; ++j) Diff[i] += Rm[i][j] * Lam[j] * dZ[j];
    }
    F[0] -= T(0.5) * Diff[0];
    F[1] -= T(0.5) * Diff[1];
    F[2] -= T(0.5) * Diff[2];
    if (P.fix_zmom)
      F[3] -= T(0.5) * Diff[3];
    else
      F[2] -= T(0.5) * Diff[3];  // kepes_tecno4_roe.cc:170 (z dissipation into y)
    F[4] -= T(0.5) * Diff[4];
#pragma unroll
    for (int v = 0; v < 5; ++v) F[v] *= area;
  }

  // viscous node fluxes (momentum x, y, z and energy of the x, y and z
  // fluxes, already times -area) from the 8 cells c[a][b][q] around a node
  __device__ static void node(const Phys<T>& P, const T* ld, int NB, const int c[2][2][2],
                              T X[4], T Y[4], T Z[4]) {
    const T dx = P.h[0], dy = P.h[1], dz = P.h[2];
    const T f1 = T(1.0 / 8.0);
    if (P.model == 12) {
      // nodal_vlux_ES vflux_ES.cc:15-151: node averages and differences of V
      const T fx = T(1.0 / 4.0) / dx, fy = T(1.0 / 4.0) / dy, fz = T(1.0 / 4.0) / dz;
      T avg[5], Tx[5], Ty[5], Tz[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        auto V = [&](int a, int b, int q) { return e3_vent<MIX>(ld, NB, c[a][b][q], k); };
        avg[k] = V(1, 1, 1) * f1 + V(0, 1, 1) * f1 + V(1, 0, 1) * f1 + V(1, 1, 0) * f1;
        avg[k] += V(0, 0, 1) * f1 + V(0, 1, 0) * f1 + V(1, 0, 0) * f1 + V(0, 0, 0) * f1;
        Tx[k] = V(1, 1, 1) * fx + V(1, 1, 0) * fx + V(1, 0, 1) * fx + V(1, 0, 0) * fx;
        Tx[k] += V(0, 1, 1) * -fx + V(0, 1, 0) * -fx + V(0, 0, 1) * -fx + V(0, 0, 0) * -fx;
        Ty[k] = V(1, 1, 1) * fy + V(1, 1, 0) * fy + V(0, 1, 1) * fy + V(0, 1, 0) * fy;
        Ty[k] += V(1, 0, 1) * -fy + V(1, 0, 0) * -fy + V(0, 0, 1) * -fy + V(0, 0, 0) * -fy;
        Tz[k] = V(1, 1, 1) * fz + V(1, 0, 1) * fz + V(0, 1, 1) * fz + V(0, 0, 1) * fz;
        Tz[k] += V(1, 1, 0) * -fz + V(1, 0, 0) * -fz + V(0, 1, 0) * -fz + V(0, 0, 0) * -fz;
      }
      const T T_avg = T(-1) / avg[4] / P.R;
      const T mu = T(visc_mu(P, (double)T_avg));
      const T kappa = mu * T(P.Cp) / T(P.prandtl);
      const T c1 = T(4) * mu / T(3), c2 = T(-2) * mu / T(3), c3 = kappa / P.R, e4 = T(1) / avg[4],
              e1 = avg[1] * e4, e2 = avg[2] * e4, e3 = avg[3] * e4;
      // x flux: K11 V_x + K12 V_y + K13 V_z
      X[0] = c1 * e4 * (-Tx[1] + e1 * Tx[4]);
      X[1] = mu * e4 * (-Tx[2] + e2 * Tx[4]);
      X[2] = mu * e4 * (-Tx[3] + e3 * Tx[4]);
      X[3] = e4 * (c1 * e1 * Tx[1] + mu * e2 * Tx[2] + mu * e3 * Tx[3] +
                   (-c1 * e1 * e1 - mu * (e2 * e2 + e3 * e3) + c3 * e4) * Tx[4]);
      X[0] += c2 * e4 * (-Ty[2] + e2 * Ty[4]);
      X[1] += mu * e4 * (-Ty[1] + e1 * Ty[4]);
      X[3] += e4 * (mu * e2 * Ty[1] + c2 * e1 * Ty[2] - (mu + c2) * e1 * e2 * Ty[4]);
      X[0] += c2 * e4 * (-Tz[3] + e3 * Tz[4]);
      X[2] += mu * e4 * (-Tz[1] + e1 * Tz[4]);
      X[3] += e4 * (mu * e3 * Tz[1] + c2 * e1 * Tz[3] - (mu + c2) * e1 * e3 * Tz[4]);
      // y flux: the reference's K21 part is overwritten by the K22 assignment
      Y[0] = mu * e4 * (-Ty[1] + e1 * Ty[4]);
      Y[1] = c1 * e4 * (-Ty[2] + e2 * Ty[4]);
      Y[2] = mu * e4 * (-Ty[3] + e3 * Ty[4]);
      Y[3] = e4 * (mu * e1 * Ty[1] + c1 * e2 * Ty[2] + mu * e3 * Ty[3] +
                   (-c1 * e2 * e2 - mu * (e1 * e1 + e3 * e3) + c3 * e4) * Ty[4]);
      Y[1] += c2 * e4 * (-Tz[3] + e3 * Tz[4]);
      Y[2] += mu * e4 * (-Tz[2] + e2 * Tz[4]);
      Y[3] += e4 * (mu * e3 * Tz[2] + c2 * e2 * Tz[3] - (mu + c2) * e2 * e3 * Tz[4]);
      if (P.fix_visc) {  // K21 V_x (vflux_ES.cc:103-108), kept
        Y[0] += mu * e4 * (-Tx[2] + e2 * Tx[4]);
        Y[1] += c2 * e4 * (-Tx[1] + e1 * Tx[4]);
        Y[3] += e4 * (c2 * e2 * Tx[1] + mu * e1 * Tx[2] - (mu + c2) * e1 * e2 * Tx[4]);
      }
      // z flux: K31 and K32 parts overwritten by the K33 assignment
      Z[0] = mu * e4 * (-Tz[1] + e1 * Tz[4]);
      Z[1] = mu * e4 * (-Tz[2] + e2 * Tz[4]);
      Z[2] = c1 * e4 * (-Tz[3] + e3 * Tz[4]);
      Z[3] = e4 * (mu * e1 * Tz[1] + mu * e2 * Tz[2] + c1 * e3 * Tz[3] +
                   (-c1 * e3 * e3 - mu * (e1 * e1 + e2 * e2) + c3 * e4) * Tz[4]);
      if (P.fix_visc) {  // K31 V_x and K32 V_y (vflux_ES.cc:127-138), kept
        Z[0] += mu * e4 * (-Tx[3] + e3 * Tx[4]);
        Z[2] += c2 * e4 * (-Tx[1] + e1 * Tx[4]);
        Z[3] += e4 * (c2 * e3 * Tx[1] + mu * e1 * Tx[3] - (mu + c2) * e1 * e3 * Tx[4]);
        Z[1] += mu * e4 * (-Ty[3] + e3 * Ty[4]);
        Z[2] += c2 * e4 * (-Ty[2] + e2 * Ty[4]);
        Z[3] += e4 * (c2 * e3 * Ty[2] + mu * e2 * Ty[3] - (mu + c2) * e2 * e3 * Ty[4]);
      }
    } else {
      // nodal_vlux_KE vflux_KE.cc:15-96 with nodal_avg_and_vel_derivatives
      const T fx = T(1) / (T(4) * dx), fy = T(1) / (T(4) * dy), fz = T(1) / (T(4) * dz);
      auto Q = [&](int r, int a, int b, int q) { return ld[r * NB + c[a][b][q]]; };
      T avg[5];
      const int rr[5] = {E3_RHO, E3_U, E3_V, E3_W, E3_P};
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const int r = rr[k];
        avg[k] = Q(r, 1, 1, 1) * f1 + Q(r, 0, 1, 1) * f1 + Q(r, 1, 0, 1) * f1 + Q(r, 1, 1, 0) * f1;
        avg[k] += Q(r, 0, 0, 1) * f1 + Q(r, 0, 1, 0) * f1 + Q(r, 1, 0, 0) * f1 + Q(r, 0, 0, 0) * f1;
      }
      T Tx[3], Ty[3], Tz[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const int r = E3_U + d;
        Tx[d] = Q(r, 1, 1, 1) * fx + Q(r, 1, 1, 0) * fx + Q(r, 1, 0, 1) * fx + Q(r, 1, 0, 0) * fx;
        Tx[d] += Q(r, 0, 1, 1) * -fx + Q(r, 0, 1, 0) * -fx + Q(r, 0, 0, 1) * -fx + Q(r, 0, 0, 0) * -fx;
        Ty[d] = Q(r, 1, 1, 1) * fy + Q(r, 1, 1, 0) * fy + Q(r, 0, 1, 1) * fy + Q(r, 0, 1, 0) * fy;
        Ty[d] += Q(r, 1, 0, 1) * -fy + Q(r, 1, 0, 0) * -fy + Q(r, 0, 0, 1) * -fy + Q(r, 0, 0, 0) * -fy;
        Tz[d] = Q(r, 1, 1, 1) * fz + Q(r, 1, 0, 1) * fz + Q(r, 0, 1, 1) * fz + Q(r, 0, 0, 1) * fz;
        Tz[d] += Q(r, 1, 1, 0) * -fz + Q(r, 1, 0, 0) * -fz + Q(r, 0, 1, 0) * -fz + Q(r, 0, 0, 0) * -fz;
      }
      auto Tc = [&](int a, int b, int q) { return Q(E3_P, a, b, q) / (P.R * Q(E3_RHO, a, b, q)); };
      const T T_avg = avg[4] / (P.R * avg[0]);
      const T mu = T(visc_mu(P, (double)T_avg));
      const T kappa = mu * T(P.Cp) / T(P.prandtl);
      const T Txt = fx * (Tc(1, 1, 1) + Tc(1, 1, 0) + Tc(1, 0, 1) + Tc(1, 0, 0) - Tc(0, 1, 1) -
                          Tc(0, 1, 0) - Tc(0, 0, 1) - Tc(0, 0, 0));
      const T Tyt = fy * (Tc(1, 1, 1) + Tc(1, 1, 0) + Tc(0, 1, 1) + Tc(0, 1, 0) - Tc(1, 0, 1) -
                          Tc(1, 0, 0) - Tc(0, 0, 1) - Tc(0, 0, 0));
      const T Tzt = fz * (Tc(1, 1, 1) + Tc(1, 0, 1) + Tc(0, 1, 1) + Tc(0, 0, 1) - Tc(1, 1, 0) -
                          Tc(1, 0, 0) - Tc(0, 1, 0) - Tc(0, 0, 0));
      const T gv = Tx[0] + Ty[1] + Tz[2];
      const T txx = mu * (T(2) * Tx[0] - (T(2) / T(3)) * gv);
      const T tyy = mu * (T(2) * Ty[1] - (T(2) / T(3)) * gv);
      const T tzz = mu * (T(2) * Tz[2] - (T(2) / T(3)) * gv);
      const T txy = mu * (Ty[0] + Tx[1]), txz = mu * (Tx[2] + Tz[0]), tyz = mu * (Ty[2] + Tz[1]);
      const T* u = &avg[1];
      X[0] = txx; X[1] = txy; X[2] = txz;
      X[3] = u[0] * txx + u[1] * txy + u[2] * txz + kappa * Txt;
      Y[0] = txy; Y[1] = tyy; Y[2] = tyz;
      Y[3] = u[0] * txy + u[1] * tyy + u[2] * tyz + kappa * Tyt;
      Z[0] = txz; Z[1] = tyz; Z[2] = tzz;
      Z[3] = u[0] * txz + u[1] * tyz + u[2] * tzz + kappa * Tzt;
    }
    const T ax_ = -(dy * dz), ay_ = -(dx * dz), az_ = -(dx * dy);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      X[v] *= ax_;
      Y[v] *= ay_;
      Z[v] *= az_;
    }
  }

  // compute_dt fv.cc:366-383 (viscous term for the NS models)
  __device__ static T speed(const Phys<T>& P, const T* U, int64_t N, int64_t idx) {
    const T rho = U[idx], E = U[4 * N + idx];
    const T inv = T(1) / rho;
    T u[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[d] = U[(1 + d) * N + idx] * inv;
    const T m2 = U[N + idx] * U[N + idx] + U[2 * N + idx] * U[2 * N + idx] +
                 U[3 * N + idx] * U[3 * N + idx];
    const T p = (P.gamma - T(1)) * (E - T(0.5) * m2 / rho);
    T s = m_sqrt(vdot3(u, u)) + m_sqrt(P.gamma * p / rho);
    if (P.model == 11 || P.model == 12)
      s += T(visc_mu(P, (double)(p / (P.R * rho)))) / rho / P.hmin;
    return s;
  }
  __device__ static bool bad(const T* U, int64_t N, int64_t idx) {
    const T rho = U[idx];
    const T m2 = U[N + idx] * U[N + idx] + U[2 * N + idx] * U[2 * N + idx] +
                 U[3 * N + idx] * U[3 * N + idx];
    const T e = U[4 * N + idx] - T(0.5) * m2 / rho;
    return !(rho > T(0)) || !(e > T(0)) || !(e < T(INFINITY));
  }
};

// ================================================================ 3D tile
template <class PH, int TX, int TY, int TZ, int G, bool VISC>
struct Tile3 {
  using T = typename PH::TS;
  static constexpr int BX = TX + 2 * G, BY = TY + 2 * G, BZ = TZ + 2 * G;
  static constexpr int NB = BX * BY * BZ;
  static constexpr int NFX = (TX + 1) * TY * TZ, NFY = TX * (TY + 1) * TZ;
  static constexpr int NFZ = TX * TY * (TZ + 1), NF = NFX + NFY + NFZ;
  static constexpr int ROUNDS = (NF + 255) / 256;
  static constexpr int NN = (TX + 1) * (TY + 1) * (TZ + 1);
  static constexpr int NCELL = TX * TY * TZ;
  static constexpr int NV = PH::NV;
  static constexpr size_t REC_BYTES = (size_t)PH::NREC * NB * sizeof(T);
  static constexpr size_t FLUX_BYTES = (size_t)NV * NF * sizeof(T);
  static constexpr size_t NODE_BYTES = VISC ? (size_t)12 * NN * sizeof(T) : 0;
  static constexpr size_t LDS_BYTES =
      REC_BYTES > FLUX_BYTES + NODE_BYTES ? REC_BYTES : FLUX_BYTES + NODE_BYTES;
  static_assert(NN <= 256, "one node per lane");

  __device__ static int box(int x, int y, int z) { return (z * BY + y) * BX + x; }

  __device__ static void load(const StageArgs<T>& A, const Phys<T>& P, const GridDev& g, T* ld,
                              int tid, int x0, int y0, int z0) {
    for (int c = tid; c < NB; c += 256) {
      const int bx = c % BX, by = (c / BX) % BY, bz = c / (BX * BY);
      const int lx = bx - G, ly = by - G, lz = bz - G;
      const int out = (lx < 0 || lx >= TX) + (ly < 0 || ly >= TY) + (lz < 0 || lz >= TZ);
      if (out > (VISC ? 3 : 1)) continue;  // outside every face stencil
      int fx, fy, fz;
      const int sx = bc_map(x0 + lx, g.n[0], g.bc[0], g.bc[1], fx);
      const int sy = bc_map(y0 + ly, g.n[1], g.bc[2], g.bc[3], fy);
      const int sz = bc_map(z0 + lz, g.n[2], g.bc[4], g.bc[5], fz);
      const int64_t idx = ((int64_t)sz * g.n[1] + sy) * g.n[0] + sx;
      PH::load(P, A.U, g.N, idx, fx | (fy << 1) | (fz << 2), ld, NB, c);
    }
  }

  __device__ static void faces(const Phys<T>& P, const T* ld, int tid, T Fr[ROUNDS][NV],
                               unsigned& nfb) {
#pragma unroll
    for (int rd = 0; rd < ROUNDS; ++rd) {
      int f = rd * 256 + tid;
      if (rd * 256 + (tid & ~63) >= NF) break;  // whole wave past the end
      f = f < NF ? f : NF - 1;
      int base, stride, ax;
      if (f < NFX) {
        const int x = f % (TX + 1), y = (f / (TX + 1)) % TY, z = f / ((TX + 1) * TY);
        base = box(x, y + G, z + G);  // cells x .. x + 2G - 1 along x
        stride = 1;
        ax = 0;
      } else if (f < NFX + NFY) {
        const int q = f - NFX;
        const int x = q % TX, y = (q / TX) % (TY + 1), z = q / (TX * (TY + 1));
        base = box(x + G, y, z + G);
        stride = BX;
        ax = 1;
      } else {
        const int q = f - NFX - NFY;
        const int x = q % TX, y = (q / TX) % TY, z = q / (TX * TY);
        base = box(x + G, y + G, z);
        stride = BX * BY;
        ax = 2;
      }
      // lanes past the end evaluate the last face again (the whole wave
      // takes part in the matrix-core network); their count is not kept
      unsigned nf = 0;
      if constexpr (PH::FR_SELECT) {
        // constant indices into Fr: the compiler keeps this loop rolled for
        // the FP32-network faces, and Fr[rd] would live in scratch memory
        // (TGV 128^3 FP32 DSP-WENO +7%; the FP64 faces are faster without)
        T F[NV];
        PH::face(P, ld, NB, base, stride, ax, F, nf);
#pragma unroll
        for (int r2 = 0; r2 < ROUNDS; ++r2)
#pragma unroll
          for (int v = 0; v < NV; ++v) Fr[r2][v] = (r2 == rd) ? F[v] : Fr[r2][v];
      } else {
        PH::face(P, ld, NB, base, stride, ax, Fr[rd], nf);
      }
      if (rd * 256 + tid < NF) nfb += nf;
    }
  }

  // node (a, b, q) of the tile lies between cells a-1 and a (local) along x;
  // a node on an edge of two periodic boundaries counts twice, as in
  // compute_visc_residual's boundary loops (fv.cc:821-1156), unless
  // fix_edges; nodes on neumann / wall boundaries count once
  __device__ static void node(const Phys<T>& P, const GridDev& g, const T* ld, int tid, int x0,
                              int y0, int z0, T X[4], T Y[4], T Z[4]) {
    const int nx = tid % (TX + 1), ny = (tid / (TX + 1)) % (TY + 1), nz = tid / ((TX + 1) * (TY + 1));
    int c[2][2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int q = 0; q < 2; ++q) c[a][b][q] = box(nx - 1 + a + G, ny - 1 + b + G, nz - 1 + q + G);
    PH::node(P, ld, NB, c, X, Y, Z);
    const int gx = x0 + nx - 1, gy = y0 + ny - 1, gz = z0 + nz - 1;
    const bool ox = gx == -1 || gx == g.n[0] - 1, oy = gy == -1 || gy == g.n[1] - 1,
               oz = gz == -1 || gz == g.n[2] - 1;
    const int nb = ox + oy + oz;
    const int np = (ox && g.bc[0] == 0) + (oy && g.bc[2] == 0) + (oz && g.bc[4] == 0);
    if (nb == 2 && np == 2 && !P.fix_edges) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        X[v] *= T(2);
        Y[v] *= T(2);
        Z[v] *= T(2);
      }
    }
  }

  __device__ static void update(const StageArgs<T>& A, const Phys<T>& P, const GridDev& g,
                                const T* lf, const T* lnode, int tid, int x0, int y0, int z0,
                                double dt, T& smax, unsigned& nbad) {
    const int64_t N = g.N;
    const double factor = dt / P.vol_d;
    const T a = T(A.a), b = T(A.b), fb = T(factor * A.c);
    // U is read where the stage weighs it (b != 0), U^n where a != 0 or for the
    // classical RK4 forms: SSPRK3's first stage and rk1 (a = 0) skip U^n,
    // jameson_rk4 (b = 0) skips U
    const bool rd_u = A.r2 == 0 && A.b != 0.0, rd_u0 = A.r2 != 0 || A.a != 0.0;
    for (int c = tid; c < NCELL; c += 256) {
      const int lx = c % TX, ly = (c / TX) % TY, lz = c / (TX * TY);
      const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
      if (gx >= g.n[0] || gy >= g.n[1] || gz >= g.n[2]) continue;
      const int64_t idx = ((int64_t)gz * g.n[1] + gy) * g.n[0] + gx;
      const int fxl = (lz * TY + ly) * (TX + 1) + lx;
      const int fyl = NFX + (lz * (TY + 1) + ly) * TX + lx;
      const int fzl = NFX + NFY + (lz * TY + ly) * TX + lx;
      // every state read of the cell is issued before its first store (Uout
      // may be U^n: a read after a store would wait for it); the FP64 FAM 0
      // kernel keeps its reads next to their use (register budget)
      T u[NV], u0[NV], un[NV];
      if constexpr (PH::REG_SLACK) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          u[v] = rd_u ? A.U[v * N + idx] : T(0);
          u0[v] = rd_u0 ? A.U0[v * N + idx] : T(0);
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        T r = -lf[v * NF + fxl];
        r += lf[v * NF + fxl + 1];
        r -= lf[v * NF + fyl];
        r += lf[v * NF + fyl + TX];
        r -= lf[v * NF + fzl];
        r += lf[v * NF + fzl + TX * TY];
        if (VISC && v > 0) {
          // the 8 corner nodes: (+1/4) on the upper side of a direction
#pragma unroll
          for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
              for (int aa = 0; aa < 2; ++aa) {
                const int nd = ((lz + q) * (TY + 1) + (ly + bb)) * (TX + 1) + (lx + aa);
                const T sx = aa ? T(0.25) : T(-0.25), sy = bb ? T(0.25) : T(-0.25),
                        sz = q ? T(0.25) : T(-0.25);
                r += lnode[(v - 1) * NN + nd] * sx + lnode[(4 + v - 1) * NN + nd] * sy +
                     lnode[(8 + v - 1) * NN + nd] * sz;
              }
        }
        if constexpr (!PH::REG_SLACK) {
          u[v] = A.U[v * N + idx];
          u0[v] = A.U0[v * N + idx];
        }
        un[v] = rk_update<NV>(A, v * N + idx, u[v], u0[v], r, a, b, fb, factor);
        A.Uout[v * N + idx] = un[v];
      }
      if (A.last) {
        // from the registers of the new state (a one-cell view)
        smax = m_max(smax, PH::speed(P, un, 1, 0));
        nbad += PH::bad(un, 1, 0) ? 1u : 0u;
      }
    }
  }
};

#ifndef TECNO_HOST_BUILD
template <class PH, int TX, int TY, int TZ, int G, bool VISC>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PH::MIN_WAVES3)))
void stage3_kernel(StageArgs<typename PH::TS> A, Phys<typename PH::TP> P, GridDev g) {
  using TL = Tile3<PH, TX, TY, TZ, G, VISC>;
  using T = typename PH::TS;
  __shared__ __attribute__((aligned(16))) unsigned char lds[TL::LDS_BYTES];
  __shared__ double red[4];
  if (!A.ctl->active) return;
  const int tid = threadIdx.x;
  int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
  if constexpr (PH::REG_SLACK) tile_of_block(bx, by, bz);
  const int x0 = bx * TX, y0 = by * TY, z0 = bz * TZ;
  T* ld = (T*)lds;
  TL::load(A, P, g, ld, tid, x0, y0, z0);
  __syncthreads();
  T Fr[TL::ROUNDS][TL::NV];
  unsigned nfb = 0;
  TL::faces(P, ld, tid, Fr, nfb);
  if (nfb) atomicAdd(&A.ctl->nfallback, (unsigned long long)nfb);
  T X[4], Y[4], Z[4];
  if (VISC && tid < TL::NN) TL::node(P, g, ld, tid, x0, y0, z0, X, Y, Z);
  __syncthreads();
  T* lf = ld;
  T* lnode = ld + (size_t)TL::NV * TL::NF;
#pragma unroll
  for (int rd = 0; rd < TL::ROUNDS; ++rd) {
    const int f = rd * 256 + tid;
    if (f < TL::NF)
#pragma unroll
      for (int v = 0; v < TL::NV; ++v) lf[v * TL::NF + f] = Fr[rd][v];
  }
  if (VISC && tid < TL::NN)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      lnode[v * TL::NN + tid] = X[v];
      lnode[(4 + v) * TL::NN + tid] = Y[v];
      lnode[(8 + v) * TL::NN + tid] = Z[v];
    }
  __syncthreads();
  T smax = T(0);
  unsigned nbad = 0;
  TL::update(A, P, g, lf, lnode, tid, x0, y0, z0, A.ctl->dt, smax, nbad);
  if (A.last) {
    double s = (double)smax;
    for (int off = 32; off > 0; off >>= 1) s = fmax(s, __shfl_xor(s, off));
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    if (tid == 0) {
      double m = red[0];
      for (int w = 1; w < 4; ++w) m = fmax(m, red[w]);
      atomic_max_pos(&A.ctl->smax_bits, m);
    }
    if (nbad) atomicAdd(&A.ctl->nbad, (unsigned long long)nbad);
  }
}
#endif

}  // namespace tecno
