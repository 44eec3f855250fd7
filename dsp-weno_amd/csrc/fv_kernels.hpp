// fv_kernels.hpp — fused SSPRK3 stage kernels for gfx950 and their launchers.
//
// One launch advances one Runge-Kutta stage of the whole grid
// (FV::compute_residual + FV::update_solution, enflo/src/fv.cc:591-767 and
// 1163-1175; enflo_scalar/src_scalar/fv.cc:317-476).  A workgroup owns a
// TX x TY tile of cells:
//   1. it reads the conserved state of the tile and its G-cell cross-shaped
//      halo from HBM (structure of arrays, x fastest; ghost cells are produced
//      by index mapping: periodic wrap, Neumann copy, wall mirror), converts
//      each cell to a record (primitive, entropy variables ...) in LDS;
//   2. every lane evaluates one face flux per round (x faces and y faces in
//      one flat, wave-uniform loop) from LDS and keeps it in registers;
//   3. the fluxes go to LDS (aliasing the records) and each cell accumulates
//      its residual in the reference's order (-Fx(i-1/2) + Fx(i+1/2)
//      - Fy(j-1/2) + Fy(j+1/2)), applies the SSPRK3 update and writes back;
//      the last stage also reduces max(|u| + c) for the next CFL step.
// Faces on tile edges are evaluated by both neighbouring tiles; the tile
// shapes are chosen so that the face count fills whole 256-lane rounds.
#pragma once
#include "fv_physics.hpp"

namespace tecno {

// ---------------------------------------------------------------- control
struct Ctl {
  double dt, time, t_save, t_int, final_time;
  unsigned long long smax_bits;  // max wave speed of the current state (bits of a double >= 0)
  int iter, max_iter, active, pad;
  unsigned long long nbad;       // cells with non-finite or non-positive rho / p
  unsigned long long nfallback;  // DSP-WENO components replaced by SP-WENO
};

struct GridDev {
  int n[3];    // cells held per direction (1 for collapsed axes)
  int bc[6];
  int64_t N;
};

// coefficients of one Runge-Kutta stage (update_solution, enflo/src/fv.cc:1160-1198,
// enflo_scalar/src_scalar/fv.cc:459-499):
//   rk1 / ssprk3:  u = b*u + a*u0 - (dt/vol)*c*res, c = b (ext_constants.h:7-8)
//   jameson_rk4:   a = 1, b = 0, c = jameson_rks[k] (ext_constants.h:9)
//   rk4:           R2 accumulates w*res with w = 1/rk4_rks[k]; stages 0-2
//                  take u = u0 - (dt/vol)*rk4_rks[k+1]*res (a = 1, b = 0),
//                  the last u = u0 - (dt/vol)*R2/6
struct StageCoef {
  double a, b, c;
  double w;   // weight of res in the classical RK4 accumulator
  int last;   // reduce the CFL wave speed of Uout
  int r2;     // classical RK4 accumulator: 0 unused, 1 first stage, 2 middle, 3 last
};

template <typename T>
struct StageArgs {
  const T* U;    // U^(k-1), NV x N
  const T* U0;   // U^n
  T* Uout;       // U^(k)
  Ctl* ctl;
  double a, b, c;
  int last;
  T* R2;         // classical RK4 residual accumulator, NV x N (r2 != 0)
  double w;
  int r2;
};

// the stage's new value of one variable: r is the cell's residual (the
// reference's residual[i][j][k] for that variable), factor = dt/vol
template <int NV, typename TS>
__device__ __forceinline__ TS rk_update(const StageArgs<TS>& A, int64_t at, TS u, TS u0, TS r,
                                        TS a, TS b, TS fb, double factor) {
  if (A.r2 == 0) {
    TS un = u * b;
    un += (u0 * a - fb * r);
    return un;
  }
  // residual2 += residual/rk4_rks[rk] (scalar fv.cc:492) and
  // residual2.sadd(residual, 1/rk4_rks[rk]) (systems fv.cc:1191): the
  // weights 1 and 2 make both exact
  TS r2 = r * TS(A.w);
  if (A.r2 != 1) r2 = A.R2[at] + r2;
  if (A.r2 != 3) {
    A.R2[at] = r2;
    return u0 - fb * r;
  }
  // last stage: the scalar solver rounds old - factor*R2/6 (fv.cc:496),
  // the systems solver old + R2*(-factor/6) (fv.cc:1195)
  if (NV == 1) return u0 - TS(factor) * r2 / TS(6);
  return u0 + r2 * TS(-factor / 6.0);
}

// map a (possibly ghost) coordinate along one axis to a source cell; returns
// the sign to apply to the normal momentum (-1 for a wall mirror)
__device__ __forceinline__ int bc_map(int c, int n, int bcl, int bcr, int& flip) {
  flip = 0;
  if (c < 0) {
    if (bcl == 0) c += n;                 // periodic
    else if (bcl == 3) { c = -1 - c; flip = 1; }  // wall: mirror
    else c = 0;                           // neumann: first interior cell
  } else if (c >= n) {
    if (bcr == 0) c -= n;
    else if (bcr == 3) { c = 2 * n - 1 - c; flip = 1; }
    else c = n - 1;
  }
  // tiles that overhang a small domain: keep every access inside the grid
  c = c < 0 ? 0 : (c >= n ? n - 1 : c);
  return c;
}

// ================================================================ stage kernel
template <int TX, int TY, int DIMS, int G>
struct TileShape {
  static constexpr int BX = TX + 2 * G;
  static constexpr int BY = (DIMS >= 2) ? TY + 2 * G : 1;
  static constexpr int NB = BX * BY;
  static constexpr int NFX = (TX + 1) * TY;
  static constexpr int NFY = (DIMS >= 2) ? TX * (TY + 1) : 0;
  static constexpr int NF = NFX + NFY;
  static constexpr int ROUNDS = (NF + 255) / 256;
  static constexpr int NCELL = TX * TY;
  // distance-2 cell pairs (c, c+2) along each axis: the pairs (0,2) and (1,3)
  // of the 4-cell stencil of a face are the pairs p and p+1 of its row
  // (x faces) or column (y faces), so neighbouring faces share one
  static constexpr int NPX = (TX + 2) * TY;
  static constexpr int NPY = (DIMS >= 2) ? TX * (TY + 2) : 0;
  static constexpr int NP = NPX + NPY;
};

// PH::PAIRS: the face flux takes the two-point fluxes of the distance-2 pairs
// of its stencil from LDS (computed once per pair) instead of evaluating them
template <class PH, class = void>
struct has_pairs {
  static constexpr bool value = false;
};
template <class PH>
struct has_pairs<PH, decltype((void)PH::PAIRS)> {
  static constexpr bool value = PH::PAIRS;
};

#ifndef TECNO_HOST_BUILD
__device__ __forceinline__ double atomic_max_pos(unsigned long long* addr, double v) {
  // non-negative doubles order like their bit patterns (NaN sorts above +inf)
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return __longlong_as_double((long long)atomicMax(addr, b));
}
#endif

// The three phases of a stage on one tile.  They are separate functions so
// that the same code also runs in a host emulation of the workgroup.
template <class PH, int DIMS, int TX, int TY, int G>
struct Tile {
  using SH = TileShape<TX, TY, DIMS, G>;
  using TS = typename PH::TS;
  using TF = typename PH::TF;
  using TP = typename PH::TP;
  static constexpr int NB = SH::NB, NF = SH::NF, NV = PH::NV;
  static constexpr bool PAIRS = has_pairs<PH>::value;
  static constexpr int NP = SH::NP;
  static constexpr size_t REC_BYTES = PH::rec_bytes(NB);
  // pair fluxes follow the records (TF[NV][NP]); the face fluxes overwrite both
  static constexpr size_t PAIR_OFF = (REC_BYTES + 15) / 16 * 16;
  static constexpr size_t PAIR_BYTES = PAIRS ? (size_t)NV * NP * sizeof(TF) : 0;
  static constexpr size_t IN_BYTES = PAIRS ? PAIR_OFF + PAIR_BYTES : REC_BYTES;
  static constexpr size_t FLUX_BYTES = (size_t)NV * NF * sizeof(TF);
  static constexpr size_t LDS_BYTES = IN_BYTES > FLUX_BYTES ? IN_BYTES : FLUX_BYTES;
  static constexpr int ROUNDS = SH::ROUNDS;

  // 1. cell records of the tile and its cross-shaped halo
  __device__ static void load(const StageArgs<TS>& A, const Phys<TP>& P, const GridDev& g,
                              void* lds, int tid, int bxi, int byi) {
    const int x0 = bxi * TX, y0 = byi * TY;
    for (int c = tid; c < NB; c += 256) {
      const int bx = c % SH::BX, by = c / SH::BX;
      const int lx = bx - G, ly = (DIMS >= 2) ? by - G : 0;
      const bool inx = (lx >= 0 && lx < TX), iny = (ly >= 0 && ly < TY);
      if (!inx && !iny) continue;  // corner of the box: not in any stencil
      int fx, fy = 0;
      const int sx = bc_map(x0 + lx, g.n[0], g.bc[0], g.bc[1], fx);
      int sy = 0;
      if (DIMS >= 2) sy = bc_map(y0 + ly, g.n[1], g.bc[2], g.bc[3], fy);
      const int64_t idx = (int64_t)sy * g.n[0] + sx;
      PH::load(P, A.U, g.N, idx, fx, fy, lds, NB, c);
    }
  }

  // 1b. two-point fluxes of the distance-2 pairs (PH::PAIRS only): x pairs
  // (TX+2 per row) then y pairs (TX per row of TY+2 rows), both row by row,
  // one pair per lane per round: neighbouring lanes read neighbouring cells,
  // so the FP64 record reads are free of LDS bank conflicts
  __device__ static void pairs(const Phys<TP>& P, void* lds, int tid) {
    if constexpr (PAIRS) {
      TF* out = (TF*)((unsigned char*)lds + PAIR_OFF);
      for (int q = tid; q < NP; q += 256) {
        int ca, stride, ax;
        if (q < SH::NPX) {
          const int fy = q / (TX + 2), px = q - fy * (TX + 2);
          ca = (DIMS >= 2 ? (fy + G) * SH::BX : 0) + px;
          stride = 1;
          ax = 0;
        } else {
          const int r = q - SH::NPX;
          const int py = r / TX, fx = r - py * TX;
          ca = py * SH::BX + fx + G;
          stride = SH::BX;
          ax = 1;
        }
        PH::pair(P, lds, NB, ca, ca + 2 * stride, ax, out, NP, q);
      }
    }
  }

  // 2. face fluxes, one face per lane per round, kept in registers
  __device__ static void faces(const Phys<TP>& P, const void* lds, int tid, TF Fr[ROUNDS][NV],
                               unsigned& nfb) {
#pragma unroll
    for (int rd = 0; rd < ROUNDS; ++rd) {
      int f = rd * 256 + tid;
      if (rd * 256 + (tid & ~63) >= NF) break;  // whole wave past the end
      f = f < NF ? f : NF - 1;
      int base, stride, ax, pq, ps;
      if (f < SH::NFX) {
        const int fy = f / (TX + 1), fx = f - fy * (TX + 1);
        base = (DIMS >= 2 ? (fy + G) * SH::BX : 0) + fx;
        stride = 1;
        ax = 0;
        pq = fy * (TX + 2) + fx;
        ps = 1;
      } else {
        const int q = f - SH::NFX;
        const int fy = q / TX, fx = q - fy * TX;
        base = fy * SH::BX + fx + G;
        stride = SH::BX;
        ax = 1;
        pq = SH::NPX + fy * TX + fx;
        ps = TX;
      }
      unsigned c = 0;
      if constexpr (PAIRS)
        PH::face(P, lds, NB, base, stride, ax, Fr[rd], c,
                 (const TF*)((const unsigned char*)lds + PAIR_OFF), NP, pq, ps);
      else
        PH::face(P, lds, NB, base, stride, ax, Fr[rd], c);
      if (rd * 256 + tid < NF) nfb += c;
    }
  }

  // face fluxes -> LDS (after the records are dead)
  __device__ static void store(void* ldsv, int tid, const TF Fr[ROUNDS][NV]) {
    TF* lds = (TF*)ldsv;
#pragma unroll
    for (int rd = 0; rd < ROUNDS; ++rd) {
      const int f = rd * 256 + tid;
      if (f < NF) {
#pragma unroll
        for (int v = 0; v < NV; ++v) lds[v * NF + f] = Fr[rd][v];
      }
    }
  }

  // 3. residual in the reference's order, SSPRK3 update, CFL speed
  __device__ static void update(const StageArgs<TS>& A, const Phys<TP>& P, const GridDev& g,
                                const void* ldsv, int tid, int bxi, int byi, double dt, TS& smax,
                                unsigned& nbad) {
    const TF* lds = (const TF*)ldsv;
    const int x0 = bxi * TX, y0 = byi * TY;
    const int64_t N = g.N;
    const double factor = dt / P.vol_d;
    const TS a = TS(A.a), b = TS(A.b), fb = TS(factor * A.c);
    // U is read where the stage weighs it (b != 0), U^n where a != 0 or for the
    // classical RK4 forms: SSPRK3's first stage and rk1 (a = 0) skip U^n,
    // jameson_rk4 (b = 0) skips U
    const bool rd_u = A.r2 == 0 && A.b != 0.0, rd_u0 = A.r2 != 0 || A.a != 0.0;
    for (int c = tid; c < SH::NCELL; c += 256) {
      const int ly = c / TX, lx = c - ly * TX;
      const int gx = x0 + lx, gy = y0 + ly;
      if (gx >= g.n[0] || (DIMS >= 2 && gy >= g.n[1])) continue;
      const int64_t idx = (int64_t)gy * g.n[0] + gx;
      const int fxl = ly * (TX + 1) + lx;
      const int fyl = SH::NFX + ly * TX + lx;
      // every state read of the cell is issued before its first store: Uout
      // may be U^n (the last SSPRK3 stage writes in place), so a read after a
      // store would wait for it, one memory round trip per variable
      TS u[NV], u0[NV], un[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        u[v] = rd_u ? A.U[v * N + idx] : TS(0);
        u0[v] = rd_u0 ? A.U0[v * N + idx] : TS(0);
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        TS r = -TS(lds[v * NF + fxl]);
        r += TS(lds[v * NF + fxl + 1]);
        if (DIMS >= 2) {
          r -= TS(lds[v * NF + fyl]);
          r += TS(lds[v * NF + fyl + TX]);
        }
        un[v] = rk_update<NV>(A, v * N + idx, u[v], u0[v], r, a, b, fb, factor);
        A.Uout[v * N + idx] = un[v];
      }
      if (A.last) {
        // wave speed and admissibility of the new state from its registers
        // (a one-cell view: element v of un is variable v)
        smax = m_max(smax, PH::speed(P, un, 1, 0));
        nbad += PH::bad(un, 1, 0) ? 1u : 0u;
      }
    }
  }
};

// Tile of a workgroup.  The dispatcher deals workgroups out round-robin
// over the chip's 8 XCDs (workgroup i to XCD i mod 8), each with its own L2,
// so neighbouring tiles of the launch grid land on different XCDs and the
// halo cells that two tiles share are fetched from HBM by both.  The linear
// workgroup id is remapped so that each XCD takes chunks of XCD_ROWS whole
// tile rows (2D) or tile planes (3D), the chunks dealt out to the XCDs in
// turn: inside a chunk a tile's left and lower neighbours were loaded
// shortly before by the same XCD and the halo they share is an L2 hit, while
// every XCD still gets rows from all over the grid (the cost of a tile
// depends on the flow in it, so contiguous eighths of the grid would leave
// XCDs idle).  Tiles past the last whole round of chunks are split into 8
// contiguous runs.  A bijection of [0, gx gy gz) for any grid (host test).
#ifndef TECNO_XCD_ROWS
#define TECNO_XCD_ROWS 2
#endif
__device__ __forceinline__ unsigned xcd_tile(unsigned lin, unsigned gx, unsigned gy, unsigned gz) {
  if (TECNO_XCD_ROWS <= 0) return lin;
  const unsigned T = gx * gy * gz;
  const unsigned C = (unsigned)TECNO_XCD_ROWS * (gz > 1 ? gx * gy : gx);
  const unsigned xcd = lin % 8u, k = lin / 8u, q = T / 8u, r = T % 8u;
  const unsigned J = q / C;  // whole chunks that every XCD takes
  if (k < J * C) return ((k / C) * 8u + xcd) * C + k % C;
  const unsigned k2 = k - J * C, q2 = q - J * C;
  return 8u * J * C + (xcd < r ? xcd * (q2 + 1u) : r * (q2 + 1u) + (xcd - r) * q2) + k2;
}

#ifndef TECNO_HOST_BUILD
__device__ __forceinline__ void tile_of_block(int& bx, int& by, int& bz) {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned t = xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx, gy, gridDim.z);
  bx = (int)(t % gx);
  by = (int)((t / gx) % gy);
  bz = (int)(t / (gx * gy));
}

template <class PH, int DIMS, int TX, int TY, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(PH::MIN_WAVES)))
void stage_kernel(StageArgs<typename PH::TS> A, Phys<typename PH::TP> P, GridDev g) {
  using TL = Tile<PH, DIMS, TX, TY, G>;
  using TS = typename PH::TS;
  __shared__ __attribute__((aligned(16))) unsigned char lds[TL::LDS_BYTES];
  __shared__ double red[4];
  if (!A.ctl->active) return;
  const int tid = threadIdx.x;
  int bx, by, bz;
  tile_of_block(bx, by, bz);
  TL::load(A, P, g, lds, tid, bx, by);
  __syncthreads();
  if constexpr (TL::PAIRS) {
    TL::pairs(P, lds, tid);
    __syncthreads();
  }
  typename PH::TF Fr[TL::ROUNDS][TL::NV];
  unsigned nfb = 0;
  TL::faces(P, lds, tid, Fr, nfb);
  if (nfb) atomicAdd(&A.ctl->nfallback, (unsigned long long)nfb);
  __syncthreads();
  TL::store(lds, tid, Fr);
  __syncthreads();
  TS smax = TS(0);
  unsigned nbad = 0;
  TL::update(A, P, g, lds, tid, bx, by, A.ctl->dt, smax, nbad);
  if (A.last) {
    // wave max, then across the 4 waves
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
