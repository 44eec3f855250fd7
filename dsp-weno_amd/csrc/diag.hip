// diag.hip — FP64 diagnostics of the systems solver: nodal vorticity and
// enstrophy, KE_diss_pr and KE_diss_stress (the globals.dat columns of
// FV::compute_global and the omg####.vtk field of FV::output).
//
// Ghost cells follow the reference's apply_bc (enflo/src/fv.cc:121-361): a
// cell outside the grid along one axis takes that axis's boundary map
// (periodic wrap, neumann copy of the first interior cell, wall mirror with
// the normal velocity negated); a cell outside along two or three axes (an
// edge or corner ghost) is filled as if every boundary were periodic
// (fv.cc:317-359), whatever the boundary types.  The viscous node fluxes of
// KE_diss_stress take their cells as the stage kernel does (boundary maps
// composed per axis), which is the same on the periodic grids the reference
// allows for Navier-Stokes.
#include "../../include/tecno.h"
#include "diag.hpp"
#include "fv_euler3.hpp"

namespace tecno {
namespace {

constexpr int NB_THREADS = 64;

struct Prim {
  double rho, u[3], p;
};

// con_to_prim (fv.cc:388-400) of held cell idx; 2D states hold rho, m_x,
// m_y, E
__device__ Prim prim_at(const double* U, int64_t N, int64_t idx, int dims, double gamma) {
  Prim q;
  const int nv = dims == 3 ? 5 : 4;
  const double rho = U[idx];
  const double inv = 1.0 / rho;
  double m[3] = {U[N + idx], U[2 * N + idx], dims == 3 ? U[3 * N + idx] : 0.0};
  q.rho = rho;
  for (int d = 0; d < 3; ++d) q.u[d] = m[d] * inv;
  const double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
  q.p = (gamma - 1.0) * (U[(nv - 1) * N + idx] - 0.5 * m2 / rho);
  return q;
}

__device__ int wrap(int c, int n) { return c < 0 ? c + n : (c >= n ? c - n : c); }

// boundary map of a ghost one cell outside an axis of n cells
__device__ int side(int c, int n, int bcl, int bcr, int& flip) {
  flip = 0;
  if (c < 0) {
    if (bcl == TECNO_BC_PERIODIC) return c + n;
    flip = bcl == TECNO_BC_WALL;
    return 0;
  }
  if (c >= n) {
    if (bcr == TECNO_BC_PERIODIC) return c - n;
    flip = bcr == TECNO_BC_WALL;
    return n - 1;
  }
  return c;
}

// held index and flipped-velocity mask of cell (ci, cj, ck); zghost marks a
// 2D cell that the reference stores in a z ghost layer.  ref_edges: edge and
// corner ghosts periodic (apply_bc); otherwise the per-axis maps compose.
__device__ int64_t lookup(const DiagGrid& g, int ci, int cj, int ck, bool zghost, bool ref_edges,
                          int& flips) {
  const bool ox = ci < 0 || ci >= g.n[0];
  const bool oy = g.dims >= 2 && (cj < 0 || cj >= g.n[1]);
  const bool oz = g.dims == 3 ? (ck < 0 || ck >= g.n[2]) : zghost;
  flips = 0;
  if (ref_edges && (int)ox + (int)oy + (int)oz >= 2) {
    ci = wrap(ci, g.n[0]);
    cj = g.dims >= 2 ? wrap(cj, g.n[1]) : 0;
    ck = g.dims == 3 ? wrap(ck, g.n[2]) : 0;
  } else {
    int f;
    ci = side(ci, g.n[0], g.bc[0], g.bc[1], f);
    flips |= f;
    if (g.dims >= 2) {
      cj = side(cj, g.n[1], g.bc[2], g.bc[3], f);
      flips |= f << 1;
    }
    if (g.dims == 3) {
      ck = side(ck, g.n[2], g.bc[4], g.bc[5], f);
      flips |= f << 2;
    } else {
      ck = 0;  // a z ghost of a 2D grid holds its interior cell (w = 0)
    }
  }
  return ((int64_t)ck * g.n[1] + cj) * g.n[0] + ci;
}

// node weight of find_node_weights (find_node_weights.cc:19-66): halved on
// every periodic boundary plane the node lies on
__device__ double axis_weight(int I, int n, int bcl, int bcr) {
  double w = 1.0;
  if (I == 0 && bcl == TECNO_BC_PERIODIC) w /= 2.0;
  if (I == n && bcr == TECNO_BC_PERIODIC) w /= 2.0;
  return w;
}

struct NodeGrad {
  double Tx[3], Ty[3], Tz[3];  // derivatives of the velocity along x, y, z
  double den;                  // 8-cell average of the density
};

// nodal_avg_and_vel_derivatives (nodal_avg_and_der.cc:16-82) for the node
// whose cells are c[a][b][q] = cell (I-1+a, J-1+b, K-1+q), in the
// reference's order of terms; den as compute_global.cc:42-50 sums it
__device__ NodeGrad node_grad(const Prim c[2][2][2], const double f[3]) {
  NodeGrad r;
  const double fx = f[0], fy = f[1], fz = f[2];
  for (int d = 0; d < 3; ++d) {
    auto V = [&](int a, int b, int q) { return c[a][b][q].u[d]; };
    r.Tx[d] = V(1, 1, 1) * fx + V(1, 1, 0) * fx + V(1, 0, 1) * fx + V(1, 0, 0) * fx;
    r.Tx[d] += V(0, 1, 1) * -fx + V(0, 1, 0) * -fx + V(0, 0, 1) * -fx + V(0, 0, 0) * -fx;
    r.Ty[d] = V(1, 1, 1) * fy + V(1, 1, 0) * fy + V(0, 1, 1) * fy + V(0, 1, 0) * fy;
    r.Ty[d] += V(1, 0, 1) * -fy + V(1, 0, 0) * -fy + V(0, 0, 1) * -fy + V(0, 0, 0) * -fy;
    r.Tz[d] = V(1, 1, 1) * fz + V(1, 0, 1) * fz + V(0, 1, 1) * fz + V(0, 0, 1) * fz;
    r.Tz[d] += V(1, 1, 0) * -fz + V(1, 0, 0) * -fz + V(0, 1, 0) * -fz + V(0, 0, 0) * -fz;
  }
  r.den = (1.0 / 8.0) * (c[0][0][0].rho + c[1][0][0].rho + c[0][1][0].rho + c[0][0][1].rho +
                         c[1][1][0].rho + c[1][0][1].rho + c[0][1][1].rho + c[1][1][1].rho);
  return r;
}

__device__ void gather(const DiagGrid& g, const double* U, double gamma, int I, int J, int K,
                       int zlayer, Prim c[2][2][2]) {
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int q = 0; q < 2; ++q) {
        const bool zg = g.dims == 2 && ((zlayer == 0 && q == 0) || (zlayer == 2 && q == 1));
        int fl;
        const int64_t idx = lookup(g, I - 1 + a, J - 1 + b, K - 1 + q, zg, true, fl);
        Prim p = prim_at(U, g.N, idx, g.dims, gamma);
        for (int d = 0; d < 3; ++d)
          if (fl & (1 << d)) p.u[d] = -p.u[d];
        c[a][b][q] = p;
      }
}

template <int DIMS, bool NS>
__global__ __launch_bounds__(NB_THREADS) void k_nodes(DiagGrid g, const double* U, double gamma,
                                                      Phys<double> P, double* omega,
                                                      double* part) {
  using PH = Euler3Ph<double, 2>;
  __shared__ double rec[NS ? E3_NREC * NB_THREADS * 8 : 1];
  __shared__ double red[NB_THREADS / 64][4];
  const int nx = g.n[0] + 1, ny = g.n[1] + 1;
  const int nz = DIMS == 3 ? g.n[2] + 1 : 3;
  const int64_t NN = (int64_t)nx * ny * nz;
  const int64_t t = (int64_t)blockIdx.x * NB_THREADS + threadIdx.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const double f[3] = {1.0 / (4.0 * g.h[0]), 1.0 / (4.0 * g.h[1]), 1.0 / (4.0 * g.h[2])};
  if (t < NN) {
    const int I = (int)(t % nx), J = (int)((t / nx) % ny), K = (int)(t / ((int64_t)nx * ny));
    Prim c[2][2][2];
    gather(g, U, gamma, I, J, DIMS == 3 ? K : 0, DIMS == 3 ? -1 : K, c);
    const NodeGrad r = node_grad(c, f);
    // vorticity[i][j][k] = (Ty.z - Tz.y, Tz.x - Tx.z, Tx.y - Ty.x), find_vort.cc:22-24
    const double w[3] = {r.Ty[2] - r.Tz[1], r.Tz[0] - r.Tx[2], r.Tx[1] - r.Ty[0]};
    if (omega)
      for (int d = 0; d < 3; ++d) omega[d * NN + t] = w[d];
    double wt = axis_weight(I, g.n[0], g.bc[0], g.bc[1]) * axis_weight(J, g.n[1], g.bc[2], g.bc[3]);
    if (DIMS == 3) wt *= axis_weight(K, g.n[2], g.bc[4], g.bc[5]);
    const double ens = r.den * (w[0] * w[0] + w[1] * w[1] + w[2] * w[2]) * wt;
    acc[DIMS == 3 ? 0 : K] = ens;
    if constexpr (NS && DIMS == 3) {
      // viscous node fluxes from the 8 cells' records (vflux_ES.cc / vflux_KE.cc)
      double* ld = rec;
      const int NB = NB_THREADS * 8;
      int cidx[2][2][2];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
          for (int q = 0; q < 2; ++q) {
            const int slot = threadIdx.x * 8 + a * 4 + b * 2 + q;
            int fl;
            const int64_t idx = lookup(g, I - 1 + a, J - 1 + b, K - 1 + q, false, false, fl);
            PH::load(P, U, g.N, idx, fl, ld, NB, slot);
            cidx[a][b][q] = slot;
          }
      double X[4], Y[4], Z[4];
      PH::node(P, ld, NB, cidx, X, Y, Z);
      // an edge node (two coordinates on periodic boundary planes) is visited
      // by two of the reference's boundary loops; the last two corners take
      // their derivatives and weight at node (I, Nz, Nz) (fv.cc:1143-1156)
      const int ox = I == 0 || I == g.n[0], oy = J == 0 || J == g.n[1], oz = K == 0 || K == g.n[2];
      const int np = (ox && g.bc[0] == TECNO_BC_PERIODIC) + (oy && g.bc[2] == TECNO_BC_PERIODIC) +
                     (oz && g.bc[4] == TECNO_BC_PERIODIC);
      const double mult = (ox + oy + oz == 2 && np == 2 && !g.fix_edges) ? 2.0 : 1.0;
      NodeGrad rg = r;
      double wg = wt;
      if (!g.fix_edges && J == g.n[1] && K == g.n[2] && (I == 0 || I == g.n[0]) &&
          g.n[2] <= g.n[1]) {
        Prim c2[2][2][2];
        gather(g, U, gamma, I, g.n[2], g.n[2], -1, c2);
        rg = node_grad(c2, f);
        wg = axis_weight(I, g.n[0], g.bc[0], g.bc[1]) * axis_weight(g.n[2], g.n[1], g.bc[2], g.bc[3]) *
             axis_weight(g.n[2], g.n[2], g.bc[4], g.bc[5]);
      }
      const double tx = rg.Tx[0] * X[0] + rg.Tx[1] * X[1] + rg.Tx[2] * X[2];
      const double ty = rg.Ty[0] * Y[0] + rg.Ty[1] * Y[1] + rg.Ty[2] * Y[2];
      const double tz = rg.Tz[0] * Z[0] + rg.Tz[1] * Z[1] + rg.Tz[2] * Z[2];
      acc[3] = -mult * ((tx * g.h[0] + ty * g.h[1] + tz * g.h[2]) * wg);
    }
  }
  for (int s = 0; s < 4; ++s) {
    double x = acc[s];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][s] = x;
  }
  __syncthreads();
  if (threadIdx.x < 4) part[(int64_t)blockIdx.x * 4 + threadIdx.x] = red[0][threadIdx.x];
}

__global__ __launch_bounds__(256) void k_ke_diss_pr(int dims, int64_t N, const double* U,
                                                    const double* R, double* part) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double rho = U[i], inv = 1.0 / rho;
    double u[3] = {U[N + i] * inv, U[2 * N + i] * inv, dims == 3 ? U[3 * N + i] * inv : 0.0};
    double rm[3] = {R[N + i], R[2 * N + i], dims == 3 ? R[3 * N + i] : 0.0};
    acc -= 0.5 * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * R[i] +
           (u[0] * rm[0] + u[1] * rm[1] + u[2] * rm[2]);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

hipError_t diag_nodes(hipStream_t s, const DiagGrid& g, const double* U, double gamma,
                      const Phys<double>& P, int ns, double* omega, double* part, int* nblk) {
  const int64_t nx = g.n[0] + 1, ny = g.n[1] + 1, nz = g.dims == 3 ? g.n[2] + 1 : 3;
  const int64_t NN = nx * ny * nz;
  const unsigned nb = (unsigned)((NN + NB_THREADS - 1) / NB_THREADS);
  if (g.dims == 3) {
    if (ns)
      hipLaunchKernelGGL((k_nodes<3, true>), dim3(nb), dim3(NB_THREADS), 0, s, g, U, gamma, P, omega,
                         part);
    else
      hipLaunchKernelGGL((k_nodes<3, false>), dim3(nb), dim3(NB_THREADS), 0, s, g, U, gamma, P, omega,
                         part);
  } else {
    hipLaunchKernelGGL((k_nodes<2, false>), dim3(nb), dim3(NB_THREADS), 0, s, g, U, gamma, P, omega,
                       part);
  }
  *nblk = (int)nb;
  return hipGetLastError();
}

hipError_t diag_ke_diss_pr(hipStream_t s, const DiagGrid& g, const double* U, const double* R,
                           double* part, int nblk) {
  hipLaunchKernelGGL(k_ke_diss_pr, dim3(nblk), dim3(256), 0, s, g.dims, g.N, U, R, part);
  return hipGetLastError();
}

}  // namespace tecno
