// diag.hpp — FP64 diagnostics of the systems solver on the device (the
// quantities FV::compute_global writes to globals.dat and FV::output writes
// to omg####.vtk), launched from capi.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include "fv_device.hpp"

namespace tecno {

struct DiagGrid {
  int dims;      // 2 or 3
  int n[3];      // cells held per direction (1 on a collapsed axis)
  int bc[6];
  int64_t N;     // cells held
  double h[3];   // cell sizes as param.in gives them
  int fix_edges; // Navier-Stokes edge nodes once, each corner its own node
};

// Nodal vorticity and the enstrophy sums (find_vorticity, find_vort.cc:9-29
// with nodal_avg_and_vel_derivatives, nodal_avg_and_der.cc:16-82;
// compute_global.cc:32-65) and, for Navier-Stokes, the viscous dissipation
// rate KE_diss_stress (fv.cc:792-1156).  Nodes of a 3D grid are
// (n0+1)(n1+1)(n2+1); a 2D grid has (n0+1)(n1+1) nodes in three z layers:
// the first node plane of the reference's Nz = 2 storage (its lower cells
// are z ghosts), an interior plane, and the last plane (upper cells are z
// ghosts).  omega (optional): 3 x nodes, x fastest, in that layout.  part:
// per block {enstrophy of z layer 0, 1, 2 (3D: all in 0), KE_diss_stress}.
// Returns the number of blocks through *nblk.
hipError_t diag_nodes(hipStream_t s, const DiagGrid& g, const double* U, double gamma,
                      const Phys<double>& P, int ns, double* omega, double* part, int* nblk);

// KE_diss_pr (fv.cc:756-764): per block -sum over cells of
// (|u|^2/2 R_mass + u . R_momentum) with u the velocity of U and R the
// inviscid residual of U (same layout as U).
hipError_t diag_ke_diss_pr(hipStream_t s, const DiagGrid& g, const double* U, const double* R,
                           double* part, int nblk);

}  // namespace tecno
