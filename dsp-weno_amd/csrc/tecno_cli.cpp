// tecno_cli.cpp — `tecno -i param.in`: the reference's command-line and file
// contract (enflo / enfloS main.cc, FV::run fv.cc:1231-1324) on libtecno.
//
// Reads param.in, runs each mesh level (all levels with OOC_study yes, the
// first otherwise) on the GPU and writes, under
// <output_path>/SOLN_FILES_<reconstruct>/:
//   sol####.vtk     the solution at t = 0 and at every output stamp
//                   (output.cc:10-90 layout; systems: density, velocity,
//                   pressure, physical_entropy; scalar: solution)
//   omg####.vtk     systems: the nodal vorticity with every sol file
//                   (output.cc:131-160, find_vort.cc:9-29)
//   1Dsol####.dat   along the first row: x and u (scalar, enflo_scalar
//                   output.cc:64-66), x, rho, u, p (systems, output.cc:164-169)
//   globals.dat     at the start of every step, on every mesh level: time,
//                   total entropy, its relative change and, for systems,
//                   kinetic energy, KE_diss_pr, KE_diss_stress, enstrophy
//                   (compute_global.cc:10-89, evaluated as the reference does
//                   at the first stage; the stream's precision as the
//                   reference sets it)
//   l1_err.dat, l2_err.dat, linf_err.dat   OOC tables (compute_ooc.cc)
// A 2D or 1D problem is solved on its active axes; files repeat the
// collapsed layers the reference stores (Nz = 2 for "2D"), and the error
// norms count them, as the reference's sums do.
// Options after the file: --fp32 (FP32 path), --fp32-state64 (scalar: FP32
// reconstruction on an FP64 state), --mlp fp64|fp32|bf16|fp16, --fix-zmom,
// --fix-edges, --fix-visc, --sign-fallback, --quiet.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tecno.h"
#include "paramin.hpp"

using tecno::cli::Param;

namespace {

struct Run {
  const Param* p;
  tecno_ctx* c = nullptr;
  int dims = 1, nprim = 1;
  int Nx = 1, Ny = 1, Nz = 1;   // reference cell counts
  int ax = 1, ay = 1, az = 1;   // active cell counts
  double dx = 1, dy = 1, dz = 1;
  double layers = 1;            // reference cells per active cell
  std::vector<double> prim;     // active SoA
};

void check(int rc, const char* what) {
  if (rc != 0) {
    std::ostringstream s;
    s << what << ": " << tecno_last_error();
    throw std::runtime_error(s.str());
  }
}

double entropy_of(double rho, double p, double gamma) { return std::log(p) - gamma * std::log(rho); }

std::string counter_name(int counter) {
  std::ostringstream s;
  s << std::setw(4) << std::setfill('0') << counter;
  return s.str();
}

// output.cc: ASCII legacy VTK, cell data in (k, j, i) order
void write_vtk(const Run& r, const std::string& dir, int counter, double time) {
  const Param& p = *r.p;
  const std::string name = dir + "/sol" + counter_name(counter) + ".vtk";
  std::ofstream vtk(name);
  vtk << "# vtk DataFile Version 3.0\nenflo\nASCII\nDATASET STRUCTURED_POINTS\n";
  vtk << "FIELD FieldData 1\nTIME 1 1 double\n" << time << "\n";
  vtk << "DIMENSIONS " << (r.Nx + 1) << " " << (r.Ny + 1) << " " << (r.Nz + 1) << "\n";
  vtk << "SPACING " << r.dx << " " << r.dy << " " << r.dz << " \n";
  vtk << "ORIGIN " << p.xmin << " " << p.ymin << " " << p.zmin << " \n";
  vtk << "CELL_DATA  " << (int64_t)r.Nx * r.Ny * r.Nz << "\n";
  const int64_t N = (int64_t)r.ax * r.ay * r.az;
  auto at = [&](int v, int i, int j, int k) {
    const int jj = r.dims >= 2 ? j : 0, kk = r.dims >= 3 ? k : 0;
    return r.prim[(size_t)v * N + ((int64_t)kk * r.ay + jj) * r.ax + i];
  };
  auto loop = [&](auto f) {
    for (int k = 0; k < r.Nz; ++k)
      for (int j = 0; j < r.Ny; ++j)
        for (int i = 0; i < r.Nx; ++i) f(i, j, k);
  };
  if (p.scalar) {
    vtk << "SCALARS solution double 1\nLOOKUP_TABLE default\n";
    loop([&](int i, int j, int k) { vtk << at(0, i, j, k) << "\n"; });
    return;
  }
  const int iw = r.nprim == 5 ? 3 : -1, ip = r.nprim - 1;
  vtk << "SCALARS density float 1\nLOOKUP_TABLE default\n";
  loop([&](int i, int j, int k) { vtk << at(0, i, j, k) << "\n"; });
  vtk << "VECTORS velocity float\n";
  loop([&](int i, int j, int k) {
    vtk << at(1, i, j, k) << " " << at(2, i, j, k) << " " << (iw > 0 ? at(iw, i, j, k) : 0.0)
        << "\n";
  });
  vtk << "SCALARS pressure float 1\nLOOKUP_TABLE default\n";
  loop([&](int i, int j, int k) { vtk << at(ip, i, j, k) << "\n"; });
  vtk << "SCALARS physical_entropy float 1\nLOOKUP_TABLE default\n";
  loop([&](int i, int j, int k) {
    vtk << entropy_of(at(0, i, j, k), at(ip, i, j, k), p.base.gamma) << "\n";
  });
}

void fetch(Run& r) {
  check(tecno_get_state(r.c, r.prim.data(), tecno_num_cells(r.c)), "tecno_get_state");
}

void write_1d(const Run& r, const std::string& dir, int counter) {
  std::ofstream f(dir + "/1Dsol" + counter_name(counter) + ".dat");
  const int64_t N = (int64_t)r.ax * r.ay * r.az;
  for (int i = 0; i < r.Nx; ++i) {
    f << (r.p->xmin + r.dx / 2) + r.dx * i << "\t";
    if (r.p->scalar)
      f << r.prim[(size_t)i] << "\t";
    else
      f << r.prim[(size_t)i] << "\t" << r.prim[(size_t)N + i] << "\t"
        << r.prim[(size_t)(r.nprim - 1) * N + i];
    f << "\n";
  }
}

// output.cc:131-160: the nodal vorticity, (k, j, i) order
void write_omg(const Run& r, const std::string& dir, int counter, double time) {
  const Param& p = *r.p;
  const int64_t nn = (int64_t)(r.Nx + 1) * (r.Ny + 1) * (r.Nz + 1);
  std::vector<double> w((size_t)3 * nn);
  check(tecno_vorticity(r.c, w.data(), nn), "tecno_vorticity");
  std::ofstream vtk(dir + "/omg" + counter_name(counter) + ".vtk");
  vtk << "# vtk DataFile Version 3.0\nenflo\nASCII\nDATASET STRUCTURED_POINTS\n";
  vtk << "FIELD FieldData 1\nTIME 1 1 double\n" << time << "\n";
  vtk << "DIMENSIONS " << (r.Nx + 1) << " " << (r.Ny + 1) << " " << (r.Nz + 1) << "\n";
  vtk << "SPACING " << r.dx << " " << r.dy << " " << r.dz << " \n";
  vtk << "ORIGIN " << p.xmin << " " << p.ymin << " " << p.zmin << " \n";
  vtk << "POINT_DATA  " << nn << "\nVECTORS vorticity float\n";
  for (int64_t q = 0; q < nn; ++q) vtk << w[q] << " " << w[nn + q] << " " << w[2 * nn + q] << "\n";
}

// one solution output (FV::output, output.cc:10-172)
void write_outputs(Run& r, const std::string& dir, int counter, double time) {
  fetch(r);
  write_vtk(r, dir, counter, time);
  if (!r.p->scalar) write_omg(r, dir, counter, time);
  write_1d(r, dir, counter);
}

// one globals.dat line (compute_global.cc:80-88; enflo_scalar
// compute_global.cc:34-37).  The reference writes the time at the stream's
// current precision and then sets 10 digits, which the stream keeps: the
// first line's time has 6 significant digits, every later value 10.
void write_globals(std::ofstream& glob, const Param& p, tecno_ctx* c, double& ent0) {
  tecno_globals g;
  check(tecno_compute_globals(c, &g), "globals");
  if (g.time == 0) {
    ent0 = g.total_entropy;
    if (ent0 == 0) ent0 = 1.0;  // compute_global.cc:63-67
  }
  glob << g.time << " " << std::setprecision(10) << g.total_entropy << " "
       << (g.total_entropy - ent0) / (p.scalar ? ent0 : std::fabs(ent0)) << " ";
  if (!p.scalar)
    glob << g.kinetic_energy << " " << g.ke_diss_pr << " " << g.ke_diss_stress << " "
         << g.enstrophy << " ";
  glob << std::endl;
}


struct Err {
  double l1[5] = {0, 0, 0, 0, 0}, l2[5] = {0, 0, 0, 0, 0}, linf[5] = {0, 0, 0, 0, 0};
};

// compute_errors.cc:9-34: L1 / L2 (times dx dy dz, over every reference
// cell) and Linf of the primitive variables against the exact solution at the
// cell centres
Err errors(Run& r, double t) {
  const Param& p = *r.p;
  fetch(r);
  Err e;
  const int64_t N = (int64_t)r.ax * r.ay * r.az;
  const int nv = p.scalar ? 1 : 5;
  for (int k = 0; k < r.az; ++k)
    for (int j = 0; j < r.ay; ++j)
      for (int i = 0; i < r.ax; ++i) {
        const double x = p.xmin + 0.5 * r.dx + i * r.dx, y = p.ymin + 0.5 * r.dy + j * r.dy,
                     z = p.zmin + 0.5 * r.dz + k * r.dz;
        const int64_t q = ((int64_t)k * r.ay + j) * r.ax + i;
        for (int v = 0; v < nv; ++v) {
          double got;
          if (p.scalar) got = r.prim[q];
          else if (v == 3 && r.nprim == 4) got = 0.0;  // w of a 2D context
          else {
            const int pv = (r.nprim == 4 && v == 4) ? 3 : v;
            got = r.prim[(size_t)pv * N + q];
          }
          const double d = std::fabs(p.exact[v](x, y, z, t) - got);
          e.l1[v] += d * r.layers;
          e.l2[v] += d * d * r.layers;
          e.linf[v] = std::max(e.linf[v], d);
        }
      }
  const double dv = r.dx * r.dy * r.dz;
  for (int v = 0; v < nv; ++v) {
    e.l1[v] *= dv;
    e.l2[v] = std::sqrt(e.l2[v] * dv);
  }
  return e;
}

// compute_ooc.cc: one table per norm
void write_ooc(const Param& p, const std::vector<Err>& E, int which, const std::string& path) {
  std::ofstream f(path);
  auto val = [&](const Err& e, int v) {
    return which == 0 ? e.l1[v] : which == 1 ? e.l2[v] : e.linf[v];
  };
  const int nv = p.scalar ? 1 : 5;
  const int width = p.scalar ? 40 : 112;
  f << std::left << std::setw(width) << std::setfill('-') << "" << "\n";
  f << std::left << std::setw(6) << std::setfill(' ') << "Nx" << std::setw(6) << "Ny" << std::setw(6)
    << "Nz";
  const char* names[5] = {"DENSITY_ERR", "VELX_ERR", "VELY_ERR", "VELZ_ERR", "PRESSURE_ERR"};
  for (int v = 0; v < nv; ++v)
    f << std::setw(17) << (p.scalar ? "ERR" : names[v]) << std::setw(8) << "OOC";
  f << "\n" << std::left << std::setw(width) << std::setfill('-') << "" << "\n" << std::setfill(' ');
  for (size_t i = 0; i < E.size(); ++i) {
    f << std::left << std::setw(6) << p.Nx[i] << std::setw(6) << p.Ny[i] << std::setw(6) << p.Nz[i];
    double factor = 1.0;
    if (i > 0) {
      auto h2 = [&](size_t l) {
        double s = std::pow((p.xmax - p.xmin) / (double)p.Nx[l], 2.0);
        if (p.pdim >= 2) s += std::pow((p.ymax - p.ymin) / (double)p.Ny[l], 2.0);
        if (p.pdim >= 3) s += std::pow((p.zmax - p.zmin) / (double)p.Nz[l], 2.0);
        return s;
      };
      factor = std::log(std::sqrt(h2(i)) / std::sqrt(h2(i - 1)));
    }
    for (int v = 0; v < nv; ++v) {
      f << std::scientific << std::left << std::setw(17) << std::setprecision(5) << val(E[i], v);
      if (i == 0) {
        f << std::left << std::setw(8) << " - ";
      } else {
        const double ratio = val(E[i], v) < 1.0e-14 ? 0.0 : std::log(val(E[i], v) / val(E[i - 1], v));
        f << std::fixed << std::left << std::setw(8) << std::setprecision(4) << ratio / factor;
      }
    }
    f << "\n";
  }
}

}  // namespace

int main(int argc, char** argv) {
  std::string file;
  int precision = TECNO_FP64, mlp = -1, fix_zmom = 0, sign_fallback = 0, fix_edges = 0,
      fix_visc = 0;
  bool quiet = false, dry_run = false;  // dry run: parse, evaluate the IC, touch no GPU
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    if (s == "-i" && a + 1 < argc) file = argv[++a];
    else if (s == "--fp32") precision = TECNO_FP32;
    else if (s == "--fp32-state64") precision = TECNO_FP32_STATE64;
    else if (s == "--fix-visc") fix_visc = 1;
    else if (s == "--fix-zmom") fix_zmom = 1;
    else if (s == "--fix-edges") fix_edges = 1;
    else if (s == "--sign-fallback") sign_fallback = 1;
    else if (s == "--quiet") quiet = true;
    else if (s == "--dry-run") dry_run = true;
    else if (s == "--mlp" && a + 1 < argc) {
      const std::string m = argv[++a];
      mlp = m == "fp64" ? TECNO_MLP_FP64 : m == "fp32" ? TECNO_MLP_FP32
          : m == "bf16" ? TECNO_MLP_BF16 : m == "fp16" ? TECNO_MLP_FP16 : -2;
      if (mlp == -2) {
        std::cerr << "unknown --mlp " << m << "\n";
        return 2;
      }
    } else if (file.empty() && s[0] != '-') file = s;
    else {
      std::cerr << "usage: tecno -i param.in [--fp32 | --fp32-state64] [--mlp fp64|fp32|bf16|fp16] "
                   "[--fix-zmom] [--fix-edges] [--fix-visc] [--sign-fallback] [--quiet]\n";
      return 2;
    }
  }
  if (file.empty()) {
    std::cerr << "usage: tecno -i param.in\n";
    return 2;
  }
  try {
    const Param p = tecno::cli::read_param(file);
    if (dry_run) {
      // the parsed contract and a checksum of the initial condition at the
      // cell centres of the first level (x fastest, then y, then z)
      const int nx = p.Nx[0], ny = p.Ny[0], nz = p.Nz[0];
      const double dx = (p.xmax - p.xmin) / nx, dy = (p.ymax - p.ymin) / ny,
                   dz = (p.zmax - p.zmin) / nz;
      std::cout << std::setprecision(17);
      std::cout << "system " << (p.scalar ? "scalar" : "euler") << "\nflux " << p.base.flux
                << "\nmodel " << p.base.model << "\nrecon " << p.base.recon << " "
                << p.base.spweno_corr << " " << p.base.eno_level << "\ntime_scheme "
                << p.base.time_scheme << "\nreconstruction "
                << p.reconstruction << "\nlevels " << p.mesh_levels << "\nlisted " << p.Nx.size()
                << "\nfinal_time "
                << p.final_time << "\nt_stamps " << p.t_stamps << "\nooc " << p.ooc
                << "\npdim " << p.pdim << "\ncfl " << p.base.cfl << "\ngamma " << p.base.gamma
                << "\nmu " << p.base.mu_model << " " << p.base.mu_ref << "\nbc";
      for (int i = 0; i < 6; ++i) std::cout << " " << p.base.bc[i];
      std::cout << "\nexact " << p.exact_available << "\n";
      for (size_t v = 0; v < p.ic.size(); ++v) {
        double s1 = 0.0, s2 = 0.0;
        for (int k = 0; k < nz; ++k)
          for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
              const double u = p.ic[v](p.xmin + 0.5 * dx + i * dx, p.ymin + 0.5 * dy + j * dy,
                                       p.zmin + 0.5 * dz + k * dz, 0.0);
              s1 += u;
              s2 += u * u;
            }
        std::cout << "ic " << v << " " << s1 << " " << s2 << "\n";
      }
      return 0;
    }
    const std::string dir = p.output_path + "/SOLN_FILES_" + p.reconstruction;
    std::filesystem::remove_all(dir);
    std::filesystem::create_directories(dir);
    std::ofstream glob(dir + "/globals.dat");
    double ent0 = 1.0;  // the reference's total_ent0, set at time 0 of each level
    std::vector<Err> errs;
    for (int rl = 0; rl < p.mesh_levels; ++rl) {
      Run r;
      r.p = &p;
      r.Nx = p.Nx[rl];
      r.Ny = p.Ny[rl];
      r.Nz = p.Nz[rl];
      r.dx = (p.xmax - p.xmin) / r.Nx;
      r.dy = (p.ymax - p.ymin) / r.Ny;
      r.dz = (p.zmax - p.zmin) / r.Nz;
      tecno_cfg cfg = p.base;
      r.dims = p.scalar ? p.pdim : std::max(p.pdim, 2);
      if (p.scalar && r.dims == 3) throw std::runtime_error("scalar: 3D problems are not supported (1D and 2D)");
      cfg.dims = r.dims;
      cfg.n[0] = r.Nx;
      cfg.n[1] = r.Ny;
      cfg.n[2] = r.Nz;
      cfg.lo[0] = p.xmin; cfg.hi[0] = p.xmax;
      cfg.lo[1] = p.ymin; cfg.hi[1] = p.ymax;
      cfg.lo[2] = p.zmin; cfg.hi[2] = p.zmax;
      cfg.precision = precision;
      cfg.fix_zmom = fix_zmom;
      cfg.sign_fallback = sign_fallback;
      cfg.fix_edges = fix_edges;
      cfg.fix_visc = fix_visc;
      cfg.mlp_dtype = mlp >= 0 ? mlp : (precision == TECNO_FP32 ? TECNO_MLP_FP32 : TECNO_MLP_FP64);
      r.ax = r.Nx;
      r.ay = r.dims >= 2 ? r.Ny : 1;
      r.az = r.dims >= 3 ? r.Nz : 1;
      r.layers = (double)r.Nx * r.Ny * r.Nz / ((double)r.ax * r.ay * r.az);
      if (!quiet)
        std::cout << "\n\n   Solving on mesh with size " << r.Nx << "x" << r.Ny << "x" << r.Nz
                  << std::endl;
      check(tecno_create(&cfg, &r.c), "tecno_create");
      if (cfg.recon == TECNO_REC_SP_WENO_DL) check(tecno_load_net(r.c, p.net_file.c_str()), "net_file");
      r.nprim = tecno_num_vars(r.c);
      const int64_t N = tecno_num_cells(r.c);
      r.prim.assign((size_t)r.nprim * N, 0.0);
      // initial condition at the cell centres (fv.cc:18-110)
      for (int k = 0; k < r.az; ++k)
        for (int j = 0; j < r.ay; ++j)
          for (int i = 0; i < r.ax; ++i) {
            const double x = p.xmin + 0.5 * r.dx + i * r.dx, y = p.ymin + 0.5 * r.dy + j * r.dy,
                         z = p.zmin + 0.5 * r.dz + k * r.dz;
            const int64_t q = ((int64_t)k * r.ay + j) * r.ax + i;
            if (p.scalar) {
              r.prim[q] = p.ic[0](x, y, z, 0.0);
            } else {
              int pv = 0;
              for (int v = 0; v < 5; ++v) {
                if (v == 3 && r.nprim == 4) continue;  // w of a 2D context
                r.prim[(size_t)pv * N + q] = p.ic[v](x, y, z, 0.0);
                ++pv;
              }
            }
          }
      check(tecno_set_state(r.c, r.prim.data(), N), "tecno_set_state");
      const auto t0 = std::chrono::steady_clock::now();
      int iters = 0;
      double time = 0.0;
      // FV::run (fv.cc:1231-1324): the solution files at t = 0 and at every
      // stamp unless OOC_study is on (print_soln, fv.h:50-63); globals on
      // every mesh level, at the start of every step
      // (systems); the scalar solver writes them at t = 0 and after every
      // step, the final state included (enflo_scalar fv.cc:522-607)
      int counter = 0;
      if (!p.ooc) write_outputs(r, dir, counter++, 0.0);
      if (p.scalar) write_globals(glob, p, r.c, ent0);
      const double t_int = p.final_time / p.t_stamps;
      double t_save = t_int;
      while (time < p.final_time && iters < p.max_iter) {
        if (!p.scalar) write_globals(glob, p, r.c, ent0);
        check(tecno_run_steps(r.c, 1, p.final_time, p.t_stamps), "step");
        check(tecno_get_time(r.c, &time, &iters), "time");
        if (p.scalar) write_globals(glob, p, r.c, ent0);
        if (time == t_save) {
          if (!p.ooc) write_outputs(r, dir, counter++, time);
          t_save += t_int;
        }
      }
      const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (!quiet)
        std::cout << "   " << iters << " steps to t = " << time << " in " << sec << " s ("
                  << (double)N * iters / sec << " cell-updates/s)" << std::endl;
      tecno_audit a;
      check(tecno_audit_state(r.c, &a), "audit");
      if (a.n_nonfinite) throw std::runtime_error("non-finite state");
      if (p.ooc) errs.push_back(errors(r, time));
      tecno_destroy(r.c);
    }
    if (p.ooc) {
      write_ooc(p, errs, 0, dir + "/l1_err.dat");
      write_ooc(p, errs, 1, dir + "/l2_err.dat");
      write_ooc(p, errs, 2, dir + "/linf_err.dat");
    }
    if (!quiet) std::cout << "\n\n---------------Solver has finished!!----------------------\n";
  } catch (const std::exception& e) {
    std::cerr << "tecno: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
