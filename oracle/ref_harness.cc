// ref_harness.cc — drives the reference solver's own member functions to
// produce full-precision golden vectors.  Compiled (by oracle/Makefile) against
// the reference sources where they lie under /root/reference, with outputs in
// oracle/_ref/ only.  Nothing of the reference is copied: this file is the
// project's own driver.  Built twice: for enflo/src (systems, NVAR=5) and with
// -DHARNESS_SCALAR for enflo_scalar/src_scalar (NVAR=1).
//
// Modes (all data little-endian raw doubles):
//   run   <param.in> <nsteps> <out.bin>   FV::run's time loop (fv.cc:1262-1310),
//          dumping the interior primitive state before the first and after
//          every step: header int32[5] = {nx, ny, nz, nvar, nsteps}, then
//          initial state, then per step {time, dt, state}.
//   recon <param.in> <n> <ng> <in.bin> <out.bin>  FV::reconstruct on n
//          stencils of 2*ng values -> n x {l, r}.
//   flux  <param.in> <n> <axis> <in.bin> <out.bin>  FV::num_flux on n faces.
//          systems: in = n x 2ng x {P[5], V[5]}, out = n x 5;
//          scalar:  in = n x 2ng, out = n x 1.
//   ent   <param.in> <n> <in.bin> <out.bin>  prim_to_ent on n primitive
//          states (systems only; via the entropy_var array).
//   time  <param.in> <nsteps>   runs nsteps steps of the time loop and prints
//          the wall time of the steps alone (initialisation excluded).
//   globals <param.in> <nsteps> <out.bin>   FV::run's time loop with its
//          diagnostics: find_vorticity before the loop and after every step,
//          compute_global at the first stage of every step (fv.cc:1244-1306).
//          header int32[4] = {nx, ny, nz, nsteps}; per step {time, total_ent,
//          KE, KE_diss_pr, KE_diss_stress, Enstrophy} (scalar: {time,
//          total_ent}); systems then the nodal vorticity before the first and
//          after the last step, 3 x (nx+1)(ny+1)(nz+1) each, x fastest.
//          KE_diss_stress is set to 0 before the loop: the Euler model never
//          assigns it (fv.cc:591-601) and prints whatever the member holds.
//   full  <param.in>   FV::run itself (the reference's own driver, main.cc:23-24)
//          with the same KE_diss_stress initialisation; writes the
//          reference's output files under the param.in's output_path.

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>
#include <algorithm>

#define private public
#include "fv.h"
#undef private

/* the expression parser's constant table is defined by the reference main.cc,
 * which this driver replaces */
std::map<std::string, double> constants;

static std::vector<double> read_bin(const char* path, size_t n) {
  std::vector<double> v(n);
  FILE* f = fopen(path, "rb");
  if (!f || fread(v.data(), sizeof(double), n, f) != n) {
    fprintf(stderr, "harness: cannot read %zu doubles from %s\n", n, path);
    exit(2);
  }
  fclose(f);
  return v;
}

static void write_bin(FILE* f, const double* p, size_t n) {
  if (fwrite(p, sizeof(double), n, f) != n) {
    fprintf(stderr, "harness: write failed\n");
    exit(2);
  }
}

static void setup(FV& fv) {
  fv.Nx_cell = fv.param.Nx_cell[0];
  fv.Ny_cell = fv.param.Ny_cell[0];
  fv.Nz_cell = fv.param.Nz_cell[0];
  fv.initialize();
}

#ifdef HARNESS_SCALAR
static void dump_state(FV& fv, FILE* f) {
  int g = fv.ng_cell;
  std::vector<double> buf;
  for (int k = 0; k < fv.Nz_cell; ++k)
    for (int j = 0; j < fv.Ny_cell; ++j)
      for (int i = 0; i < fv.Nx_cell; ++i) buf.push_back(fv.solution[i + g][j + g][k + g]);
  write_bin(f, buf.data(), buf.size());
}
#else
static void dump_state(FV& fv, FILE* f) {
  int g = fv.ng_cell;
  int64_t N = (int64_t)fv.Nx_cell * fv.Ny_cell * fv.Nz_cell;
  std::vector<double> buf(5 * N);
  int64_t q = 0;
  for (int k = 0; k < fv.Nz_cell; ++k)
    for (int j = 0; j < fv.Ny_cell; ++j)
      for (int i = 0; i < fv.Nx_cell; ++i, ++q) {
        const PrimVar& P = fv.primitive[i + g][j + g][k + g];
        buf[q] = P.density;
        buf[N + q] = P.velocity.x;
        buf[2 * N + q] = P.velocity.y;
        buf[3 * N + q] = P.velocity.z;
        buf[4 * N + q] = P.pressure;
      }
  write_bin(f, buf.data(), buf.size());
}
#endif

// classical RK4 clears its accumulated residual at the start of every step
// (fv.cc:1267-1274; enflo_scalar fv.cc:558-566)
static void clear_residual2(FV& fv) {
  if (fv.param.time_scheme != rk4) return;
  for (int i = 0; i < fv.Nx_cell; ++i)
    for (int j = 0; j < fv.Ny_cell; ++j)
      for (int k = 0; k < fv.Nz_cell; ++k)
#ifdef HARNESS_SCALAR
        fv.residual2[i][j][k] = 0.0;
#else
        fv.residual2[i][j][k].zero();
#endif
}

static int mode_run(char* param, int nsteps, const char* out) {
  FV fv(param);
  setup(fv);
#ifndef HARNESS_SCALAR
  fv.prim_to_con();
  fv.prim_to_ent();
#endif
  fv.time = 0.0;
  fv.iter = 0;
  FILE* f = fopen(out, "wb");
  int32_t hdr[5] = {fv.Nx_cell, fv.Ny_cell, fv.Nz_cell, NVAR, 0};
  fwrite(hdr, sizeof hdr, 1, f);
  dump_state(fv, f);
  double t_int = fv.param.final_time / fv.param.t_stamps;
  double t_save = t_int;
  int done = 0;
  while (done < nsteps && fv.time < fv.param.final_time && fv.iter < fv.param.max_iter) {
#ifdef HARNESS_SCALAR
    fv.solution_old = fv.solution;
#else
    fv.conserved_old = fv.conserved;
#endif
    fv.compute_dt();
    if (fv.time + fv.dt > t_save) fv.dt = t_save - fv.time;
    clear_residual2(fv);
    for (int rk = 0; rk < fv.param.n_rks; ++rk) {
#ifdef HARNESS_SCALAR
      fv.compute_residual(rk);
      fv.update_solution(rk);
      fv.apply_bc();
#else
      fv.prim_to_ent();
      fv.compute_residual(rk);
      fv.update_solution(rk);
      fv.con_to_prim();
#endif
    }
    fv.time += fv.dt;
    ++fv.iter;
    if (fv.time == t_save) t_save += t_int;
    double td[2] = {fv.time, fv.dt};
    write_bin(f, td, 2);
    dump_state(fv, f);
    ++done;
  }
  fseek(f, 0, SEEK_SET);
  hdr[4] = done;
  fwrite(hdr, sizeof hdr, 1, f);
  fclose(f);
  return 0;
}

#ifndef HARNESS_SCALAR
static void dump_vorticity(FV& fv, FILE* f) {
  const int nx = fv.Nx_cell + 1, ny = fv.Ny_cell + 1, nz = fv.Nz_cell + 1;
  std::vector<double> buf((size_t)3 * nx * ny * nz);
  const size_t NN = (size_t)nx * ny * nz;
  size_t q = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++q) {
        const Vector& w = fv.vorticity[i][j][k];
        buf[q] = w.x;
        buf[NN + q] = w.y;
        buf[2 * NN + q] = w.z;
      }
  write_bin(f, buf.data(), buf.size());
}
#endif

static int mode_globals(char* param, int nsteps, const char* out) {
  FV fv(param);
  setup(fv);
#ifndef HARNESS_SCALAR
  fv.prim_to_con();
  fv.prim_to_ent();
  fv.find_vorticity();
  fv.KE_diss_stress = 0.0;
#endif
  fv.time = 0.0;
  fv.iter = 0;
  std::vector<double> rows;
#ifndef HARNESS_SCALAR
  std::vector<double> vort0;
  {
    char* mem = nullptr;
    size_t len = 0;
    FILE* m = open_memstream(&mem, &len);
    dump_vorticity(fv, m);
    fclose(m);
    vort0.assign((double*)mem, (double*)(mem + len));
    free(mem);
  }
#endif
  double t_int = fv.param.final_time / fv.param.t_stamps;
  double t_save = t_int;
  int done = 0;
  while (done < nsteps && fv.time < fv.param.final_time && fv.iter < fv.param.max_iter) {
#ifdef HARNESS_SCALAR
    fv.solution_old = fv.solution;
#else
    fv.conserved_old = fv.conserved;
#endif
    fv.compute_dt();
    if (fv.time + fv.dt > t_save) fv.dt = t_save - fv.time;
    clear_residual2(fv);
    for (int rk = 0; rk < fv.param.n_rks; ++rk) {
#ifdef HARNESS_SCALAR
      fv.compute_residual(rk);
      if (rk == 0) {
        fv.compute_global();
        rows.push_back(fv.time);
        rows.push_back(fv.total_ent);
      }
      fv.update_solution(rk);
      fv.apply_bc();
#else
      fv.prim_to_ent();
      fv.compute_residual(rk);
      if (rk == 0) {
        fv.compute_global();
        const double r[6] = {fv.time, fv.total_ent, fv.KE, fv.KE_diss_pr, fv.KE_diss_stress,
                             fv.Enstrophy};
        rows.insert(rows.end(), r, r + 6);
      }
      fv.update_solution(rk);
      fv.con_to_prim();
#endif
    }
    fv.time += fv.dt;
    ++fv.iter;
#ifndef HARNESS_SCALAR
    fv.find_vorticity();
#endif
    if (fv.time == t_save) t_save += t_int;
    ++done;
  }
  FILE* f = fopen(out, "wb");
  int32_t hdr[4] = {fv.Nx_cell, fv.Ny_cell, fv.Nz_cell, done};
  fwrite(hdr, sizeof hdr, 1, f);
  write_bin(f, rows.data(), rows.size());
#ifndef HARNESS_SCALAR
  write_bin(f, vort0.data(), vort0.size());
  dump_vorticity(fv, f);
#endif
  fclose(f);
  return 0;
}

static int mode_full(char* param) {
  FV fv(param);
#ifndef HARNESS_SCALAR
  fv.KE_diss_stress = 0.0;
#endif
  fv.run();
  return 0;
}

static int mode_time(char* param, int nsteps) {
  FV fv(param);
  setup(fv);
#ifndef HARNESS_SCALAR
  fv.prim_to_con();
  fv.prim_to_ent();
#endif
  fv.time = 0.0;
  fv.iter = 0;
  double t_int = fv.param.final_time / fv.param.t_stamps;
  double t_save = t_int;
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int s = 0; s < nsteps; ++s) {
#ifdef HARNESS_SCALAR
    fv.solution_old = fv.solution;
#else
    fv.conserved_old = fv.conserved;
#endif
    fv.compute_dt();
    if (fv.time + fv.dt > t_save) fv.dt = t_save - fv.time;
    clear_residual2(fv);
    for (int rk = 0; rk < fv.param.n_rks; ++rk) {
#ifdef HARNESS_SCALAR
      fv.compute_residual(rk);
      fv.update_solution(rk);
      fv.apply_bc();
#else
      fv.prim_to_ent();
      fv.compute_residual(rk);
      fv.update_solution(rk);
      fv.con_to_prim();
#endif
    }
    fv.time += fv.dt;
    ++fv.iter;
    if (fv.time == t_save) t_save += t_int;
  }
  clock_gettime(CLOCK_MONOTONIC, &b);
  double sec = (b.tv_sec - a.tv_sec) + 1e-9 * (b.tv_nsec - a.tv_nsec);
  fprintf(stdout, "steps %d seconds %.6f cells %lld\n", nsteps, sec,
          (long long)fv.Nx_cell * fv.Ny_cell * fv.Nz_cell);
  return 0;
}

static int mode_recon(char* param, int n, int ng, const char* in, const char* out) {
  FV fv(param);
  setup(fv);
  std::vector<double> st = read_bin(in, (size_t)n * 2 * ng);
  FILE* f = fopen(out, "wb");
  for (int q = 0; q < n; ++q) {
#ifdef HARNESS_SCALAR
    std::vector<double> s(st.begin() + (size_t)q * 2 * ng, st.begin() + (size_t)(q + 1) * 2 * ng);
    double lr[2];
    fv.reconstruct(fv.param.reconstruct_scheme, fv.dx, s, lr[0], lr[1]);
#else
    std::vector<std::vector<double> > s(2 * ng, std::vector<double>(NVAR, 0.0));
    for (int j = 0; j < 2 * ng; ++j) s[j][0] = st[(size_t)q * 2 * ng + j];
    double l[NVAR], r[NVAR];
    fv.reconstruct(fv.param.reconstruct_scheme, fv.dx, s, l, r);
    double lr[2] = {l[0], r[0]};
#endif
    write_bin(f, lr, 2);
  }
  fclose(f);
  return 0;
}

static int mode_flux(char* param, int n, int axis, const char* in, const char* out) {
  FV fv(param);
  setup(fv);
  int s = 2 * fv.ng_cell;
  Vector nrm;
  nrm = 0;
  if (axis == 0) nrm.x = 1;
  if (axis == 1) nrm.y = 1;
  if (axis == 2) nrm.z = 1;
  FILE* f = fopen(out, "wb");
#ifdef HARNESS_SCALAR
  std::vector<double> data = read_bin(in, (size_t)n * s);
  for (int q = 0; q < n; ++q) {
    std::vector<double> S(data.begin() + (size_t)q * s, data.begin() + (size_t)(q + 1) * s);
    double F;
    fv.num_flux(S, nrm, F);
    write_bin(f, &F, 1);
  }
#else
  std::vector<double> data = read_bin(in, (size_t)n * s * 10);
  std::vector<PrimVar> P(s);
  std::vector<EntVar> V(s);
  for (int q = 0; q < n; ++q) {
    for (int l = 0; l < s; ++l) {
      const double* r = &data[((size_t)q * s + l) * 10];
      P[l].density = r[0];
      P[l].velocity.x = r[1];
      P[l].velocity.y = r[2];
      P[l].velocity.z = r[3];
      P[l].pressure = r[4];
      V[l].entf = r[5];
      V[l].entv.x = r[6];
      V[l].entv.y = r[7];
      V[l].entv.z = r[8];
      V[l].entl = r[9];
    }
    Flux F;
    F.zero();
    fv.num_flux(P, V, nrm, F);
    double o[5] = {F.mass_flux, F.momentum_flux.x, F.momentum_flux.y, F.momentum_flux.z,
                   F.energy_flux};
    write_bin(f, o, 5);
  }
#endif
  fclose(f);
  return 0;
}

#ifndef HARNESS_SCALAR
static int mode_ent(char* param, int n, const char* in, const char* out) {
  FV fv(param);
  setup(fv);
  std::vector<double> data = read_bin(in, (size_t)n * 5);
  int g = fv.ng_cell;
  FILE* f = fopen(out, "wb");
  for (int q = 0; q < n; ++q) {
    PrimVar& P = fv.primitive[g][g][g];
    P.density = data[5 * q];
    P.velocity.x = data[5 * q + 1];
    P.velocity.y = data[5 * q + 2];
    P.velocity.z = data[5 * q + 3];
    P.pressure = data[5 * q + 4];
    fv.prim_to_ent();
    const EntVar& V = fv.entropy_var[g][g][g];
    double o[5] = {V.entf, V.entv.x, V.entv.y, V.entv.z, V.entl};
    write_bin(f, o, 5);
  }
  fclose(f);
  return 0;
}
#endif

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s run|recon|flux|ent <param.in> ...\n", argv[0]);
    return 2;
  }
  std::cout.setstate(std::ios_base::failbit); /* silence the solver's log */
  std::string m = argv[1];
  if (m == "time" && argc == 4) return mode_time(argv[2], atoi(argv[3]));
  if (m == "globals" && argc == 5) return mode_globals(argv[2], atoi(argv[3]), argv[4]);
  if (m == "full" && argc == 3) return mode_full(argv[2]);
  if (m == "run" && argc == 5) return mode_run(argv[2], atoi(argv[3]), argv[4]);
  if (m == "recon" && argc == 7)
    return mode_recon(argv[2], atoi(argv[3]), atoi(argv[4]), argv[5], argv[6]);
  if (m == "flux" && argc == 7)
    return mode_flux(argv[2], atoi(argv[3]), atoi(argv[4]), argv[5], argv[6]);
#ifndef HARNESS_SCALAR
  if (m == "ent" && argc == 6) return mode_ent(argv[2], atoi(argv[3]), argv[4], argv[5]);
#endif
  fprintf(stderr, "harness: bad arguments\n");
  return 2;
}
