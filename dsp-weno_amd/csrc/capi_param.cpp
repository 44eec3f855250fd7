// capi_param.cpp — tecno_read_param: the param.in reader of the command line
// (paramin.cpp) behind the C ABI, host code only.
#include <algorithm>
#include <cstring>
#include <exception>
#include <string>

#include "../../include/tecno.h"
#include "paramin.hpp"

namespace tecno {
int set_last_error(int code, const std::string& msg);
}

extern "C" int tecno_read_param(const char* path, int level, tecno_cfg* cfg, int* n_levels,
                                double* final_time, int* time_stamps, int* max_iter,
                                char* net_file, int net_len) {
  if (!path || !cfg) return tecno::set_last_error(TECNO_EINVAL, "null argument");
  try {
    const tecno::cli::Param p = tecno::cli::read_param(path);
    if (level < 0 || level >= (int)p.Nx.size())
      return tecno::set_last_error(TECNO_EINVAL, "mesh level out of range");
    tecno_cfg c = p.base;
    c.dims = p.scalar ? p.pdim : std::max(p.pdim, 2);
    c.n[0] = p.Nx[level];
    c.n[1] = p.Ny[level];
    c.n[2] = p.Nz[level];
    c.lo[0] = p.xmin;
    c.hi[0] = p.xmax;
    c.lo[1] = p.ymin;
    c.hi[1] = p.ymax;
    c.lo[2] = p.zmin;
    c.hi[2] = p.zmax;
    if (net_file && net_len > 0) {
      if ((int)p.net_file.size() >= net_len)
        return tecno::set_last_error(TECNO_EINVAL, "net_file buffer too small");
      std::memcpy(net_file, p.net_file.c_str(), p.net_file.size() + 1);
    }
    *cfg = c;
    // mesh levels FV::run runs: all with OOC_study yes, else the first
    // (parameter.cc:636-639); `level` may still address any listed mesh
    if (n_levels) *n_levels = p.mesh_levels;
    if (final_time) *final_time = p.final_time;
    if (time_stamps) *time_stamps = p.t_stamps;
    if (max_iter) *max_iter = p.max_iter;
  } catch (const std::exception& e) {
    return tecno::set_last_error(TECNO_EINVAL, e.what());
  }
  return 0;
}
