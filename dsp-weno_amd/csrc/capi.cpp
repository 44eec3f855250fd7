// capi.cpp — the C ABI of libtecno.so (include/tecno.h): contexts, device
// buffers, time stepping, net_file handling, batched seams, timing.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <random>
#include <vector>

#include "../../include/tecno.h"
#include "fv_launch.hpp"
#include "diag.hpp"
#include "ensemble.hpp"
#include "mlp.hpp"
#include "netinit.hpp"

namespace tecno {
extern template struct Launch<double>;
extern template struct Launch<float>;
extern template struct StateOps<double>;
extern template struct StateOps<float>;
}  // namespace tecno

using namespace tecno;

static thread_local std::string g_err;

static int fail(int code, const std::string& msg);
namespace tecno {
// for the host-only entry points in other files of the library
int set_last_error(int code, const std::string& msg) { return fail(code, msg); }
}  // namespace tecno
static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(TECNO_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

struct HostNet {
  int nl = 0;
  int w[TECNO_NET_MAXL + 1] = {0};
  std::vector<double> W = std::vector<double>(TECNO_NET_MAXL * TECNO_NET_MAXW * TECNO_NET_MAXW, 0.0);
  std::vector<double> b = std::vector<double>(TECNO_NET_MAXL * TECNO_NET_MAXW, 0.0);
};

struct tecno_ctx {
  tecno_cfg cfg;
  int device = 0;  // the GPU current at tecno_create
  int dims = 2;
  int nvar = 4;    // conserved variables held
  int nprim = 4;   // primitive variables exchanged with the host
  int n[3] = {1, 1, 1};
  int64_t N = 0;
  int G = 2;
  double h[3] = {0, 0, 0};
  double hmin = 0, vol = 0;
  int esize = 8;   // bytes per state value
  hipStream_t stream = nullptr;
  void* U[3] = {nullptr, nullptr, nullptr};
  Ctl* ctl = nullptr;
  double* scratch = nullptr;  // host<->device staging, audit partials
  size_t scratch_bytes = 0;
  HostNet net;
  bool has_net = false;
  double* dW64 = nullptr;
  double* db64 = nullptr;
  float* dW32 = nullptr;
  float* db32 = nullptr;
  unsigned short* dWmf = nullptr;  // MFMA-packed weights: [bf16 | fp16][4][64][8]
  Phys<double> P64;
  Phys<float> P32;
  LaunchInfo L{};
  bool state_set = false;
  bool configured_final = false;
  double cfg_final = 0;
  int cfg_stamps = 0;
  // timing
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
  double t_ms_acc = 0;
  int64_t t_n = 0;
  MlpTrainer* trainer = nullptr;
};

// ------------------------------------------------------------------ helpers
static int flux_ng(int flux) {
  switch (flux) {
    case TECNO_FLUX_LIN_EC4: case TECNO_FLUX_LIN_TECNO4: case TECNO_FLUX_BUR_EC4:
    case TECNO_FLUX_BUR_TECNO4: case TECNO_FLUX_KEPES_TECNO4_ROE:
    case TECNO_FLUX_KEPES_TECNO4_RUSANOV: case TECNO_FLUX_ROE_TECNO4_ROE:
    case TECNO_FLUX_KEPEC4: case TECNO_FLUX_ROE_EC4: case TECNO_FLUX_KEP4: case TECNO_FLUX_KEPS4:
      return 2;
    default:
      return 1;
  }
}

template <typename T>
static void fill_phys(const tecno_ctx* c, Phys<T>& P) {
  const tecno_cfg& g = c->cfg;
  P.system = g.system;
  P.model = g.model;
  P.flux = g.flux;
  P.recon = g.recon == TECNO_REC_SP_WENO ? REC_SPWENO
            : g.recon == TECNO_REC_ENO   ? REC_ENO
            : g.recon == TECNO_REC_SP_WENO_DL ? REC_DSP
            : g.recon == TECNO_REC_SECOND ? REC_SECOND
            : g.recon == TECNO_REC_MINMOD ? REC_MINMOD
            : g.recon == TECNO_REC_MINABS ? REC_MINABS
                                               : REC_FIRST;
  P.corr = g.spweno_corr;
  P.eno_k = g.eno_level;
  P.ec1_fix = g.ec1_fix;
  P.fix_zmom = g.fix_zmom;
  P.sign_fallback = g.sign_fallback;
  P.fix_edges = g.fix_edges;
  P.fix_visc = g.fix_visc;
  P.mlp_mode = g.mlp_dtype;
  P.gamma = T(g.gamma);
  P.R = T(g.gas_const);
  P.gm1 = T(g.gamma - 1.0);
  P.inv_gm1 = T(1.0 / (g.gamma - 1.0));
  for (int d = 0; d < 3; ++d) {
    P.vel[d] = T(g.vel[d]);
    P.h[d] = T(c->h[d]);
  }
  P.area[0] = T(c->h[1] * c->h[2]);
  P.area[1] = T(c->h[0] * c->h[2]);
  P.area[2] = T(c->h[0] * c->h[1]);
  P.area_d[0] = c->h[1] * c->h[2];
  P.area_d[1] = c->h[0] * c->h[2];
  P.area_d[2] = c->h[0] * c->h[1];
  P.hmin = T(c->hmin);
  P.cfl = T(g.cfl);
  P.vol = T(c->vol);
  P.vol_d = c->vol;
  P.hmin_d = c->hmin;
  P.cfl_d = g.cfl;
  P.s05 = T(std::sqrt(0.5));
  P.sgm1 = T(std::sqrt(g.gamma - 1.0));
  P.gpgm = T((g.gamma + 1.0) / (g.gamma - 1.0));
  P.inv_gamma = T(1.0 / g.gamma);
  P.gamma_d = g.gamma;
  P.gm1_d = g.gamma - 1.0;
  P.inv_gm1_d = 1.0 / (g.gamma - 1.0);
  P.R_d = g.gas_const;
  P.s05_d = std::sqrt(0.5);
  P.sgm1_d = std::sqrt(g.gamma - 1.0);
  P.inv_gamma_d = 1.0 / g.gamma;
  P.vel_norm = std::sqrt(g.vel[0] * g.vel[0] + g.vel[1] * g.vel[1] + g.vel[2] * g.vel[2]);
  P.mu_model = g.mu_model;
  P.mu_ref = g.mu_ref;
  P.T_ref = g.T_ref;
  P.T_0 = g.T_0;
  P.omega = g.omega;
  P.prandtl = g.prandtl;
  P.Cp = g.gamma * g.gas_const / (g.gamma - 1.0);  // parameter.cc:286
  P.arch5 = 0;
  P.net.nl = c->net.nl;
  for (int i = 0; i <= TECNO_NET_MAXL; ++i) P.net.w[i] = c->net.w[i];
  if (c->has_net && c->net.nl == 4) {
    bool all5 = true;
    for (int i = 0; i <= 4; ++i) all5 = all5 && c->net.w[i] == 5;
    P.arch5 = all5;
  }
}

static void refresh_phys(tecno_ctx* c) {
  fill_phys(c, c->P64);
  fill_phys(c, c->P32);
  c->P64.net.W = c->dW64;
  c->P64.net.b = c->db64;
  c->P32.net.W = c->dW32;
  c->P32.net.b = c->db32;
  c->P32.mfma_w[0] = c->dWmf;
  c->P32.mfma_w[1] = c->dWmf ? c->dWmf + 4 * 64 * 8 : nullptr;
  // the 3D FP32 face runs on the FP64 state (Phys<double>) with the FP32
  // and matrix-core weights
  c->P64.mfma_w[0] = c->P32.mfma_w[0];
  c->P64.mfma_w[1] = c->P32.mfma_w[1];
  for (Phys<double>* P : {&c->P64}) {
    P->net32.nl = c->P32.net.nl;
    for (int i = 0; i <= TECNO_NET_MAXL; ++i) P->net32.w[i] = c->P32.net.w[i];
    P->net32.W = c->dW32;
    P->net32.b = c->db32;
  }
  c->P32.net32 = c->P32.net;
}

static int ensure_scratch(tecno_ctx* c, size_t bytes) {
  if (bytes <= c->scratch_bytes) return 0;
  if (c->scratch) (void)hipFree(c->scratch);
  c->scratch = nullptr;
  c->scratch_bytes = 0;
  HIPCHK(hipMalloc(&c->scratch, bytes));
  c->scratch_bytes = bytes;
  return 0;
}

static int upload_net(tecno_ctx* c) {
  const size_t nw = TECNO_NET_MAXL * TECNO_NET_MAXW * TECNO_NET_MAXW, nb = TECNO_NET_MAXL * TECNO_NET_MAXW;
  if (!c->dW64) {
    HIPCHK(hipMalloc(&c->dW64, nw * sizeof(double)));
    HIPCHK(hipMalloc(&c->db64, nb * sizeof(double)));
    HIPCHK(hipMalloc(&c->dW32, nw * sizeof(float)));
    HIPCHK(hipMalloc(&c->db32, nb * sizeof(float)));
  }
  std::vector<float> w32(c->net.W.begin(), c->net.W.end()), b32(c->net.b.begin(), c->net.b.end());
  HIPCHK(hipMemcpy(c->dW64, c->net.W.data(), nw * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->db64, c->net.b.data(), nb * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->dW32, w32.data(), nw * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->db32, b32.data(), nb * sizeof(float), hipMemcpyHostToDevice));
  // per-lane packed weights for the matrix-core network (5-5-5-5-5 only)
  if (c->net.nl == 4) {
    bool all5 = true;
    for (int i = 0; i <= 4; ++i) all5 = all5 && c->net.w[i] == 5;
    if (all5) {
      std::vector<unsigned short> pk(2 * 4 * 64 * 8, 0);
      for (int l = 0; l < 4; ++l)
        for (int lane = 0; lane < 64; ++lane) {
          const int m = lane % 32, half = lane / 32;
          int j = -1;  // output neuron of row m for this lane's half
          if (half == 0 && (m < 4 || m == 8)) j = m < 4 ? m : 4;
          if (half == 1 && ((m >= 4 && m < 8) || m == 12)) j = m < 8 ? m - 4 : 4;
          for (int e = 0; e < 6; ++e) {  // 5 weights, then the bias (input 1 at slot 5)
            const float w = j < 0   ? 0.0f
                            : e < 5 ? (float)c->net.W[(size_t)l * TECNO_NET_MAXW * TECNO_NET_MAXW +
                                                      j * TECNO_NET_MAXW + e]
                                    : (float)c->net.b[(size_t)l * TECNO_NET_MAXW + j];
            uint32_t u;
            std::memcpy(&u, &w, 4);
            u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
            pk[((size_t)l * 64 + lane) * 8 + e] = (unsigned short)(u >> 16);
            _Float16 h = (_Float16)w;
            unsigned short hb;
            std::memcpy(&hb, &h, 2);
            pk[4 * 64 * 8 + ((size_t)l * 64 + lane) * 8 + e] = hb;
          }
        }
      if (!c->dWmf) HIPCHK(hipMalloc(&c->dWmf, pk.size() * sizeof(unsigned short)));
      HIPCHK(hipMemcpy(c->dWmf, pk.data(), pk.size() * sizeof(unsigned short), hipMemcpyHostToDevice));
    }
  }
  refresh_phys(c);
  return 0;
}

// 2D Euler FP32 with a TeCNO flux other than the fused kepes_tecno4_roe
// stage (or with ec1_fix): the generic face with FP32 jumps, FP64 kernel family
static bool euler2_generic_fp32(const tecno_cfg& g) {
  return g.system == TECNO_SYS_EULER && g.dims == 2 && g.precision == TECNO_FP32 &&
         (g.flux != TECNO_FLUX_KEPES_TECNO4_ROE || g.ec1_fix);
}

// the FP64 kernel family: FP64, and the FP32 paths that keep an FP64 state
// and differ from FP64 in the face only (scalar L.scalar_mixed, 3D L.mix3,
// 2D generic TeCNO L.mix2)
static bool fp64(const tecno_ctx* c) {
  return c->cfg.precision == TECNO_FP64 || c->cfg.precision == TECNO_FP32_STATE64 ||
         (c->cfg.system == TECNO_SYS_EULER && c->cfg.dims == 3) || euler2_generic_fp32(c->cfg);
}

// cells the reference stores per held cell: the param.in cell counts of the
// collapsed axes (a "2D" run holds Nx x Ny x Nz_cell = 2 cells in the
// reference), so sums over held cells times this are the reference's sums
static double collapsed_cells(const tecno_ctx* c) {
  double l = 1.0;
  for (int d = c->dims; d < 3; ++d) l *= c->cfg.n[d];
  return l;
}

static hipError_t launch_stage(tecno_ctx* c, const void* U, const void* U0, void* Uo,
                               const StageCoef& k) {
  if (fp64(c)) return Launch<double>::stage(c->L, c->P64, U, U0, Uo, k);
  return Launch<float>::stage(c->L, c->P32, U, U0, Uo, k);
}

static hipError_t st_prim_to_con(tecno_ctx* c, const double* prim, void* U) {
  return c->L.state_fp64 ? StateOps<double>::prim_to_con(c->L, prim, U, c->cfg.gamma)
                         : StateOps<float>::prim_to_con(c->L, prim, U, c->cfg.gamma);
}
static hipError_t st_con_to_prim(tecno_ctx* c, const void* U, double* prim) {
  return c->L.state_fp64 ? StateOps<double>::con_to_prim(c->L, U, prim, c->cfg.gamma)
                         : StateOps<float>::con_to_prim(c->L, U, prim, c->cfg.gamma);
}

// ------------------------------------------------------------------ API
extern "C" {

const char* tecno_last_error(void) { return g_err.c_str(); }
const char* tecno_version(void) { return "tecno-mi355x 0.1 (gfx950)"; }

void tecno_cfg_default(tecno_cfg* cfg) {
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->system = TECNO_SYS_EULER;
  cfg->model = TECNO_MODEL_EULER;
  cfg->flux = TECNO_FLUX_KEPES_TECNO4_ROE;
  cfg->recon = TECNO_REC_SP_WENO;
  cfg->eno_level = 2;
  cfg->time_scheme = TECNO_TS_SSPRK3;
  cfg->precision = TECNO_FP64;
  cfg->mlp_dtype = TECNO_MLP_FP64;
  cfg->dims = 2;
  for (int d = 0; d < 3; ++d) {
    cfg->n[d] = 2;
    cfg->lo[d] = 0.0;
    cfg->hi[d] = 1.0;
  }
  cfg->cfl = 0.5;
  cfg->vel[0] = 1.0;
  cfg->gamma = 1.4;
  cfg->gas_const = 1.0;
  cfg->prandtl = 0.72;
}

int tecno_create(const tecno_cfg* cfg, tecno_ctx** out) {
  if (!cfg || !out) return fail(TECNO_EINVAL, "null argument");
  *out = nullptr;
  const tecno_cfg& g = *cfg;
  if (g.dims < 1 || g.dims > 3) return fail(TECNO_EINVAL, "dims must be 1, 2 or 3");
  if (g.system == TECNO_SYS_EULER && g.dims == 1)
    return fail(TECNO_EINVAL, "Euler: 2D or 3D grids");
  if (g.system == TECNO_SYS_EULER && g.dims == 3) {
    // the 3D FP32 path reconstructs in FP32 on the FP64 state: TeCNO fluxes
    // of the kepec / roe_ec / kep family with sp_weno (any correction) or
    // sp_weno_dl
    if (g.precision == TECNO_FP32 &&
        (g.flux < 100 || g.flux > 106 ||
         (g.recon != TECNO_REC_SP_WENO && g.recon != TECNO_REC_SP_WENO_DL)))
      return fail(TECNO_EINVAL, "3D FP32 path: TeCNO fluxes (kepes_tecno*, roe_tecno*, "
                                "kep_tecno_roe) with sp_weno or sp_weno_dl only");
    if (g.precision != TECNO_FP64 && g.precision != TECNO_FP32)
      return fail(TECNO_EINVAL, "3D Euler / Navier-Stokes: precision FP64 or FP32");
    if (g.recon == TECNO_REC_ENO && g.eno_level == 3)
      return fail(TECNO_EINVAL, "3D: reconstructions with 2 ghost cells (eno 3 is 2D only)");
  }
  const bool ns = g.model == TECNO_MODEL_NS_ES || g.model == TECNO_MODEL_NS_KEP;
  if (g.system == TECNO_SYS_EULER && ns) {
    if (g.dims != 3) return fail(TECNO_EINVAL, "Navier-Stokes runs on 3D grids");
    // neumann and wall boundaries extend the reference (parameter.cc:564-568
    // allows periodic only): the viscous nodes on them read the same ghost
    // cells as the faces, composed per axis at edges and corners
    for (int d = 0; d < 3; ++d)
      if (g.n[d] < 2) return fail(TECNO_EINVAL, "Navier-Stokes: at least 2 cells per direction");
    if (!(g.prandtl > 0.0)) return fail(TECNO_EINVAL, "prandtl must be positive");
    if (g.mu_model < TECNO_MU_CONSTANT || g.mu_model > TECNO_MU_POWER)
      return fail(TECNO_EINVAL, "unknown viscosity model");
  }
  if (g.system == TECNO_SYS_SCALAR && g.dims == 3)
    return fail(TECNO_EINVAL, "scalar: 3D grids are not supported (1D and 2D)");
  if (g.system == TECNO_SYS_SCALAR && (g.flux < 0 || g.flux > 7))
    return fail(TECNO_EINVAL, "unknown scalar flux");
  if (g.system == TECNO_SYS_EULER && !((g.flux >= 100 && g.flux <= 106) || (g.flux >= 120 && g.flux <= 128)))
    return fail(TECNO_EINVAL, "unknown Euler flux");
  if (g.recon < TECNO_REC_SP_WENO || g.recon > TECNO_REC_MINABS)
    return fail(TECNO_EINVAL, "unknown reconstruction");
  if (g.system == TECNO_SYS_EULER && g.model != TECNO_MODEL_EULER && !ns)
    return fail(TECNO_EINVAL, "unknown Euler model");
  if (g.recon == TECNO_REC_ENO && g.eno_level != 2 && g.eno_level != 3)
    return fail(TECNO_EINVAL, "eno k: k must be 2 or 3");
  if (g.time_scheme < TECNO_TS_RK1 || g.time_scheme > TECNO_TS_RK4)
    return fail(TECNO_EINVAL, "unknown time scheme");
  if (g.precision != TECNO_FP64 && g.precision != TECNO_FP32 && g.precision != TECNO_FP32_STATE64)
    return fail(TECNO_EINVAL, "precision must be FP64, FP32 or FP32_STATE64");
  if (g.precision == TECNO_FP32_STATE64 &&
      (g.system != TECNO_SYS_SCALAR || g.recon != TECNO_REC_SP_WENO))
    return fail(TECNO_EINVAL, "FP32_STATE64: scalar problems with sp_weno, sp_wenoc or corr2 only");
  for (int i = 0; i < 6; ++i)
    if (g.bc[i] == TECNO_BC_DIRICHLET || g.bc[i] < 0 || g.bc[i] > 3)
      return fail(TECNO_EINVAL, "boundary type must be periodic, neumann or wall");
  for (int d = 0; d < 3; ++d)  // parameter.cc:570-581
    if ((g.bc[2 * d] == TECNO_BC_PERIODIC) != (g.bc[2 * d + 1] == TECNO_BC_PERIODIC))
      return fail(TECNO_EINVAL, "periodic boundaries must be paired");
  for (int d = 0; d < 3; ++d) {
    if (g.n[d] < 1) return fail(TECNO_EINVAL, "cell counts must be positive");
    if (!(g.hi[d] > g.lo[d])) return fail(TECNO_EINVAL, "empty domain");
  }
  if (!(g.cfl > 0)) return fail(TECNO_EINVAL, "cfl must be positive");

  tecno_ctx* c = new tecno_ctx();
  c->cfg = g;
  c->dims = g.dims;
  for (int d = 0; d < 3; ++d) {
    c->h[d] = (g.hi[d] - g.lo[d]) / g.n[d];
    c->n[d] = d < g.dims ? g.n[d] : 1;
  }
  c->hmin = std::min(c->h[0], std::min(c->h[1], c->h[2]));
  c->vol = c->h[0] * c->h[1] * c->h[2];
  c->N = (int64_t)c->n[0] * c->n[1] * c->n[2];
  c->nvar = g.system == TECNO_SYS_SCALAR ? 1 : (g.dims == 3 ? 5 : 4);
  c->nprim = c->nvar;
  int rec_ng = g.recon == TECNO_REC_ENO ? g.eno_level : g.recon == TECNO_REC_FIRST ? 1 : 2;
  c->G = std::max(flux_ng(g.flux), rec_ng);
  if (c->G < 2) c->G = 2;
  for (int d = 0; d < g.dims; ++d)
    if (c->n[d] < c->G) {
      delete c;
      return fail(TECNO_EINVAL, "grid too small for the stencil");
    }
  // the mixed-precision Euler path keeps the state in FP64
  c->L.state_fp64 = (fp64(c) || g.system == TECNO_SYS_EULER) ? 1 : 0;
  c->esize = c->L.state_fp64 ? 8 : 4;
  if (g.system == TECNO_SYS_EULER && g.dims == 2 && g.precision == TECNO_FP32 &&
      (g.flux < 100 || g.flux > 106 ||
       (g.recon != TECNO_REC_SP_WENO && g.recon != TECNO_REC_SP_WENO_DL))) {
    delete c;
    return fail(TECNO_EINVAL, "FP32 Euler path: TeCNO fluxes (kepes_tecno*, roe_tecno*, "
                              "kep_tecno_roe) with sp_weno, sp_wenoc, corr2 or sp_weno_dl only "
                              "(use FP64 for the other options)");
  }
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return fail(TECNO_EHIP, "hipGetDevice failed");
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(TECNO_EHIP, "hipStreamCreate failed");
  }
  size_t bytes = (size_t)c->N * c->nvar * c->esize;
  for (int i = 0; i < 3; ++i)
    if (hipMalloc(&c->U[i], bytes) != hipSuccess) {
      tecno_destroy(c);
      return fail(TECNO_ENOMEM, "device allocation of the state failed");
    }
  if (hipMalloc(&c->ctl, sizeof(Ctl)) != hipSuccess) {
    tecno_destroy(c);
    return fail(TECNO_ENOMEM, "device allocation failed");
  }
  // classical RK4 keeps the accumulated residual (FV::residual2, fv.h:221)
  if (g.time_scheme == TECNO_TS_RK4 && hipMalloc(&c->L.R2, bytes) != hipSuccess) {
    tecno_destroy(c);
    return fail(TECNO_ENOMEM, "device allocation of the RK4 accumulator failed");
  }
  Ctl h{};
  (void)hipMemcpy(c->ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice);
  c->L.system = g.system == TECNO_SYS_SCALAR ? 0 : 1;
  c->L.dims = g.dims;
  c->L.G = c->G;
  c->L.fast_flux = g.flux == TECNO_FLUX_KEPES_TECNO4_ROE && !g.ec1_fix ? FV_KEPES4_ROE : FV_GENERIC;
  c->L.scalar_mixed = g.precision == TECNO_FP32_STATE64 ? 1 : 0;
  c->L.mix3 = c->L.mix2 = M3_NONE;
  {
    const int mix = g.recon == TECNO_REC_SP_WENO        ? M3_SPW
                    : g.mlp_dtype == TECNO_MLP_BF16 ? M3_DSP_BF16
                    : g.mlp_dtype == TECNO_MLP_FP16 ? M3_DSP_FP16
                                                    : M3_DSP_SIMT;
    if (g.system == TECNO_SYS_EULER && g.dims == 3 && g.precision == TECNO_FP32) c->L.mix3 = mix;
    if (euler2_generic_fp32(g)) c->L.mix2 = mix;
  }
  c->L.mlp_mode = g.mlp_dtype == TECNO_MLP_BF16 ? MM_BF16
                  : g.mlp_dtype == TECNO_MLP_FP16 ? MM_FP16
                                                   : MM_SIMT;
  c->L.fast_rec = (g.recon == TECNO_REC_SP_WENO) ? RV_SPWENO
                  : (g.recon == TECNO_REC_SP_WENO_DL) ? RV_DSP
                                                      : RV_GENERIC;
  for (int d = 0; d < 3; ++d) c->L.grid.n[d] = c->n[d];
  for (int i = 0; i < 6; ++i) c->L.grid.bc[i] = g.bc[i];
  c->L.grid.N = c->N;
  c->L.stream = c->stream;
  c->L.ctl = c->ctl;
  refresh_phys(c);
  *out = c;
  return 0;
}

void tecno_destroy(tecno_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 3; ++i)
    if (c->U[i]) (void)hipFree(c->U[i]);
  if (c->ctl) (void)hipFree(c->ctl);
  if (c->L.R2) (void)hipFree(c->L.R2);
  if (c->scratch) (void)hipFree(c->scratch);
  if (c->dW64) (void)hipFree(c->dW64);
  if (c->db64) (void)hipFree(c->db64);
  if (c->dW32) (void)hipFree(c->dW32);
  if (c->db32) (void)hipFree(c->db32);
  if (c->dWmf) (void)hipFree(c->dWmf);
  for (auto& e : c->ev) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  delete c->trainer;
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// net_file reader: Neural_Network::load_network (enflo/src/neural_network.cc:82-130)
static int parse_net(const char* path, HostNet& net, std::string& err) {
  std::ifstream in(path);
  if (!in.good()) {
    err = std::string("DSP-WENO network file does not exist: ") + path;
    return TECNO_EIO;
  }
  net = HostNet();
  bool weight = true;
  int nw = 0, nb = 0;
  std::string line;
  while (std::getline(in, line)) {
    while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
    if (line == "BEGIN") {
      std::vector<std::vector<double>> rows;
      std::vector<double> bias;
      while (std::getline(in, line)) {
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ')) line.pop_back();
        if (line == "END") break;
        std::istringstream iss(line);
        std::vector<double> row;
        double v;
        while (iss >> v) row.push_back(v);
        if (weight)
          rows.push_back(row);
        else
          bias.insert(bias.end(), row.begin(), row.end());
      }
      if (weight) {
        if (nw >= TECNO_NET_MAXL || rows.empty()) { err = "net_file: bad weight block"; return TECNO_EIO; }
        int out = (int)rows.size(), inn = (int)rows[0].size();
        if (out > TECNO_NET_MAXW || inn > TECNO_NET_MAXW) { err = "net_file: layer wider than 16"; return TECNO_EIO; }
        for (int o = 0; o < out; ++o) {
          if ((int)rows[o].size() != inn) { err = "net_file: ragged weight block"; return TECNO_EIO; }
          for (int i = 0; i < inn; ++i) net.W[(size_t)nw * TECNO_NET_MAXW * TECNO_NET_MAXW + o * TECNO_NET_MAXW + i] = rows[o][i];
        }
        if (nw > 0 && net.w[nw] != inn) { err = "net_file: layer widths do not chain"; return TECNO_EIO; }
        net.w[nw] = inn;
        net.w[nw + 1] = out;
        nw++;
      } else {
        if (nb >= nw || (int)bias.size() != net.w[nb + 1]) { err = "net_file: bias block does not match its layer"; return TECNO_EIO; }
        for (size_t o = 0; o < bias.size(); ++o) net.b[(size_t)nb * TECNO_NET_MAXW + o] = bias[o];
        nb++;
      }
    } else if (line == "Biases") {
      weight = false;
    }
  }
  if (nw == 0 || nb != nw) { err = "net_file: weight and bias blocks do not match"; return TECNO_EIO; }
  if (net.w[0] != 5 || net.w[nw] != 5) { err = "net_file: the network must map 5 features to 5 vertex weights"; return TECNO_EIO; }
  net.nl = nw;
  return 0;
}

int tecno_load_net(tecno_ctx* c, const char* path) {
  if (!c || !path) return fail(TECNO_EINVAL, "null argument");
  HostNet net;
  std::string err;
  int rc = parse_net(path, net, err);
  if (rc) return fail(rc, err);
  c->net = net;
  c->has_net = true;
  return upload_net(c);
}

int tecno_set_net(tecno_ctx* c, int nl, const int32_t* widths, const double* W, const double* b) {
  if (!c || !widths || !W || !b) return fail(TECNO_EINVAL, "null argument");
  if (nl < 1 || nl > TECNO_NET_MAXL) return fail(TECNO_EINVAL, "bad layer count");
  HostNet net;
  net.nl = nl;
  size_t ow = 0, ob = 0;
  for (int l = 0; l <= nl; ++l) {
    if (widths[l] < 1 || widths[l] > TECNO_NET_MAXW) return fail(TECNO_EINVAL, "bad layer width");
    net.w[l] = widths[l];
  }
  if (net.w[0] != 5 || net.w[nl] != 5) return fail(TECNO_EINVAL, "network must map 5 -> 5");
  for (int l = 0; l < nl; ++l) {
    for (int o = 0; o < net.w[l + 1]; ++o) {
      for (int i = 0; i < net.w[l]; ++i)
        net.W[(size_t)l * TECNO_NET_MAXW * TECNO_NET_MAXW + o * TECNO_NET_MAXW + i] = W[ow + o * net.w[l] + i];
      net.b[(size_t)l * TECNO_NET_MAXW + o] = b[ob + o];
    }
    ow += (size_t)net.w[l] * net.w[l + 1];
    ob += net.w[l + 1];
  }
  c->net = net;
  c->has_net = true;
  return upload_net(c);
}

int tecno_get_net(const tecno_ctx* c, int* nl, int32_t* widths, double* W, double* b) {
  if (!c || !nl || !widths || !W || !b) return fail(TECNO_EINVAL, "null argument");
  if (!c->has_net) return fail(TECNO_ESTATE, "no network loaded");
  *nl = c->net.nl;
  size_t ow = 0, ob = 0;
  for (int l = 0; l <= c->net.nl; ++l) widths[l] = c->net.w[l];
  for (int l = 0; l < c->net.nl; ++l) {
    for (int o = 0; o < c->net.w[l + 1]; ++o) {
      for (int i = 0; i < c->net.w[l]; ++i)
        W[ow + o * c->net.w[l] + i] = c->net.W[(size_t)l * TECNO_NET_MAXW * TECNO_NET_MAXW + o * TECNO_NET_MAXW + i];
      b[ob + o] = c->net.b[(size_t)l * TECNO_NET_MAXW + o];
    }
    ow += (size_t)c->net.w[l] * c->net.w[l + 1];
    ob += c->net.w[l + 1];
  }
  return 0;
}

int64_t tecno_num_cells(const tecno_ctx* c) { return c ? c->N : 0; }
int tecno_num_vars(const tecno_ctx* c) { return c ? c->nprim : 0; }

static int need_net(tecno_ctx* c) {
  if (c->cfg.recon == TECNO_REC_SP_WENO_DL && !c->has_net)
    return fail(TECNO_ESTATE, "sp_weno_dl needs a network (tecno_load_net)");
  // every kernel that evaluates the network dereferences these pointers
  if (c->cfg.recon == TECNO_REC_SP_WENO_DL &&
      (!c->P64.net.W || !c->P64.net.b || !c->P32.net.W || !c->P32.net.b))
    return fail(TECNO_ESTATE, "network weights are not on the device");
  if (c->cfg.recon == TECNO_REC_SP_WENO_DL && c->L.mlp_mode != MM_SIMT &&
      (!c->P32.mfma_w[0] || !c->P32.mfma_w[1]))
    return fail(TECNO_ESTATE, "matrix-core weights are not on the device");
  if (c->cfg.recon == TECNO_REC_SP_WENO_DL && c->L.mlp_mode != MM_SIMT && !c->P32.arch5)
    return fail(TECNO_EINVAL, "the matrix-core network needs the 5-5-5-5-5 architecture");
  return 0;
}

int tecno_set_state(tecno_ctx* c, const double* prim, int64_t n) {
  if (!c || !prim) return fail(TECNO_EINVAL, "null argument");
  if (n != c->N) return fail(TECNO_EINVAL, "cell count mismatch");
  size_t bytes = (size_t)c->N * c->nprim * sizeof(double);
  if (int rc = ensure_scratch(c, bytes)) return rc;
  HIPCHK(hipMemcpyAsync(c->scratch, prim, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(st_prim_to_con(c, c->scratch, c->U[0]));
  Ctl h{};
  h.time = 0.0;
  h.iter = 0;
  h.final_time = INFINITY;
  h.max_iter = 0x7fffffff;
  h.t_save = INFINITY;
  h.t_int = 0.0;
  HIPCHK(hipMemcpyAsync(c->ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
  if (!(c->cfg.system == TECNO_SYS_SCALAR && c->cfg.model == TECNO_MODEL_LINADV)) {
    if (fp64(c))
      HIPCHK(Launch<double>::speed(c->L, c->P64, c->U[0]));
    else
      HIPCHK(Launch<float>::speed(c->L, c->P32, c->U[0]));
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  c->state_set = true;
  c->configured_final = false;
  return 0;
}

int tecno_get_state(tecno_ctx* c, double* prim, int64_t n) {
  if (!c || !prim) return fail(TECNO_EINVAL, "null argument");
  if (n != c->N) return fail(TECNO_EINVAL, "cell count mismatch");
  if (!c->state_set) return fail(TECNO_ESTATE, "no state");
  size_t bytes = (size_t)c->N * c->nprim * sizeof(double);
  if (int rc = ensure_scratch(c, bytes)) return rc;
  HIPCHK(st_con_to_prim(c, c->U[0], c->scratch));
  HIPCHK(hipMemcpyAsync(prim, c->scratch, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// one step, asynchronous (device-side dt and time)
static int launch_step(tecno_ctx* c) {
  const bool linadv = c->cfg.system == TECNO_SYS_SCALAR && c->cfg.model == TECNO_MODEL_LINADV;
  HIPCHK(launch_dt(c->L, c->hmin, c->cfg.cfl, linadv ? 1 : 0, c->P64.vel_norm));
  void *A = c->U[0], *B = c->U[1], *C = c->U[2];
  auto timed = [&](auto fn) -> int {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c->timing) {
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
      HIPCHK(hipEventRecord(e0, c->stream));
    }
    HIPCHK(fn());
    if (c->timing) {
      HIPCHK(hipEventRecord(e1, c->stream));
      c->ev.emplace_back(e0, e1);
    }
    return 0;
  };
  int rc;
  auto st = [&](const void* U, void* Uo, StageCoef k) {
    return timed([&] { return launch_stage(c, U, A, Uo, k); });
  };
  switch (c->cfg.time_scheme) {
    case TECNO_TS_RK1:
      if ((rc = st(A, B, {0.0, 1.0, 1.0, 0.0, 1, 0}))) return rc;
      HIPCHK(hipMemcpyAsync(A, B, (size_t)c->N * c->nvar * c->esize, hipMemcpyDeviceToDevice,
                            c->stream));
      break;
    case TECNO_TS_JAMESON_RK4:
      // every stage restarts from U^n: u = U^n - (dt/vol)*c_k*res(u_prev),
      // c = {1/4, 1/3, 1/2, 1} (fv.cc:1176-1184, ext_constants.h:9).  The last
      // stage writes U^n in place: each cell reads only its own U^n.
      if ((rc = st(A, B, {1.0, 0.0, 1.0 / 4.0, 0.0, 0, 0}))) return rc;
      if ((rc = st(B, C, {1.0, 0.0, 1.0 / 3.0, 0.0, 0, 0}))) return rc;
      if ((rc = st(C, B, {1.0, 0.0, 1.0 / 2.0, 0.0, 0, 0}))) return rc;
      if ((rc = st(B, A, {1.0, 0.0, 1.0, 0.0, 1, 0}))) return rc;
      break;
    case TECNO_TS_RK4:
      // classical RK4 (fv.cc:1185-1197, ext_constants.h:10): stage k adds
      // res/rk4_rks[k] to the accumulator and, before the last, moves to
      // U^n - (dt/vol)*rk4_rks[k+1]*res; the last takes U^n - (dt/vol)*R2/6
      if ((rc = st(A, B, {1.0, 0.0, 1.0 / 2.0, 1.0, 0, 1}))) return rc;
      if ((rc = st(B, C, {1.0, 0.0, 1.0 / 2.0, 2.0, 0, 2}))) return rc;
      if ((rc = st(C, B, {1.0, 0.0, 1.0, 2.0, 0, 2}))) return rc;
      if ((rc = st(B, A, {1.0, 0.0, 1.0, 1.0, 1, 3}))) return rc;
      break;
    default:
      // SSPRK3 (ext_constants.h:7-8): a = {0, 3/4, 1/3}, b = {1, 1/4, 2/3}
      if ((rc = st(A, B, {0.0, 1.0, 1.0, 0.0, 0, 0}))) return rc;
      if ((rc = st(B, C, {3.0 / 4.0, 1.0 / 4.0, 1.0 / 4.0, 0.0, 0, 0}))) return rc;
      if ((rc = st(C, A, {1.0 / 3.0, 2.0 / 3.0, 2.0 / 3.0, 0.0, 1, 0}))) return rc;
      break;
  }
  HIPCHK(launch_finish(c->L));
  return 0;
}

static int set_schedule(tecno_ctx* c, double final_time, int t_stamps, int max_iter) {
  Ctl h{};
  HIPCHK(hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  h.final_time = final_time;
  h.t_int = final_time / t_stamps;
  h.t_save = h.t_int;
  while (h.time >= h.t_save && h.t_save < final_time) h.t_save += h.t_int;
  h.max_iter = max_iter;
  HIPCHK(hipMemcpyAsync(c->ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
  c->configured_final = true;
  c->cfg_final = final_time;
  c->cfg_stamps = t_stamps;
  return 0;
}

int tecno_step(tecno_ctx* c, double t_stamp, double* dt_out) {
  if (!c) return fail(TECNO_EINVAL, "null argument");
  if (!c->state_set) return fail(TECNO_ESTATE, "tecno_set_state first");
  if (int rc = need_net(c)) return rc;
  Ctl h{};
  HIPCHK(hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  h.final_time = INFINITY;
  h.max_iter = 0x7fffffff;
  h.t_save = t_stamp;
  h.t_int = 0.0;
  HIPCHK(hipMemcpyAsync(c->ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
  c->configured_final = false;
  if (int rc = launch_step(c)) return rc;
  HIPCHK(hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (dt_out) *dt_out = h.dt;
  if (h.nbad) return fail(TECNO_ENONFINITE, "non-finite or non-positive density/pressure");
  return 0;
}

int tecno_run_steps(tecno_ctx* c, int n, double final_time, int t_stamps) {
  if (!c) return fail(TECNO_EINVAL, "null argument");
  if (!c->state_set) return fail(TECNO_ESTATE, "tecno_set_state first");
  if (int rc = need_net(c)) return rc;
  if (t_stamps < 1) return fail(TECNO_EINVAL, "t_stamps must be >= 1");
  if (!c->configured_final || c->cfg_final != final_time || c->cfg_stamps != t_stamps)
    if (int rc = set_schedule(c, final_time, t_stamps, 0x7fffffff)) return rc;
  for (int i = 0; i < n; ++i)
    if (int rc = launch_step(c)) return rc;
  return 0;
}

int tecno_advance(tecno_ctx* c, double final_time, int t_stamps, int max_iter, int* iters_out,
                  double* time_out) {
  if (!c) return fail(TECNO_EINVAL, "null argument");
  if (!c->state_set) return fail(TECNO_ESTATE, "tecno_set_state first");
  if (int rc = need_net(c)) return rc;
  if (t_stamps < 1) return fail(TECNO_EINVAL, "t_stamps must be >= 1");
  if (int rc = set_schedule(c, final_time, t_stamps, max_iter)) return rc;
  Ctl h{};
  for (;;) {
    for (int i = 0; i < 32; ++i)
      if (int rc = launch_step(c)) return rc;
    HIPCHK(hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (h.nbad) return fail(TECNO_ENONFINITE, "non-finite or non-positive density/pressure");
    if (!(h.time < h.final_time && h.iter < h.max_iter)) break;
  }
  if (iters_out) *iters_out = h.iter;
  if (time_out) *time_out = h.time;
  return 0;
}

int tecno_get_time(tecno_ctx* c, double* t, int* it) {
  if (!c) return fail(TECNO_EINVAL, "null argument");
  Ctl h{};
  HIPCHK(hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (t) *t = h.time;
  if (it) *it = h.iter;
  return 0;
}

int tecno_synchronize(tecno_ctx* c) {
  if (!c) return fail(TECNO_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

}  // extern "C"
static int finish_audit(tecno_ctx* c, const double acc[9], const Ctl& h, tecno_audit* out);
extern "C" {

int tecno_audit_state(tecno_ctx* c, tecno_audit* out) {
  if (!c || !out) return fail(TECNO_EINVAL, "null argument");
  if (!c->state_set) return fail(TECNO_ESTATE, "no state");
  const int nblk = 1024;
  if (int rc = ensure_scratch(c, (size_t)nblk * 9 * sizeof(double))) return rc;
  if (c->L.state_fp64)
    HIPCHK(StateOps<double>::audit(c->L, c->U[0], c->cfg.gamma, c->scratch, nblk));
  else
    HIPCHK(StateOps<float>::audit(c->L, c->U[0], c->cfg.gamma, c->scratch, nblk));
  std::vector<double> part((size_t)nblk * 9);
  HIPCHK(hipMemcpyAsync(part.data(), c->scratch, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  Ctl h{};
  HIPCHK(hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < nblk; ++b)
    for (int v = 0; v < 9; ++v) acc[v] = (v == 7) ? std::max(acc[v], part[b * 9 + v]) : acc[v] + part[b * 9 + v];
  return finish_audit(c, acc, h, out);
}

// ---- allocation-free audit
// The partial sums of the audit's blocks and their reduced row live in a
// file-scope device array (73.8 KB; every GPU the library runs on holds its
// own copy until the process ends).  Host code takes the array's per-device
// address from hipGetSymbolAddress and never uses its name.  Each call writes
// every row it reads, so nothing has to start at zero; calls that overlap on
// one GPU share the array and race.  The blocks stride over the cells, so
// the array's size does not depend on the grid.
}  // extern "C"

namespace {
constexpr int AUDIT_BLOCKS = 1024, AUDIT_ROW = 9, MAX_DEVICES = 64;
}
static __device__ __align__(256) double scratch[(AUDIT_BLOCKS + 1) * AUDIT_ROW];
static std::mutex g_scratch_mu;
static double* g_scratch_dev[MAX_DEVICES] = {};

// the rows of the blocks reduced in block order, as tecno_audit_state sums
// them on the host (max for the wave speed), into row AUDIT_BLOCKS
__global__ void k_audit_rows(double* part, int nblk) {
  const int v = threadIdx.x;
  if (v >= AUDIT_ROW) return;
  double acc = 0.0;
  for (int b = 0; b < nblk; ++b) {
    const double x = part[(int64_t)b * AUDIT_ROW + v];
    acc = (v == 7) ? (acc < x ? x : acc) : acc + x;
  }
  part[(int64_t)nblk * AUDIT_ROW + v] = acc;
}

// this device's address of the array (first call per device, with the
// device current)
static int audit_scratch(int dev, double** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return fail(TECNO_EINVAL, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(TECNO_EHIP, "hipGetSymbolAddress returned no address");
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

extern "C" {

int tecno_audit_state_noalloc(tecno_ctx* c, tecno_audit* out) {
  if (!c || !out) return fail(TECNO_EINVAL, "null argument");
  if (!c->state_set) return fail(TECNO_ESTATE, "no state");
  int cur = -1;
  HIPCHK(hipGetDevice(&cur));
  if (cur != c->device) HIPCHK(hipSetDevice(c->device));
  double* part = nullptr;
  int rc = audit_scratch(c->device, &part);
  double acc[AUDIT_ROW];
  Ctl h{};
  if (rc == 0) {
    hipError_t e = c->L.state_fp64
                       ? StateOps<double>::audit(c->L, c->U[0], c->cfg.gamma, part, AUDIT_BLOCKS)
                       : StateOps<float>::audit(c->L, c->U[0], c->cfg.gamma, part, AUDIT_BLOCKS);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_audit_rows, dim3(1), dim3(64), 0, c->stream, part, AUDIT_BLOCKS);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(acc, part + (size_t)AUDIT_BLOCKS * AUDIT_ROW, sizeof(acc),
                         hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) rc = fail(TECNO_EHIP, std::string("audit: ") + hipGetErrorString(e));
  }
  if (cur != c->device) (void)hipSetDevice(cur);
  if (rc) return rc;
  return finish_audit(c, acc, h, out);
}

}  // extern "C"

// the audit's sums (block rows reduced) to the reported quantities
static int finish_audit(tecno_ctx* c, const double acc[9], const Ctl& h, tecno_audit* out) {
  const double dV = c->vol * collapsed_cells(c);
  out->total_entropy = acc[0] * dV;
  out->kinetic_energy = 0.5 * acc[1] * dV;
  out->mass = acc[2] * dV;
  out->momentum[0] = acc[3] * dV;
  out->momentum[1] = acc[4] * dV;
  out->momentum[2] = acc[5] * dV;
  out->energy = acc[6] * dV;
  out->max_speed = (c->cfg.system == TECNO_SYS_SCALAR && c->cfg.model == TECNO_MODEL_LINADV)
                       ? c->P64.vel_norm : acc[7];
  out->n_nonfinite = (int64_t)acc[8];
  out->n_sign_fallback = (int64_t)h.nfallback;
  return 0;
}

extern "C" {

int tecno_l1_error(tecno_ctx* c, const double* exact, int64_t n, double* l1) {
  if (!c || !exact || !l1) return fail(TECNO_EINVAL, "null argument");
  if (n != c->N) return fail(TECNO_EINVAL, "cell count mismatch");
  const int nblk = 1024;
  size_t eb = (size_t)c->N * c->nprim * sizeof(double);
  if (int rc = ensure_scratch(c, eb + (size_t)nblk * 5 * sizeof(double))) return rc;
  double* part = c->scratch + (size_t)c->N * c->nprim;
  HIPCHK(hipMemcpyAsync(c->scratch, exact, eb, hipMemcpyHostToDevice, c->stream));
  if (c->L.state_fp64)
    HIPCHK(StateOps<double>::l1(c->L, c->U[0], c->scratch, c->cfg.gamma, part, nblk));
  else
    HIPCHK(StateOps<float>::l1(c->L, c->U[0], c->scratch, c->cfg.gamma, part, nblk));
  std::vector<double> hp((size_t)nblk * 5);
  HIPCHK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int v = 0; v < c->nprim; ++v) {
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += hp[b * 5 + v];
    l1[v] = s * c->vol * collapsed_cells(c);
  }
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ globals
static DiagGrid diag_grid(const tecno_ctx* c) {
  DiagGrid g{};
  g.dims = c->dims;
  for (int d = 0; d < 3; ++d) {
    g.n[d] = c->n[d];
    g.h[d] = c->h[d];
  }
  for (int i = 0; i < 6; ++i) g.bc[i] = c->cfg.bc[i];
  g.N = c->N;
  g.fix_edges = c->cfg.fix_edges;
  return g;
}

static int64_t diag_nodes_held(const tecno_ctx* c) {
  return (int64_t)(c->n[0] + 1) * (c->n[1] + 1) * (c->dims == 3 ? c->n[2] + 1 : 3);
}

// the 2D node plane K (0..Nz of the reference's storage) is held as layer
// 0 (K = 0), 2 (K = Nz) or 1 (in between)
static int plane_layer(int K, int Nz) { return K == 0 ? 0 : (K == Nz ? 2 : 1); }

// nodal diagnostics into the scratch buffer: omega (3 x held nodes) when
// with_omega, then the per-block partial sums; *part_out points at them
static int run_nodes(tecno_ctx* c, bool with_omega, double** omega_out, double** part_out,
                     int* nblk) {
  const int64_t NN = diag_nodes_held(c);
  const int64_t nb = (NN + 63) / 64;
  const size_t need = (size_t)(with_omega ? 3 * NN : 0) + (size_t)nb * 4 + 64;
  if (int rc = ensure_scratch(c, need * sizeof(double))) return rc;
  double* om = with_omega ? c->scratch : nullptr;
  double* part = c->scratch + (with_omega ? 3 * NN : 0);
  const bool ns = c->cfg.model == TECNO_MODEL_NS_ES || c->cfg.model == TECNO_MODEL_NS_KEP;
  HIPCHK(diag_nodes(c->stream, diag_grid(c), (const double*)c->U[0], c->cfg.gamma, c->P64,
                    ns && c->dims == 3, om, part, nblk));
  *omega_out = om;
  *part_out = part;
  return 0;
}

extern "C" {

int tecno_compute_globals(tecno_ctx* c, tecno_globals* out) {
  if (!c || !out) return fail(TECNO_EINVAL, "null argument");
  if (!c->state_set) return fail(TECNO_ESTATE, "no state");
  if (int rc = need_net(c)) return rc;
  tecno_audit a;
  if (int rc = tecno_audit_state(c, &a)) return rc;
  std::memset(out, 0, sizeof(*out));
  Ctl h{};
  HIPCHK(hipMemcpyAsync(&h, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  out->time = h.time;
  out->total_entropy = a.total_entropy;
  if (c->cfg.system == TECNO_SYS_SCALAR) return 0;
  out->kinetic_energy = a.kinetic_energy;
  const double layers = collapsed_cells(c);
  // KE_diss_pr: the inviscid residual of the state (the first stage's
  // residual, fv.cc:756-764) from one stage launch that writes the residual
  // itself: u = 0*u + 0*u0 - (dt/vol)*(-1)*res with dt = vol, exactly res.
  // The control block (dt, counters) is restored afterwards.
  {
    Ctl m = h;
    m.dt = c->vol;
    m.active = 1;
    HIPCHK(hipMemcpyAsync(c->ctl, &m, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
    const StageCoef k{0.0, 0.0, -1.0, 0.0, 0, 0};
    hipError_t e;
    if (fp64(c)) {
      Phys<double> P = c->P64;
      P.model = TECNO_MODEL_EULER;
      e = Launch<double>::stage(c->L, P, c->U[0], c->U[0], c->U[1], k);
    } else {
      Phys<float> P = c->P32;
      P.model = TECNO_MODEL_EULER;
      e = Launch<float>::stage(c->L, P, c->U[0], c->U[0], c->U[1], k);
    }
    HIPCHK(e);
    HIPCHK(hipMemcpyAsync(c->ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
  }
  const int nbr = 1024;
  if (int rc = ensure_scratch(c, (size_t)nbr * sizeof(double))) return rc;
  HIPCHK(diag_ke_diss_pr(c->stream, diag_grid(c), (const double*)c->U[0], (const double*)c->U[1],
                         c->scratch, nbr));
  std::vector<double> pr(nbr);
  HIPCHK(hipMemcpyAsync(pr.data(), c->scratch, nbr * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double kdp = 0.0;
  for (double v : pr) kdp += v;
  out->ke_diss_pr = kdp * layers;
  // nodal sums: enstrophy (compute_global.cc:32-75) and KE_diss_stress
  double *om, *part;
  int nblk = 0;
  if (int rc = run_nodes(c, false, &om, &part, &nblk)) return rc;
  std::vector<double> hp((size_t)nblk * 4);
  HIPCHK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  double S[4] = {0, 0, 0, 0};
  for (int b = 0; b < nblk; ++b)
    for (int q = 0; q < 4; ++q) S[q] += hp[(size_t)b * 4 + q];
  const tecno_cfg& g = c->cfg;
  // nodal volume: (extent / (cells + 1)) per axis, the reference's cell
  // counts (compute_global.cc:70-72)
  double nvol = 1.0;
  for (int d = 0; d < 3; ++d) nvol *= (g.hi[d] - g.lo[d]) / (g.n[d] + 1);
  double ens = S[0];
  if (c->dims == 2) {
    // every z node plane of the reference's storage, with its weight
    const int Nz = g.n[2];
    ens = 0.0;
    for (int K = 0; K <= Nz; ++K) {
      double w = 1.0;
      if (K == 0 && g.bc[4] == TECNO_BC_PERIODIC) w /= 2.0;
      if (K == Nz && g.bc[5] == TECNO_BC_PERIODIC) w /= 2.0;
      ens += w * S[plane_layer(K, Nz)];
    }
  }
  out->enstrophy = 0.5 * ens * nvol;
  out->ke_diss_stress = S[3];
  return 0;
}

int tecno_vorticity(tecno_ctx* c, double* omega, int64_t n_nodes) {
  if (!c || !omega) return fail(TECNO_EINVAL, "null argument");
  if (c->cfg.system == TECNO_SYS_SCALAR) return fail(TECNO_EINVAL, "vorticity: systems solver only");
  if (!c->state_set) return fail(TECNO_ESTATE, "no state");
  const tecno_cfg& g = c->cfg;
  const int nx = c->n[0] + 1, ny = c->n[1] + 1;
  const int nzr = (c->dims == 3 ? c->n[2] : g.n[2]) + 1;
  const int64_t NNr = (int64_t)nx * ny * nzr;
  if (n_nodes != NNr) return fail(TECNO_EINVAL, "node count mismatch");
  double *om, *part;
  int nblk = 0;
  if (int rc = run_nodes(c, true, &om, &part, &nblk)) return rc;
  const int64_t NN = diag_nodes_held(c);
  std::vector<double> h((size_t)3 * NN);
  HIPCHK(hipMemcpyAsync(h.data(), om, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->dims == 3) {
    std::memcpy(omega, h.data(), h.size() * sizeof(double));
    return 0;
  }
  const int64_t plane = (int64_t)nx * ny;
  for (int d = 0; d < 3; ++d)
    for (int K = 0; K < nzr; ++K)
      std::memcpy(omega + d * NNr + K * plane, h.data() + d * NN + plane_layer(K, nzr - 1) * plane,
                  plane * sizeof(double));
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------ seams
template <typename F>
static int with_device_io(tecno_ctx* c, std::vector<std::pair<const double*, size_t>> in,
                          std::vector<std::pair<double*, size_t>> out, F launch) {
  size_t tot = 0;
  for (auto& p : in) tot += p.second;
  for (auto& p : out) tot += p.second;
  if (int rc = ensure_scratch(c, tot * sizeof(double) + 64)) return rc;
  std::vector<double*> din, dout;
  double* q = c->scratch;
  for (auto& p : in) {
    HIPCHK(hipMemcpyAsync(q, p.first, p.second * sizeof(double), hipMemcpyHostToDevice, c->stream));
    din.push_back(q);
    q += p.second;
  }
  for (auto& p : out) {
    dout.push_back(q);
    q += p.second;
  }
  HIPCHK(launch(din, dout));
  for (size_t i = 0; i < out.size(); ++i)
    HIPCHK(hipMemcpyAsync(out[i].first, dout[i], out[i].second * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" {

int tecno_reconstruct_batch(tecno_ctx* c, const double* st, int64_t n, double* l, double* r) {
  if (!c || !st || !l || !r) return fail(TECNO_EINVAL, "null argument");
  if (n <= 0) return 0;
  if (int rc = need_net(c)) return rc;
  const int ng = c->cfg.recon == TECNO_REC_ENO ? c->cfg.eno_level : 2;
  return with_device_io(c, {{st, (size_t)n * 2 * ng}}, {{l, (size_t)n}, {r, (size_t)n}},
                        [&](std::vector<double*>& i, std::vector<double*>& o) {
                          return fp64(c) ? Launch<double>::recon_batch(c->L, c->P64, i[0], ng, n, o[0], o[1])
                                         : Launch<float>::recon_batch(c->L, c->P32, i[0], ng, n, o[0], o[1]);
                        });
}

int tecno_mlp_forward_batch(tecno_ctx* c, const double* u4, int64_t n, double* l, double* r) {
  if (!c || !u4 || !l || !r) return fail(TECNO_EINVAL, "null argument");
  if (!c->has_net) return fail(TECNO_ESTATE, "no network loaded");
  if (!c->P64.net.W || !c->P32.net.W || !c->db32)
    return fail(TECNO_ESTATE, "network weights are not on the device");
  if (n <= 0) return 0;
  if ((c->cfg.mlp_dtype == TECNO_MLP_BF16 || c->cfg.mlp_dtype == TECNO_MLP_FP16) && !c->P32.arch5)
    return fail(TECNO_EINVAL, "the matrix-core network needs the 5-5-5-5-5 architecture");
  if (c->cfg.mlp_dtype == TECNO_MLP_BF16 || c->cfg.mlp_dtype == TECNO_MLP_FP16)
    return with_device_io(c, {{u4, (size_t)n * 4}}, {{l, (size_t)n}, {r, (size_t)n}},
                          [&](std::vector<double*>& i, std::vector<double*>& o) {
                            const bool h = c->cfg.mlp_dtype == TECNO_MLP_FP16;
                            return mlp_mfma_batch(c->stream, c->P32.mfma_w[h ? 1 : 0], h, i[0], n,
                                                  o[0], o[1]);
                          });
  return with_device_io(c, {{u4, (size_t)n * 4}}, {{l, (size_t)n}, {r, (size_t)n}},
                        [&](std::vector<double*>& i, std::vector<double*>& o) {
                          return (c->cfg.mlp_dtype == TECNO_MLP_FP64)
                                     ? Launch<double>::mlp_batch(c->L, c->P64, i[0], n, o[0], o[1])
                                     : Launch<float>::mlp_batch(c->L, c->P32, i[0], n, o[0], o[1]);
                        });
}

int tecno_flux_batch(tecno_ctx* c, const double* P, const double* V, int64_t n, int axis,
                     double* F) {
  if (!c || !P || !F) return fail(TECNO_EINVAL, "null argument");
  if (axis < 0 || axis > 1) return fail(TECNO_EINVAL, "axis must be 0 or 1 (2D faces)");
  if (n <= 0) return 0;
  if (int rc = need_net(c)) return rc;
  const int ng = c->G;
  const bool sc = c->cfg.system == TECNO_SYS_SCALAR;
  if (!sc && !V) return fail(TECNO_EINVAL, "entropy variables required");
  size_t np = sc ? (size_t)n * 2 * ng : (size_t)n * 2 * ng * 5;
  std::vector<std::pair<const double*, size_t>> in = {{P, np}};
  if (!sc) in.push_back({V, np});
  return with_device_io(c, in, {{F, sc ? (size_t)n : (size_t)n * 5}},
                        [&](std::vector<double*>& i, std::vector<double*>& o) {
                          const double* vin = sc ? nullptr : i[1];
                          return fp64(c) ? Launch<double>::flux_batch(c->L, c->P64, i[0], vin, ng, n, axis, o[0])
                                         : Launch<float>::flux_batch(c->L, c->P32, i[0], vin, ng, n, axis, o[0]);
                        });
}

// ------------------------------------------------------------------ timing
int tecno_timing_enable(tecno_ctx* c, int on) {
  if (!c) return fail(TECNO_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& e : c->ev) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  c->ev.clear();
  c->t_ms_acc = 0;
  c->t_n = 0;
  c->timing = on != 0;
  return 0;
}

int tecno_timing_get(tecno_ctx* c, double* ms_avg, int64_t* n) {
  if (!c) return fail(TECNO_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& e : c->ev) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
    c->t_ms_acc += ms;
    c->t_n += 1;
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  c->ev.clear();
  if (ms_avg) *ms_avg = c->t_n ? c->t_ms_acc / c->t_n : 0.0;
  if (n) *n = c->t_n;
  return 0;
}

// ------------------------------------------------------------------ training
int tecno_train_init(tecno_ctx* c, const tecno_adam* opt) {
  if (!c || !opt) return fail(TECNO_EINVAL, "null argument");
  if (!c->has_net) return fail(TECNO_ESTATE, "no network loaded");
  if (!c->net.nl || c->net.nl > 4) return fail(TECNO_EINVAL, "training supports up to 4 layers");
  for (int l = 0; l <= c->net.nl; ++l)
    if (c->net.w[l] > 8) return fail(TECNO_EINVAL, "training supports widths up to 8");
  delete c->trainer;
  c->trainer = new MlpTrainer();
  std::string err;
  int rc = c->trainer->init(c->stream, c->net.nl, c->net.w, c->net.W.data(), c->net.b.data(),
                            opt->lr, opt->beta1, opt->beta2, opt->eps, opt->weight_decay, err);
  if (rc) {
    delete c->trainer;
    c->trainer = nullptr;
    return fail(TECNO_EHIP, err);
  }
  return 0;
}

static int sync_net_from_trainer(tecno_ctx* c) {
  std::vector<double> W, b;
  c->trainer->get_weights(W, b);
  for (int l = 0; l < c->net.nl; ++l)
    for (int o = 0; o < c->net.w[l + 1]; ++o) {
      for (int i = 0; i < c->net.w[l]; ++i)
        c->net.W[(size_t)l * TECNO_NET_MAXW * TECNO_NET_MAXW + o * TECNO_NET_MAXW + i] = W[(size_t)l * 64 + o * 8 + i];
      c->net.b[(size_t)l * TECNO_NET_MAXW + o] = b[(size_t)l * 8 + o];
    }
  return upload_net(c);
}

int tecno_train_step(tecno_ctx* c, const double* rows, int64_t n, int use_bf16, double* loss) {
  if (!c || !rows) return fail(TECNO_EINVAL, "null argument");
  if (!c->trainer) return fail(TECNO_ESTATE, "tecno_train_init first");
  if (n <= 0) return fail(TECNO_EINVAL, "empty batch");
  std::string err;
  double l = 0;
  if (c->trainer->step(rows, n, use_bf16 != 0, true, &l, nullptr, err)) return fail(TECNO_EHIP, err);
  if (loss) *loss = l;
  return sync_net_from_trainer(c);
}

int tecno_train_grad(tecno_ctx* c, const double* rows, int64_t n, int use_bf16, double* loss,
                     double* grad) {
  if (!c || !rows || !grad) return fail(TECNO_EINVAL, "null argument");
  if (!c->trainer) return fail(TECNO_ESTATE, "tecno_train_init first");
  if (n <= 0) return fail(TECNO_EINVAL, "empty batch");
  std::string err;
  std::vector<double> g;
  double l = 0;
  if (c->trainer->step(rows, n, use_bf16 != 0, false, &l, &g, err)) return fail(TECNO_EHIP, err);
  // trainer layout [l][8x8] / [l][8] -> packed W then b
  size_t q = 0;
  for (int L = 0; L < c->net.nl; ++L)
    for (int o = 0; o < c->net.w[L + 1]; ++o)
      for (int i = 0; i < c->net.w[L]; ++i) grad[q++] = g[(size_t)L * 64 + o * 8 + i];
  for (int L = 0; L < c->net.nl; ++L)
    for (int o = 0; o < c->net.w[L + 1]; ++o) grad[q++] = g[4 * 64 + (size_t)L * 8 + o];
  if (loss) *loss = l;
  return 0;
}

int tecno_train_set_data(tecno_ctx* c, const double* rows, int64_t n) {
  if (!c || !rows) return fail(TECNO_EINVAL, "null argument");
  if (!c->trainer) return fail(TECNO_ESTATE, "tecno_train_init first");
  std::string err;
  if (c->trainer->set_data(rows, n, err)) return fail(TECNO_EINVAL, err);
  return 0;
}

int tecno_train_epoch(tecno_ctx* c, int64_t batch, uint64_t seed, int use_bf16, double* loss) {
  if (!c) return fail(TECNO_EINVAL, "null argument");
  if (!c->trainer) return fail(TECNO_ESTATE, "tecno_train_init first");
  std::string err;
  double l = 0;
  if (c->trainer->epoch(batch, seed, use_bf16 != 0, &l, err)) return fail(TECNO_EHIP, err);
  if (loss) *loss = l;
  return sync_net_from_trainer(c);
}

int tecno_make_train_rows(const double* u4, const double* target, int64_t n, double* rows) {
  if (!u4 || !target || !rows) return fail(TECNO_EINVAL, "null argument");
  if (n <= 0) return 0;
  std::string err;
  if (make_train_rows(nullptr, u4, target, n, rows, err)) return fail(TECNO_EHIP, err);
  return 0;
}

int tecno_save_net(const tecno_ctx* c, const char* path) {
  if (!c || !path) return fail(TECNO_EINVAL, "null argument");
  if (!c->has_net) return fail(TECNO_ESTATE, "no network loaded");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(TECNO_EIO, std::string("cannot write ") + path);
  std::fprintf(f, "Weights\n");
  for (int l = 0; l < c->net.nl; ++l) {
    std::fprintf(f, "Layer %d\nBEGIN\n", l + 1);
    for (int o = 0; o < c->net.w[l + 1]; ++o) {
      for (int i = 0; i < c->net.w[l]; ++i)
        std::fprintf(f, i ? " %.18e" : "%.18e",
                     c->net.W[(size_t)l * TECNO_NET_MAXW * TECNO_NET_MAXW + o * TECNO_NET_MAXW + i]);
      std::fprintf(f, "\n");
    }
    std::fprintf(f, "END\n");
  }
  std::fprintf(f, "Biases\n");
  for (int l = 0; l < c->net.nl; ++l) {
    std::fprintf(f, "Layer %d\nBEGIN\n", l + 1);
    for (int o = 0; o < c->net.w[l + 1]; ++o)
      std::fprintf(f, "%.18e\n", c->net.b[(size_t)l * TECNO_NET_MAXW + o]);
    std::fprintf(f, "END\n");
  }
  const bool ok = std::fclose(f) == 0;
  return ok ? 0 : fail(TECNO_EIO, std::string("write failed: ") + path);
}

int tecno_trainer_epoch_seeds(int nl, const int32_t* widths, uint64_t seed, int n_epochs,
                              uint64_t* out) {
  if (!widths || (!out && n_epochs > 0)) return fail(TECNO_EINVAL, "null argument");
  if (nl < 1 || nl > TECNO_NET_MAXL) return fail(TECNO_EINVAL, "bad layer count");
  if (n_epochs < 0) return fail(TECNO_EINVAL, "negative epoch count");
  for (int l = 0; l <= nl; ++l)
    if (widths[l] < 1 || widths[l] > TECNO_NET_MAXW)
      return fail(TECNO_EINVAL, "layer widths must be 1..16");
  trainer_epoch_seeds(nl, widths, seed, n_epochs, out);
  return 0;
}

int tecno_init_net(tecno_ctx* c, int nl, const int32_t* widths, uint64_t seed) {
  if (!c || !widths) return fail(TECNO_EINVAL, "null argument");
  if (nl < 1 || nl > TECNO_NET_MAXL) return fail(TECNO_EINVAL, "bad layer count");
  if (widths[0] != 5 || widths[nl] != 5)
    return fail(TECNO_EINVAL, "the network maps 5 inputs to 5 vertex weights");
  for (int l = 0; l <= nl; ++l)
    if (widths[l] < 1 || widths[l] > TECNO_NET_MAXW)
      return fail(TECNO_EINVAL, "layer widths must be 1..16");
  std::vector<double> W, b;
  init_net_torch(nl, widths, seed, W, b);
  return tecno_set_net(c, nl, widths, W.data(), b.data());
}

// ------------------------------------------------------------------ ensembles
}  // extern "C"

namespace {
struct EnsGroupH {
  int prec = ENS_FP64;  // ENS_* of the members (their cfg.precision)
  int dims = 1;
  int variant = 0;
  int spw = 1;  // every member reconstructs with the SP-WENO family
  std::vector<int> members;
  void* dprobs = nullptr;
  EnsWork* dwork = nullptr;
  int nwork = 0;
  int grid = 0;
  int gridf = 1;  // workgroups launched per resident one (the stride loop takes any count)
  double tuned_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};
}  // namespace

struct tecno_ens {
  std::vector<tecno_ctx*> m;
  std::vector<EnsGroupH> g;
  EnsCtl* dctl = nullptr;
  int* dcount = nullptr;
  hipStream_t s = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  int n_cu = 0;
};

static int ens_nvariants(int dims) { return dims == 1 ? ENS_NV1 : ENS_NV2; }

static void ens_drop_graph(tecno_ens* e) {
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->gexec = nullptr;
  e->graph = nullptr;
  e->graph_valid = false;
}

// work list of a group for its current tile shape
static int ens_build_work(tecno_ens* e, EnsGroupH& G) {
  const int TX = ens_tx(G.dims, G.variant), TY = ens_ty(G.dims, G.variant);
  std::vector<EnsWork> w;
  for (int k = 0; k < (int)G.members.size(); ++k) {
    const tecno_ctx* c = e->m[G.members[k]];
    const int tx = (c->n[0] + TX - 1) / TX;
    const int ty = G.dims == 2 ? (c->n[1] + TY - 1) / TY : 1;
    for (int by = 0; by < ty; ++by)
      for (int bx = 0; bx < tx; ++bx) w.push_back(EnsWork{k, bx, by, 0});
  }
  if (G.dwork) HIPCHK(hipFree(G.dwork));
  G.dwork = nullptr;
  G.nwork = (int)w.size();
  HIPCHK(hipMalloc(&G.dwork, w.size() * sizeof(EnsWork)));
  HIPCHK(hipMemcpy(G.dwork, w.data(), w.size() * sizeof(EnsWork), hipMemcpyHostToDevice));
  const int resident = ens_grid(G.prec, G.dims, G.variant, G.spw, e->n_cu);
  G.grid = std::max(1, std::min(G.nwork, resident * G.gridf));
  ens_drop_graph(e);
  return 0;
}

template <typename T>
static int ens_upload_probs_t(tecno_ens* e, EnsGroupH& G, const Phys<T> tecno_ctx::*pm) {
  std::vector<EnsProb<T>> h(G.members.size());
  for (size_t k = 0; k < G.members.size(); ++k) {
    tecno_ctx* c = e->m[G.members[k]];
    h[k].P = c->*pm;
    h[k].g = c->L.grid;
    h[k].ctl = c->ctl;
    for (int q = 0; q < 3; ++q) h[k].U[q] = c->U[q];
  }
  if (!G.dprobs) HIPCHK(hipMalloc(&G.dprobs, h.size() * sizeof(EnsProb<T>)));
  HIPCHK(hipMemcpy(G.dprobs, h.data(), h.size() * sizeof(EnsProb<T>), hipMemcpyHostToDevice));
  return 0;
}

static int ens_upload(tecno_ens* e) {
  std::vector<EnsCtl> hc(e->m.size());
  for (size_t i = 0; i < e->m.size(); ++i) {
    tecno_ctx* c = e->m[i];
    const bool linadv = c->cfg.model == TECNO_MODEL_LINADV;
    hc[i] = EnsCtl{c->ctl, c->hmin, c->cfg.cfl, c->P64.vel_norm, linadv ? 1 : 0, 0};
  }
  if (!e->dctl) HIPCHK(hipMalloc(&e->dctl, hc.size() * sizeof(EnsCtl)));
  HIPCHK(hipMemcpy(e->dctl, hc.data(), hc.size() * sizeof(EnsCtl), hipMemcpyHostToDevice));
  for (auto& G : e->g) {
    const int rc = G.prec == ENS_FP32 ? ens_upload_probs_t<float>(e, G, &tecno_ctx::P32)
                          : ens_upload_probs_t<double>(e, G, &tecno_ctx::P64);
    if (rc) return rc;
  }
  return 0;
}

// one ensemble step on stream s: dt of every member, 3 stages per group, end
static hipError_t ens_step(tecno_ens* e, int only_group) {
  const int n = (int)e->m.size();
  hipError_t r = ens_control(e->s, e->dctl, n, 0);
  for (int st = 0; st < 3 && r == hipSuccess; ++st)
    for (int gi = 0; gi < (int)e->g.size() && r == hipSuccess; ++gi) {
      if (only_group >= 0 && gi != only_group) continue;
      const EnsGroupH& G = e->g[gi];
      r = ens_stage(e->s, G.prec, G.dims, G.variant, G.spw, G.dprobs, G.dwork, G.nwork, st,
                    G.grid);
    }
  if (r == hipSuccess) r = ens_control(e->s, e->dctl, n, 1);
  return r;
}

static int ens_capture(tecno_ens* e) {
  ens_drop_graph(e);
  HIPCHK(hipStreamBeginCapture(e->s, hipStreamCaptureModeGlobal));
  const hipError_t r = ens_step(e, -1);
  hipGraph_t graph = nullptr;
  const hipError_t r2 = hipStreamEndCapture(e->s, &graph);
  HIPCHK(r);
  HIPCHK(r2);
  e->graph = graph;
  HIPCHK(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0));
  e->graph_valid = true;
  return 0;
}

static int ens_sync_members(tecno_ens* e) {
  for (tecno_ctx* c : e->m) HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" {

int tecno_ens_create(const tecno_cfg* cfgs, int n, tecno_ens** out) {
  if (!cfgs || !out || n < 1) return fail(TECNO_EINVAL, "null argument or empty ensemble");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const tecno_cfg& g = cfgs[i];
    if (g.system != TECNO_SYS_SCALAR || g.dims > 2)
      return fail(TECNO_EINVAL, "ensembles hold 1D or 2D scalar problems");
    if (g.time_scheme != TECNO_TS_SSPRK3) return fail(TECNO_EINVAL, "ensembles step with SSPRK3");
    if (g.recon == TECNO_REC_ENO && g.eno_level == 3)
      return fail(TECNO_EINVAL, "ensembles use reconstructions with 2 ghost cells");
  }
  tecno_ens* e = new tecno_ens();
  auto bail = [&](int rc) {
    tecno_ens_destroy(e);
    return rc;
  };
  for (int i = 0; i < n; ++i) {
    tecno_ctx* c = nullptr;
    if (int rc = tecno_create(&cfgs[i], &c)) return bail(rc);
    e->m.push_back(c);
    if (c->G != 2) return bail(fail(TECNO_EINVAL, "ensembles use reconstructions with 2 ghost cells"));
    const int prec = cfgs[i].precision == TECNO_FP32           ? ENS_FP32
                     : cfgs[i].precision == TECNO_FP32_STATE64 ? ENS_FP32_STATE64
                                                               : ENS_FP64;
    const int dims = cfgs[i].dims;
    int gi = -1;
    for (int q = 0; q < (int)e->g.size(); ++q)
      if (e->g[q].prec == prec && e->g[q].dims == dims) gi = q;
    if (gi < 0) {
      e->g.emplace_back();
      gi = (int)e->g.size() - 1;
      e->g[gi].prec = prec;
      e->g[gi].dims = dims;
    }
    e->g[gi].members.push_back(i);
    if (c->P64.recon != REC_SPWENO) e->g[gi].spw = 0;
  }
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return bail(fail(TECNO_EHIP, "hipGetDeviceProperties failed"));
  e->n_cu = prop.multiProcessorCount;
  if (hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(TECNO_EHIP, "hipStreamCreate failed"));
  if (hipMalloc(&e->dcount, sizeof(int)) != hipSuccess)
    return bail(fail(TECNO_ENOMEM, "hipMalloc failed"));
  for (auto& G : e->g)
    if (int rc = ens_build_work(e, G)) return bail(rc);
  *out = e;
  return 0;
}

void tecno_ens_destroy(tecno_ens* e) {
  if (!e) return;
  if (e->s) (void)hipStreamSynchronize(e->s);
  ens_drop_graph(e);
  for (auto& G : e->g) {
    if (G.dprobs) (void)hipFree(G.dprobs);
    if (G.dwork) (void)hipFree(G.dwork);
  }
  if (e->dctl) (void)hipFree(e->dctl);
  if (e->dcount) (void)hipFree(e->dcount);
  for (tecno_ctx* c : e->m) tecno_destroy(c);
  if (e->s) (void)hipStreamDestroy(e->s);
  delete e;
}

int tecno_ens_size(const tecno_ens* e) { return e ? (int)e->m.size() : 0; }

tecno_ctx* tecno_ens_member(tecno_ens* e, int i) {
  if (!e || i < 0 || i >= (int)e->m.size()) return nullptr;
  return e->m[i];
}

int tecno_ens_num_groups(const tecno_ens* e) { return e ? (int)e->g.size() : 0; }

int tecno_ens_group_info(const tecno_ens* e, int gi, tecno_ens_group* out) {
  if (!e || !out || gi < 0 || gi >= (int)e->g.size()) return fail(TECNO_EINVAL, "bad group");
  const EnsGroupH& G = e->g[gi];
  std::memset(out, 0, sizeof(*out));
  out->precision = G.prec == ENS_FP32           ? TECNO_FP32
                  : G.prec == ENS_FP32_STATE64 ? TECNO_FP32_STATE64
                                               : TECNO_FP64;
  out->dims = G.dims;
  out->n_members = (int)G.members.size();
  out->n_tiles = G.nwork;
  out->grid = G.grid;
  out->tile_x = ens_tx(G.dims, G.variant);
  out->tile_y = ens_ty(G.dims, G.variant);
  out->n_candidates = ens_nvariants(G.dims);
  for (int v = 0; v < out->n_candidates; ++v) {
    out->cand_x[v] = ens_tx(G.dims, v);
    out->cand_y[v] = ens_ty(G.dims, v);
    out->tuned_ms[v] = G.tuned_ms[v];
  }
  return 0;
}

int tecno_ens_advance(tecno_ens* e, const double* final_times, int t_stamps, int max_iter,
                      int use_graph, int64_t* cell_updates) {
  if (!e || !final_times) return fail(TECNO_EINVAL, "null argument");
  if (t_stamps < 1) return fail(TECNO_EINVAL, "t_stamps must be >= 1");
  std::vector<int> it0(e->m.size());
  for (size_t i = 0; i < e->m.size(); ++i) {
    tecno_ctx* c = e->m[i];
    if (!c->state_set) return fail(TECNO_ESTATE, "tecno_set_state on every member first");
    if (int rc = need_net(c)) return rc;
    if (int rc = set_schedule(c, final_times[i], t_stamps, max_iter)) return rc;
  }
  if (int rc = ens_sync_members(e)) return rc;
  for (size_t i = 0; i < e->m.size(); ++i) {
    Ctl h{};
    HIPCHK(hipMemcpy(&h, e->m[i]->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    it0[i] = h.iter;
  }
  if (int rc = ens_upload(e)) return rc;
  if (use_graph && !e->graph_valid)
    if (int rc = ens_capture(e)) return rc;
  int active = 1;
  while (active) {
    for (int r = 0; r < 16; ++r) {
      if (use_graph)
        HIPCHK(hipGraphLaunch(e->gexec, e->s));
      else
        HIPCHK(ens_step(e, -1));
    }
    HIPCHK(ens_count_active(e->s, e->dctl, (int)e->m.size(), e->dcount));
    HIPCHK(hipMemcpyAsync(&active, e->dcount, sizeof(int), hipMemcpyDeviceToHost, e->s));
    HIPCHK(hipStreamSynchronize(e->s));
  }
  int64_t cu = 0;
  bool bad = false;
  for (size_t i = 0; i < e->m.size(); ++i) {
    Ctl h{};
    HIPCHK(hipMemcpy(&h, e->m[i]->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
    cu += (int64_t)(h.iter - it0[i]) * e->m[i]->N;
    bad = bad || h.nbad;
  }
  if (cell_updates) *cell_updates = cu;
  if (bad) return fail(TECNO_ENONFINITE, "non-finite state in an ensemble member");
  return 0;
}

int tecno_ens_autotune(tecno_ens* e, int steps) {
  if (!e || steps < 1) return fail(TECNO_EINVAL, "bad argument");
  for (tecno_ctx* c : e->m)
    if (!c->state_set) return fail(TECNO_ESTATE, "tecno_set_state on every member first");
  if (int rc = ens_sync_members(e)) return rc;
  // save every member's state and control block
  std::vector<void*> save(e->m.size(), nullptr);
  std::vector<Ctl> ctl(e->m.size());
  auto cleanup = [&]() {
    for (void* p : save)
      if (p) (void)hipFree(p);
  };
  for (size_t i = 0; i < e->m.size(); ++i) {
    tecno_ctx* c = e->m[i];
    const size_t bytes = (size_t)c->N * c->nvar * c->esize;
    if (hipMalloc(&save[i], bytes) != hipSuccess) {
      cleanup();
      return fail(TECNO_ENOMEM, "autotune: no memory for the state copies");
    }
    HIPCHK(hipMemcpy(save[i], c->U[0], bytes, hipMemcpyDeviceToDevice));
    HIPCHK(hipMemcpy(&ctl[i], c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost));
  }
  auto restore = [&]() -> int {
    for (size_t i = 0; i < e->m.size(); ++i) {
      tecno_ctx* c = e->m[i];
      HIPCHK(hipMemcpy(c->U[0], save[i], (size_t)c->N * c->nvar * c->esize,
                       hipMemcpyDeviceToDevice));
      Ctl h = ctl[i];
      h.final_time = INFINITY;  // keep every member stepping while timed
      h.t_save = INFINITY;
      h.t_int = INFINITY;
      h.max_iter = 0x7fffffff;
      HIPCHK(hipMemcpy(c->ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice));
    }
    return 0;
  };
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  int rc = ens_upload(e);
  for (int gi = 0; gi < (int)e->g.size() && !rc; ++gi) {
    EnsGroupH& G = e->g[gi];
    double best = 1e300;
    int bestv = 0, bestf = 1;
    // tile shape x workgroups per resident slot (1: one persistent wave of
    // workgroups as the occupancy query counts them; 2, 4: more, which the
    // hardware schedules as slots free up)
    for (int v = 0; v < ens_nvariants(G.dims) && !rc; ++v) {
      G.tuned_ms[v] = 0.0;
      for (int f = 1; f <= 4 && !rc; f *= 2) {
        G.variant = v;
        G.gridf = f;
        if ((rc = ens_build_work(e, G)) || (rc = restore())) break;
        if (ens_step(e, gi) != hipSuccess) {  // warm-up step, not timed
          rc = fail(TECNO_EHIP, "autotune launch failed");
          break;
        }
        (void)hipEventRecord(e0, e->s);
        for (int k = 0; k < steps; ++k) (void)ens_step(e, gi);
        (void)hipEventRecord(e1, e->s);
        if (hipEventSynchronize(e1) != hipSuccess) {
          rc = fail(TECNO_EHIP, "autotune: kernel failed");
          break;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (G.tuned_ms[v] == 0.0 || ms / steps < G.tuned_ms[v]) G.tuned_ms[v] = ms / steps;
        if (ms < best) {
          best = ms;
          bestv = v;
          bestf = f;
        }
      }
    }
    if (!rc) {
      G.variant = bestv;
      G.gridf = bestf;
      rc = ens_build_work(e, G);
    }
  }
  // put the states and control blocks back as they were
  for (size_t i = 0; i < e->m.size() && !rc; ++i) {
    tecno_ctx* c = e->m[i];
    if (hipMemcpy(c->U[0], save[i], (size_t)c->N * c->nvar * c->esize,
                  hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(c->ctl, &ctl[i], sizeof(Ctl), hipMemcpyHostToDevice) != hipSuccess)
      rc = fail(TECNO_EHIP, "autotune: restoring the state failed");
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  cleanup();
  return rc;
}

}  // extern "C"

