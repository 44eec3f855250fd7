/*
 * tecno.h — C ABI of the MI355X TeCNO finite-volume library (libtecno.so).
 *
 * The reference solvers (enflo / enfloS) expose no library API: the hot path
 * sits behind the CLI + param.in contract and behind three const member
 * functions of class FV that run once per face per Runge–Kutta stage:
 *   FV::num_flux          enflo/src/fv.h:180-183, enflo_scalar/src_scalar/num_flux.cc:12-57
 *   FV::reconstruct       enflo/src/fv.h:89-93,   enflo/src/reconstruct.cc:192-239
 *   Neural_Network::forward enflo/src/neural_network.h:25
 * and the driver loop FV::run (enflo/src/fv.cc:1231-1324,
 * enflo_scalar/src_scalar/fv.cc:522-607).  A per-face foreign call would forbid
 * GPU batching, so this ABI works one level up: a context owns a grid that
 * lives in HBM and is advanced one whole SSPRK3 step (three stages plus the
 * CFL time step) per call.  Batched per-face entry points mirror the
 * num_flux / reconstruct / forward seams for testing.
 *
 * Conventions
 *   - every function returns 0 on success and a negative TECNO_E* code on
 *     failure; tecno_last_error() gives the message (thread-local).
 *   - host arrays are structure-of-arrays, x fastest, interior cells only
 *     (no ghost layers):  a[v][k][j][i] at v*N + (k*ny + j)*nx + i.
 *   - a context is not thread-safe; separate contexts may be used from
 *     separate threads.  All device work of a context runs on its own stream.
 */
#ifndef TECNO_H
#define TECNO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
enum {
  TECNO_OK = 0,
  TECNO_EINVAL = -1,   /* bad argument / unsupported option combination   */
  TECNO_EHIP = -2,     /* HIP runtime error                                */
  TECNO_ENOMEM = -3,   /* device allocation failed                         */
  TECNO_EIO = -4,      /* file could not be read / parsed                  */
  TECNO_ESTATE = -5,   /* call out of order (e.g. step before set_state)   */
  TECNO_ENONFINITE = -6/* NaN / non-positive density or pressure detected  */
};

/* ---- enumerations (names follow param.in keywords) --------------------- */
/* system solved: enfloS (scalar, NVAR=1) or enflo (Euler/NS, NVAR=5)       */
enum { TECNO_SYS_SCALAR = 0, TECNO_SYS_EULER = 1 };

/* model: enflo_scalar parameter.cc:236-262, enflo/src/parameter.cc:286-298   */
enum { TECNO_MODEL_LINADV = 0, TECNO_MODEL_BURGERS = 1,
       TECNO_MODEL_EULER = 10, TECNO_MODEL_NS_KEP = 11, TECNO_MODEL_NS_ES = 12 };

/* scalar fluxes: enflo_scalar/src_scalar/ext_constants.h:17 (same order)   */
enum { TECNO_FLUX_LIN_EC = 0, TECNO_FLUX_LIN_EC4, TECNO_FLUX_LIN_TECNO,
       TECNO_FLUX_LIN_TECNO4, TECNO_FLUX_BUR_EC, TECNO_FLUX_BUR_EC4,
       TECNO_FLUX_BUR_TECNO, TECNO_FLUX_BUR_TECNO4 };

/* Euler fluxes: enflo/src/parameter.cc:308-394.  The TeCNO fluxes combine an
 * entropy-conservative flux (KEPEC, KEPEC4, Ismail-Roe, Ismail-Roe 4th order,
 * KEP) with Roe-type or Rusanov-type entropy dissipation.                    */
enum { TECNO_FLUX_KEPES_TECNO_ROE = 100, TECNO_FLUX_KEPES_TECNO4_ROE,
       TECNO_FLUX_KEPES_TECNO_RUSANOV, TECNO_FLUX_KEPES_TECNO4_RUSANOV,
       TECNO_FLUX_ROE_TECNO_ROE, TECNO_FLUX_ROE_TECNO4_ROE,
       TECNO_FLUX_KEP_TECNO_ROE,
       TECNO_FLUX_KEPEC = 120, TECNO_FLUX_KEPEC4, TECNO_FLUX_ROE_EC,
       TECNO_FLUX_ROE_EC4, TECNO_FLUX_KEP,
       /* fluxes without entropy dissipation on the reconstructed jumps:
        * Roe without entropy fix (roe.cc:10-86), KEP 4th order (kep4.cc),
        * KEPS and KEPS 4th order (keps.cc, keps4.cc)                        */
       TECNO_FLUX_ROE = 125, TECNO_FLUX_KEP4, TECNO_FLUX_KEPS, TECNO_FLUX_KEPS4 };

/* reconstruction: enflo/src/parameter.cc:155-212.  sp_wenoc and
 * sp_weno_corr2 are TECNO_REC_SP_WENO with spweno_corr = 1 / 2.  second:
 * unlimited MUSCL (kappa = 1/3), minmod / minabs: MUSCL with the limited
 * slope (reconstruct.cc:30-80, 162-186).                                  */
enum { TECNO_REC_SP_WENO = 0, TECNO_REC_ENO = 1, TECNO_REC_SP_WENO_DL = 2,
       TECNO_REC_FIRST = 3, TECNO_REC_SECOND = 4, TECNO_REC_MINMOD = 5,
       TECNO_REC_MINABS = 6 };

/* boundary types: enflo/src/ext_constants.h:23 (same values)               */
enum { TECNO_BC_PERIODIC = 0, TECNO_BC_DIRICHLET = 1, TECNO_BC_NEUMANN = 2,
       TECNO_BC_WALL = 3 };

/* time schemes: enflo/src/ext_constants.h:20 (same values).  rk4 holds one
 * more state-sized device buffer, the accumulated residual.                */
enum { TECNO_TS_RK1 = 0, TECNO_TS_SSPRK3 = 1, TECNO_TS_JAMESON_RK4 = 2, TECNO_TS_RK4 = 3 };

/* arithmetic of the face kernels.
 * TECNO_FP32: scalar problems hold and compute their state in FP32; the 2D
 *   Euler FP32 path keeps state, cell records, the entropy-conservative flux
 *   and the eigen-projection in FP64 and reconstructs in FP32.
 * TECNO_FP32_STATE64 (scalar problems with sp_weno / sp_wenoc / corr2): the
 *   Euler path's split for scalars -- FP64 state, update and
 *   entropy-conservative flux, the SP-WENO jump in FP32 from the FP64 cell
 *   jumps.  An FP32 state's rounding (~6e-8 of |u| per step) exceeds
 *   third-order errors on fine meshes; this keeps the order.               */
enum { TECNO_FP64 = 0, TECNO_FP32 = 1, TECNO_FP32_STATE64 = 2 };

/* arithmetic of the DSP-WENO network evaluation (sp_weno_dl only) */
enum { TECNO_MLP_FP64 = 0,   /* SIMT, double                               */
       TECNO_MLP_FP32 = 1,   /* SIMT, float                                */
       TECNO_MLP_BF16 = 2,   /* MFMA bf16 x bf16 -> fp32                   */
       TECNO_MLP_FP16 = 3 }; /* MFMA fp16 x fp16 -> fp32                   */

/* viscosity law: enflo/src/parameter.cc:236-276 */
enum { TECNO_MU_CONSTANT = 0, TECNO_MU_SUTHERLAND = 1, TECNO_MU_POWER = 2 };

/* ---- configuration (the param.in fields of the hot path) --------------- */
typedef struct tecno_cfg {
  int32_t system;        /* TECNO_SYS_*                                      */
  int32_t model;         /* TECNO_MODEL_*                                    */
  int32_t flux;          /* TECNO_FLUX_*                                     */
  int32_t recon;         /* TECNO_REC_*                                      */
  int32_t spweno_corr;   /* 0 sp_weno, 1 sp_wenoc, 2 sp_weno_corr2           */
  int32_t eno_level;     /* k of "eno k" (2 or 3)                            */
  int32_t time_scheme;   /* TECNO_TS_*                                       */
  int32_t precision;     /* TECNO_FP64 / TECNO_FP32 / TECNO_FP32_STATE64     */
  int32_t mlp_dtype;     /* TECNO_MLP_* (sp_weno_dl)                         */
  int32_t dims;          /* active directions: 1, 2 or 3 (leading axes)     */
  int32_t n[3];          /* Nx_cell, Ny_cell, Nz_cell as written in param.in */
  double lo[3], hi[3];   /* x_min..z_max                                     */
  int32_t bc[6];         /* xleft, xright, yleft, yright, zleft, zright      */
  double cfl;
  double vel[3];         /* linear advection velocity                        */
  double gamma, gas_const, prandtl;
  int32_t mu_model;
  double mu_ref, T_ref, T_0, omega;
  int32_t ec1_fix;       /* EC1_fix yes/no                                   */
  int32_t fix_zmom;      /* 0: z-momentum dissipation added to y-momentum as
                            the reference does; 1: corrected                */
  int32_t sign_fallback; /* sp_weno_dl: replace a face component by SP-WENO
                            where the sign property fails                   */
  int32_t fix_edges;     /* Navier-Stokes: 0 adds the viscous flux of a node
                            on an edge of two periodic boundaries twice, as
                            the reference's boundary loops do
                            (fv.cc:821-1156); 1 once.  Nodes on a neumann
                            or wall boundary count once either way.        */
  int32_t fix_visc;      /* ns_es: 0 drops the K21 V_x term of the y viscous
                            flux and the K31 V_x, K32 V_y terms of the z
                            flux, which the reference's nodal_vlux_ES adds
                            and then overwrites (vflux_ES.cc:103-145), so
                            its viscous matrix is not symmetric; 1 keeps
                            them (K symmetric positive semi-definite)      */
  int32_t reserved[6];
} tecno_cfg;

typedef struct tecno_ctx tecno_ctx;

/* Fill *cfg with defaults (2D Euler, kepes_tecno4_roe, sp_weno, fp64). */
void tecno_cfg_default(tecno_cfg* cfg);

/* Parameter::read (enflo/src/parameter.cc:16-29 with its section readers
 * 35-700; enflo_scalar/src_scalar/parameter.cc for `model linadv|burgers`
 * files): parse a param.in file into the tecno_cfg of mesh level `level`
 * (dims from problem_dim, at least 2 for systems; n, lo, hi of that level;
 * precision FP64 and the fix flags 0 - set them afterwards).  Optional
 * outputs (NULL to skip): the number of mesh levels FV::run solves (every
 * listed mesh with OOC_study yes, else 1: parameter.cc:636-639; `level` may
 * address any listed mesh), final_time,
 * time_stamps, max_iter, and net_file copied into a buffer of net_len
 * bytes.  A malformed file or an unknown name returns TECNO_EINVAL with the
 * reader's message. */
int tecno_read_param(const char* path, int level, tecno_cfg* cfg, int* n_levels,
                     double* final_time, int* time_stamps, int* max_iter, char* net_file,
                     int net_len);

/* Create a context; allocates device state.  Replaces FV::FV + FV::initialize
 * (enflo/src/fv.h:24-67, fv.cc:18-93). */
int tecno_create(const tecno_cfg* cfg, tecno_ctx** out);
void tecno_destroy(tecno_ctx* ctx);

/* Load a DSP-WENO net_file (Neural_Network::load_network,
 * enflo/src/neural_network.cc:82-130).  Layer widths are taken from the block
 * shapes; the last layer must have 5 outputs. */
int tecno_load_net(tecno_ctx* ctx, const char* net_file);
/* Same, from arrays: n_layers, widths[n_layers+1], W (row-major out x in,
 * layers concatenated), b (concatenated). */
int tecno_set_net(tecno_ctx* ctx, int n_layers, const int32_t* widths,
                  const double* W, const double* b);
int tecno_get_net(const tecno_ctx* ctx, int* n_layers, int32_t* widths,
                  double* W, double* b);
/* Write the context's network in the net_file format (the trainer's
 * save_net_weights, DSP-WENO_network_trainer.py:197-213: "Weights", per layer
 * "Layer k" / "BEGIN" / one out-row per line / "END", then "Biases" likewise,
 * values as %.18e). */
int tecno_save_net(const tecno_ctx* ctx, const char* net_file);
/* Seeded initial network of n_layers with the given widths (5 in, 5 out),
 * the trainer's initial network for seed_no = seed: its draws from PyTorch's
 * CPU generator after torch.manual_seed(seed) (DSP-WENO_network_trainer.py:
 * 296) -- nn.Linear's default initialisation of each layer in order (biases
 * U(+-1/sqrt(fan_in)); :54-56), then nn.init.xavier_uniform_ of every weight
 * (:58-60) -- reproduced bit for bit (MT19937 on the low 32 bits of seed,
 * PyTorch's 53-bit float64 uniform with a fused multiply-add). */
int tecno_init_net(tecno_ctx* ctx, int n_layers, const int32_t* widths,
                   uint64_t seed);

/* Number of interior cells held by the context (product of active n[d]). */
int64_t tecno_num_cells(const tecno_ctx* ctx);
/* Number of primitive variables per cell: 1 (scalar), 4 (2D Euler: rho,u,v,p)
 * or 5 (3D Euler: rho,u,v,w,p). */
int tecno_num_vars(const tecno_ctx* ctx);

/* Upload primitive state (interior, SoA).  Host -> device copy, caller keeps
 * ownership.  Resets time and iteration count. */
int tecno_set_state(tecno_ctx* ctx, const double* prim, int64_t n_cells);
/* Download primitive state (con_to_prim of the conserved state). */
int tecno_get_state(tecno_ctx* ctx, double* prim, int64_t n_cells);

/* One SSPRK3 step: dt = cfl*hmin/max_speed, clamped to t_stamp (FV::run,
 * enflo/src/fv.cc:1264-1266), three stages.  Returns dt taken. */
int tecno_step(tecno_ctx* ctx, double t_stamp, double* dt_out);
/* Time loop of FV::run: steps until time >= final_time or max_iter, with
 * dt clamped to the t_stamps output instants (fv.cc:1256-1310). */
int tecno_advance(tecno_ctx* ctx, double final_time, int t_stamps,
                  int max_iter, int* iters_out, double* time_out);
/* Run exactly n_steps steps with no host synchronisation in between (dt and
 * time are kept on the device; a step whose time is already >= final_time
 * does nothing).  Used for timing. */
int tecno_run_steps(tecno_ctx* ctx, int n_steps, double final_time,
                    int t_stamps);
int tecno_get_time(tecno_ctx* ctx, double* time_out, int* iters_out);
int tecno_synchronize(tecno_ctx* ctx);

/* FP64 audit (compute_global, enflo/src/compute_global.cc:10-89;
 * enflo_scalar/src_scalar/compute_global.cc:10-38). */
typedef struct tecno_audit {
  double total_entropy;  /* Euler: sum -rho*s/(gamma-1) dV; scalar: sum u^2/2 dV */
  double kinetic_energy; /* Euler: 0.5*sum rho|u|^2 dV                      */
  double mass, momentum[3], energy; /* conservation sums (times dV)         */
  double max_speed;      /* max over cells of |u|+c (scalar: |a| or |u|)    */
  int64_t n_nonfinite;   /* cells with NaN/Inf or rho<=0 or p<=0            */
  int64_t n_sign_fallback;/* sp_weno_dl face components replaced by SP-WENO
                            (accumulated since set_state)                   */
} tecno_audit;
int tecno_audit_state(tecno_ctx* ctx, tecno_audit* out);

/* tecno_audit_state without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * stepping loops that audit every step, harnesses that forbid allocation;
 * tecno_audit_state may allocate its staging buffer on first use).  Same
 * kernels and the same summation order, so the same results bit for bit.
 * The partial sums live in a fixed 73.8 KB device array that every GPU the
 * library runs on holds until the process ends; its size does not depend on
 * the grid (the audit's 1024 blocks stride over the cells), so no grid is
 * rejected.  Calls that overlap on one GPU share that array and race; run
 * them one at a time per GPU.  The call makes the context's GPU current for
 * its duration. */
int tecno_audit_state_noalloc(tecno_ctx* ctx, tecno_audit* out);

/* L1 error of the primitive state against exact values (compute_error,
 * enflo/src/compute_errors.cc:9-34): out[v] = sum |exact-prim| * dV.
 * Sums here and in tecno_audit_state are the reference's sums: a cell of
 * a collapsed axis (a 2D grid's z, a 1D grid's y and z) stands for the
 * n[d] cells the reference stores there, so dV uses that axis's extent. */
int tecno_l1_error(tecno_ctx* ctx, const double* exact_prim, int64_t n_cells,
                   double* l1_out);

/* FV::compute_global (enflo/src/compute_global.cc:10-89;
 * enflo_scalar/src_scalar/compute_global.cc:10-38) for the current state,
 * the values of one globals.dat line (the reference evaluates them at the
 * first stage of each step, i.e. for the state the step starts from):
 *   total_entropy  sum -rho s/(gamma-1) dV (scalar: sum u^2/2 dV)
 *   kinetic_energy 0.5 sum rho|u|^2 dV
 *   ke_diss_pr     fv.cc:756-764: -sum over cells of (|u|^2/2 R_mass +
 *                  u . R_momentum), R the inviscid residual of the state
 *   ke_diss_stress fv.cc:792-1156: Navier-Stokes only (0 otherwise), minus
 *                  the node sum of (grad u . viscous momentum flux) h w, edge
 *                  nodes of two periodic planes twice and the last two
 *                  corners' derivatives and weights taken at node
 *                  (i, Nz, Nz), as the reference's loops do (fix_edges:
 *                  each node once, at itself)
 *   enstrophy      compute_global.cc:32-75: 0.5 * nodal volume * node sum of
 *                  (8-cell mean density) |vorticity|^2 w
 * Systems solver: one extra stage launch (the residual) and one node
 * pass. */
typedef struct tecno_globals {
  double time;
  double total_entropy;
  double kinetic_energy;
  double ke_diss_pr;
  double ke_diss_stress;
  double enstrophy;
} tecno_globals;
int tecno_compute_globals(tecno_ctx* ctx, tecno_globals* out);

/* FV::find_vorticity (enflo/src/find_vort.cc:9-29): the vorticity at the
 * reference's (Nx+1)(Ny+1)(Nz+1) nodes (a 2D grid: Nz = n[2] as given, the
 * node planes of its z-uniform storage), omega[c*n_nodes + (k*(Ny+1) + j)*
 * (Nx+1) + i] for components c = x, y, z.  Ghost cells as the reference
 * fills them (apply_bc, fv.cc:121-361: edge and corner ghosts periodic).
 * Systems solver only. */
int tecno_vorticity(tecno_ctx* ctx, double* omega, int64_t n_nodes);

/* ---- batched seams (test / kernel-level entry points) ------------------ */
/* FV::reconstruct on n stencils of 2*ng values (ng = 2, or k for eno k):
 * st[i*2ng + j].  l[i], r[i] = reconstructed left/right face values. */
int tecno_reconstruct_batch(tecno_ctx* ctx, const double* st, int64_t n,
                            double* l, double* r);
/* FV::num_flux on n faces along axis (0,1,2): P and V are n x 2ng x 5
 * (rho,u,v,w,p / entf,entv.x,entv.y,entv.z,entl), F is n x 5 (including face
 * area).  Scalar: P is n x 2ng, V ignored, F is n x 1. */
int tecno_flux_batch(tecno_ctx* ctx, const double* P, const double* V,
                     int64_t n, int axis, double* F);
/* Neural_Network::forward on n raw 4-cell stencils (no 1e-15 skip). */
int tecno_mlp_forward_batch(tecno_ctx* ctx, const double* u4, int64_t n,
                            double* l, double* r);

/* ---- DSP-WENO training (DSP-WENO_network_trainer.py:384-450) ---------- */
/* One Adam step on a batch of rows.  Each row holds 15 features (the 5 MLP
 * inputs, 5 x-vertices, 5 y-vertices), the 4 raw cell values and the 2
 * target interface values: rows x 21 doubles.  Loss is
 * (MSE(left, tL) + MSE(right, tR)) / 2.  Master weights are fp32; the forward
 * activations are rounded to bf16 when use_bf16 != 0. */
typedef struct tecno_adam {
  double lr, beta1, beta2, eps, weight_decay;
  int64_t step;          /* number of steps taken (bias correction)         */
} tecno_adam;
int tecno_train_init(tecno_ctx* ctx, const tecno_adam* opt);
int tecno_train_step(tecno_ctx* ctx, const double* rows, int64_t n_rows,
                     int use_bf16, double* loss_out);
/* Loss and gradient without an update (gradient check).  grad layout = W then
 * b, as tecno_get_net. */
int tecno_train_grad(tecno_ctx* ctx, const double* rows, int64_t n_rows,
                     int use_bf16, double* loss_out, double* grad_out);
/* Keep n_rows training rows (n x 21, as above) on the device for
 * tecno_train_epoch. */
int tecno_train_set_data(tecno_ctx* ctx, const double* rows, int64_t n_rows);
/* One pass over the device data set in mini-batches of `batch` rows (the
 * last batch may be shorter), one Adam step per batch.  The rows are taken in
 * the order of the trainer's DataLoader(shuffle=True) (trainer:348, 392)
 * whose RandomSampler drew `seed`: torch.randperm(n_rows) under a generator
 * seeded with it (at most 214748364 rows).  loss_out = mean row loss over the
 * pass. */
int tecno_train_epoch(tecno_ctx* ctx, int64_t batch, uint64_t seed,
                      int use_bf16, double* loss_out);
/* The shuffle seeds of the trainer's first n_epochs epochs when it runs with
 * seed_no = seed and a network of n_layers with these widths
 * (DSP-WENO_network_trainer.py:296, 351, 384-392): after the network's
 * initial draws (tecno_init_net), every epoch's DataLoader iterator draws its
 * base seed and then its sampler's seed from PyTorch's global generator.
 * Passing seeds_out[e] to tecno_train_epoch of epoch e takes the trainer's
 * row order.  Host-only; no context or device needed. */
int tecno_trainer_epoch_seeds(int n_layers, const int32_t* widths, uint64_t seed,
                              int n_epochs, uint64_t* seeds_out);
/* Build training rows from raw 4-cell stencils and targets on the device:
 * u4 (n x 4), target (n x 2) -> rows (n x 21), computing features and
 * vertices as Neural_Network::forward does (scaled stencil features,
 * compute_vertices with vertex_mode 2, permutation 2). */
int tecno_make_train_rows(const double* u4, const double* target, int64_t n,
                          double* rows);

/* ---- timing ------------------------------------------------------------ */
/* Average duration (ms) of the named kernel family over the launches since
 * the last reset, measured with HIP events on the context's stream. */
int tecno_timing_enable(tecno_ctx* ctx, int on);
int tecno_timing_get(tecno_ctx* ctx, double* stage_ms_avg, int64_t* n_stages);

/* ---- batched ensembles of scalar problems -------------------------------
 * No counterpart in the reference, which runs its meshes one after the other
 * (FV::run's OOC loop, enflo_scalar/src_scalar/fv.cc:522-607): independent
 * scalar problems (1D or 2D, any sizes, FP64 or FP32, SSPRK3, a reconstruction
 * with 2 ghost cells) advance together.  Members of one precision and
 * dimension share one persistent launch per stage that strides over all
 * their tiles; one ensemble step (dt of every member, 3 stages, end of step)
 * is captured once as a hipGraph and replayed.  Each member keeps its own dt,
 * time and stamps; a member that has reached its end time skips its tiles.
 * Members are ordinary contexts for set_state / get_state / l1_error /
 * audit / get_time. */
typedef struct tecno_ens tecno_ens;
int tecno_ens_create(const tecno_cfg* cfgs, int n, tecno_ens** out);
void tecno_ens_destroy(tecno_ens* ens);
int tecno_ens_size(const tecno_ens* ens);
tecno_ctx* tecno_ens_member(tecno_ens* ens, int i);
/* Advance every member i to final_times[i] (t_stamps output stamps each,
 * at most max_iter steps).  use_graph = 0 launches the kernels one by one
 * instead of replaying the captured graph.  cell_updates_out = sum over
 * members of steps x cells. */
int tecno_ens_advance(tecno_ens* ens, const double* final_times, int t_stamps,
                      int max_iter, int use_graph, int64_t* cell_updates_out);
/* Autotuner: for every group (precision x dimension), time `steps` ensemble
 * steps with each candidate tile shape on the members' current states (which
 * are restored afterwards) and keep the fastest. */
int tecno_ens_autotune(tecno_ens* ens, int steps);
typedef struct tecno_ens_group {
  int precision;        /* TECNO_FP64 / TECNO_FP32 / TECNO_FP32_STATE64 */
  int dims;
  int n_members;
  int n_tiles;          /* work items of one stage                          */
  int grid;             /* workgroups launched (striding over the tiles)    */
  int tile_x, tile_y;   /* chosen tile shape                                */
  int n_candidates;
  int cand_x[8], cand_y[8];
  double tuned_ms[8];   /* autotuner time per candidate tile shape, best   *
                         * over 1, 2, 4 workgroups per resident slot       *
                         * (0 = not timed)                                 */
} tecno_ens_group;
int tecno_ens_num_groups(const tecno_ens* ens);
int tecno_ens_group_info(const tecno_ens* ens, int g, tecno_ens_group* out);

const char* tecno_last_error(void);
const char* tecno_version(void);

#ifdef __cplusplus
}
#endif
#endif /* TECNO_H */
