/*
 * tecno_oracle.h — CPU FP64 restatement of the reference TeCNO step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product (libtecno.so).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never calls it.
 *
 * Every function follows the reference's arithmetic in operation order
 * (plain IEEE double, no FMA contraction, glibc libm), so that on the same
 * host it reproduces the reference build bit for bit.  Citations are to
 * files under pmcharles/DSP-WENO.
 */
#ifndef TECNO_ORACLE_H
#define TECNO_ORACLE_H

#include <stdint.h>
#include "../include/tecno.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TOR_MAX_LAYERS 8
#define TOR_MAX_WIDTH 16

typedef struct tor_net {
  int n_layers;
  int width[TOR_MAX_LAYERS + 1];            /* width[0] = inputs (5)        */
  double W[TOR_MAX_LAYERS][TOR_MAX_WIDTH * TOR_MAX_WIDTH]; /* out x in      */
  double b[TOR_MAX_LAYERS][TOR_MAX_WIDTH];
} tor_net;

/* ---- point functions ---------------------------------------------------- */
double tor_logavg(double a, double b);
void tor_sp_weno(const double u[4], int corr_type, double* l, double* r);
double tor_eno_pw(int k, const double* s);
/* 0.5*minmod(a, b) or 0.5*minabs(a, b): reconstruct.cc:162-186, limiter.h:20-52 */
double tor_limited_slope(double a, double b, int use_minabs);
void tor_eno_rec(int k, const double* st, int ng, double* l, double* r);
int tor_net_load(const char* path, tor_net* net);
int tor_net_save(const char* path, const tor_net* net);
void tor_nn_forward(const tor_net* net, const double u[4], double* l,
                    double* r);
void tor_nn_vertices(double theta_m, double theta_p, double max_delta,
                     double xv[5], double yv[5]);
void tor_nn_features(const double u[4], double feat[5], double xv[5],
                     double yv[5]);
/* reconstruct() dispatch on one component stencil of 2*ng values */
void tor_reconstruct(const tecno_cfg* cfg, const tor_net* net,
                     const double* st, int ng, double* l, double* r);

/* ---- batched point functions (ctypes friendly) -------------------------- */
void tor_sp_weno_batch(int corr, const double* u4, int64_t n, double* l,
                       double* r);
void tor_reconstruct_batch(const tecno_cfg* cfg, const tor_net* net,
                           const double* st, int ng, int64_t n, double* l,
                           double* r);
void tor_nn_forward_batch(const tor_net* net, const double* u4, int64_t n,
                          double* l, double* r);
void tor_logavg_batch(const double* a, const double* b, int64_t n,
                      double* out);
/* Face fluxes.  Scalar: P is n x 2ng, F n x 1.  Euler: P, V are n x 2ng x 5,
 * F n x 5.  h[3] = dx, dy, dz. */
void tor_flux_batch(const tecno_cfg* cfg, const tor_net* net, const double* P,
                    const double* V, int ng, int64_t n, int axis,
                    const double h[3], double* F);
/* Entropy variables of primitive states (FV::prim_to_ent): n x 5 -> n x 5 */
void tor_prim_to_ent_batch(const tecno_cfg* cfg, const double* P, int64_t n,
                           double* V);

/* ---- whole-grid solver -------------------------------------------------- */
typedef struct tor_solver tor_solver;
/* faithful = 1: all three directions are swept with the param.in cell counts,
 * exactly as enflo/enfloS do (a "2D" run is Nz = 2 with z faces).
 * faithful = 0: only the first cfg->dims directions are active; inactive
 * directions hold one cell and keep the param.in spacing. */
/* the ES viscous node fluxes (momentum x, y, z and energy of the x, y and z
 * fluxes, before the face areas) from the node average of V and its
 * differences along x, y, z (vflux_ES.cc:74-145; cfg->fix_visc keeps the
 * K21, K31, K32 parts the reference overwrites) */
void tor_es_node_flux(const tecno_cfg* cfg, const double avg[5], const double Tx[5],
                      const double Ty[5], const double Tz[5], double out[12]);
tor_solver* tor_create(const tecno_cfg* cfg, const tor_net* net, int faithful);
void tor_destroy(tor_solver* s);
int64_t tor_num_cells(const tor_solver* s);
int tor_num_vars(const tor_solver* s);
/* primitive state, interior SoA (scalar: u; Euler: rho,u,v,w,p) */
int tor_set_prim(tor_solver* s, const double* prim);
int tor_get_prim(const tor_solver* s, double* prim);
int tor_step(tor_solver* s, double t_stamp, double* dt_out);
int tor_advance(tor_solver* s, double final_time, int t_stamps, int max_iter,
                int* iters_out, double* time_out);
void tor_set_threads(int n);
/* FP64 audit: out[0] total entropy, out[1] kinetic energy, out[2] mass,
 * out[3..5] momentum, out[6] energy, out[7] max speed */
void tor_audit(const tor_solver* s, double out[8]);

/* ---- trainer (DSP-WENO_trainer/DSP-WENO_network_trainer.py) --------------
 * A training row holds 21 values: the 5 network inputs (tanh theta-,
 * tanh theta+, |D0|, |D1|, |D2| of the scaled stencil), 5 x-vertices,
 * 5 y-vertices, the 4 unscaled cells and the 2 target interface values. */
#define TOR_ROW 21
void tor_train_rows(const double* u4, const double* target, int64_t n,
                    double* rows);
/* Loss 0.5 (MSE(left, tL) + MSE(right, tR)) over n rows and its gradient
 * (same layout as the network: W[l][out * TOR_MAX_WIDTH + in], b[l][out]).
 * bf16 != 0 rounds the weights and every activation to bfloat16 (nearest
 * even) before use, as a bfloat16-activation forward does. */
double tor_train_loss_grad(const tor_net* net, const double* rows, int64_t n,
                           int bf16, tor_net* grad);
/* torch.optim.Adam step t (1-based) with L2 weight decay added to the
 * gradient; m, v are the moment buffers in the network layout. */
void tor_adam_step(tor_net* net, const tor_net* grad, tor_net* m, tor_net* v,
                   int64_t t, double lr, double beta1, double beta2, double eps,
                   double weight_decay);

#ifdef __cplusplus
}
#endif
#endif
