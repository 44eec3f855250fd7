/*
 * tecno_oracle.c — CPU FP64 restatement of the reference TeCNO step.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; see tecno_oracle.h).
 *
 * Operation order follows the reference line by line where rounding depends
 * on it; the citations name the reference file and lines each function
 * restates.  Build with gcc -O2 -ffp-contract=off (no FMA), as the reference
 * is built with g++ -O3 for baseline x86-64, which has no FMA.
 */
#include "tecno_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SPW_EPS 1.0e-13 /* limiter.h:14 */

static double dmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
static double dmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

/* ------------------------------------------------------------------------ */
/* logarithmic mean: enflo/src/misc_func.h:14-32                            */
double tor_logavg(double a, double b) {
  double xi = b / a;
  double f = (xi - 1.0) / (xi + 1.0);
  double u = f * f;
  double F;
  if (u < 1.0e-2) {
    double u2 = u * u;
    double u3 = u2 * u;
    F = 1.0 + u / 3.0 + u2 / 5.0 + u3 / 7.0;
  } else {
    F = log(xi) / 2.0 / f;
  }
  return 0.5 * (a + b) / F;
}

/* ------------------------------------------------------------------------ */
/* SP-WENO: limiter.h:103-118 (fp), 153-187 (Cp), 193-252 (sp_weno_rec)     */
static double spw_fp(double theta0, double theta1) {
  double psi, val;
  if (fabs(theta0 - 1.0) < SPW_EPS)
    psi = 0.0;
  else
    psi = (1 - theta1) / (1 - theta0);
  if (fabs(theta0 - 1.0) > SPW_EPS && fabs(psi + 1.0) > SPW_EPS)
    val = 1.0 / (1.0 + psi);
  else
    val = 1.0;
  return val;
}

static void spw_cp(double theta0, double theta1, double del1, double scale,
                   double* C, double* corr) {
  double psi;
  if (fabs(theta0 - 1.0) < SPW_EPS || fabs(theta1 - 1.0) < SPW_EPS)
    psi = 0.0;
  else
    psi = (1 - theta1) / (1 - theta0);
  *corr = 0.0;
  if (fabs(psi + 1.0) > SPW_EPS && fabs(theta0 - 1.0) > SPW_EPS && psi < 0) {
    double a = spw_fp(theta0, theta1), b = spw_fp(theta1, theta0);
    *C = (a / (pow(a, 2.0) + pow(b, 2.0))) / 8.0;
    *corr += -0.25 * pow(dmin(fabs(del1 / scale), fabs(del1)), 3.0) /
             (1.0 - theta0);
  } else if (fabs(theta0 - 1.0) > SPW_EPS && fabs(psi + 1.0) < SPW_EPS) {
    *C = 0.0;
  } else if (fabs(theta0 - 1.0) < SPW_EPS ||
             (psi > -SPW_EPS && fabs(theta0) < 1 + SPW_EPS)) {
    *C = -3.0 / 8.0;
  } else if (psi > -SPW_EPS && fabs(theta0) > 1) {
    *C = 1.0 / 8.0;
  }
  /* else: "unknown region" — C keeps the caller's value (limiter.h:180-185) */
}

void tor_sp_weno(const double s[4], int corr_type, double* recl,
                 double* recr) {
  double del0 = s[1] - s[0];
  double del1 = s[2] - s[1];
  double del2 = s[3] - s[2];
  if (fabs(del1) < SPW_EPS) {
    *recl = s[1];
    *recr = s[2];
    return;
  }
  double tp = del0 / del1;
  double tm = del2 / del1;
  double scale = 0.5 * (fabs(s[1]) + fabs(s[2]));
  double C1 = 0, C2 = 0, C1c = 0, C2c = 0;
  spw_cp(tp, tm, del1, scale, &C1, &C1c);
  spw_cp(tm, tp, del1, scale, &C2, &C2c);
  if (corr_type > 0) {
    double C1n = C1 + C1c;
    double C2n = C2 + C2c;
    if (corr_type == 1) {
      C1 = C1n;
      C2 = C2n;
    }
    C1 = dmin(dmax(C1, -3.0 / 8.0), 1.0 / 8.0);
    C2 = dmin(dmax(C2, -3.0 / 8.0), 1.0 / 8.0);
  }
  double w0 = 3.0 / 4.0 + 2 * C1;
  double w0t = 1.0 / 4.0 - 2 * C2;
  *recl = w0 * 0.5 * (s[1] + s[2]) + (1 - w0) * (-0.5 * s[0] + 1.5 * s[1]);
  *recr = (1 - w0t) * 0.5 * (s[1] + s[2]) + w0t * (-0.5 * s[3] + 1.5 * s[2]);
}

/* ------------------------------------------------------------------------ */
/* pointwise ENO: limiter.h:8-12 (coefficients), 63-97                       */
static const double ENO2C[2][2] = {{-0.5, 1.5}, {0.5, 0.5}};
static const double ENO3C[3][3] = {{0.375, -1.25, 1.875},
                                   {-0.125, 0.75, 0.375},
                                   {0.375, 0.75, -0.125}};

double tor_eno_pw(int k, const double* state) {
  double dd[3][5];
  double result = 0;
  int r = 0;
  for (int j = 0; j < 2 * k - 1; ++j) dd[0][j] = state[j];
  for (int i = 1; i < k; ++i)
    for (int j = 0; j < 2 * k - 1 - i; ++j) dd[i][j] = dd[i - 1][j + 1] - dd[i - 1][j];
  int s = k - 1;
  for (int i = 1; i < k; ++i)
    if (fabs(dd[i][s - r]) > fabs(dd[i][s - r - 1])) r++;
  for (int i = 0; i < k; ++i) {
    if (k == 2)
      result += ENO2C[s - r][i] * state[s - r + i];
    else if (k == 3)
      result += ENO3C[s - r][i] * state[s - r + i];
  }
  return result;
}

/* reconstruct_eno_pw: enflo_scalar/src_scalar/reconstruct.cc:80-104 */
void tor_eno_rec(int k, const double* st, int ng, double* l, double* r) {
  double comp[5] = {0, 0, 0, 0, 0};
  int s = 2 * ng;
  for (int j = 0; j < 2 * k - 1; ++j) comp[j] = st[s / 2 - k + j];
  *l = tor_eno_pw(k, comp);
  for (int j = 0; j < 2 * k - 1; ++j) comp[2 * k - 2 - j] = st[s / 2 - k + 1 + j];
  *r = tor_eno_pw(k, comp);
}

/* ------------------------------------------------------------------------ */
/* DSP-WENO network: enflo/src/neural_network.cc                             */
int tor_net_load(const char* path, tor_net* net) {
  /* load_network: neural_network.cc:82-130.  Blocks between BEGIN and END;
   * rows of a weight block are output neurons; a "Biases" line switches to
   * bias blocks. */
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  memset(net, 0, sizeof(*net));
  char line[8192];
  int weight = 1, nw = 0, nb = 0;
  int rows = 0, cols = 0, in_block = 0, nbias = 0;
  while (fgets(line, sizeof line, f)) {
    size_t L = strlen(line);
    while (L && (line[L - 1] == '\n' || line[L - 1] == '\r')) line[--L] = 0;
    if (!in_block) {
      if (strcmp(line, "BEGIN") == 0) {
        in_block = 1;
        rows = cols = nbias = 0;
      } else if (strcmp(line, "Biases") == 0) {
        weight = 0;
      }
      continue;
    }
    if (strcmp(line, "END") == 0) {
      in_block = 0;
      if (weight) {
        if (nw >= TOR_MAX_LAYERS) { fclose(f); return -2; }
        net->width[nw] = cols;     /* inputs of layer nw      */
        net->width[nw + 1] = rows; /* outputs of layer nw     */
        nw++;
      } else {
        if (nb >= TOR_MAX_LAYERS) { fclose(f); return -2; }
        nb++;
      }
      continue;
    }
    char* p = line;
    int c = 0;
    for (;;) {
      char* e;
      double v = strtod(p, &e);
      if (e == p) break;
      p = e;
      if (weight) {
        if (rows >= TOR_MAX_WIDTH || c >= TOR_MAX_WIDTH) { fclose(f); return -2; }
        net->W[nw][rows * TOR_MAX_WIDTH + c] = v;
      } else {
        if (nbias >= TOR_MAX_WIDTH) { fclose(f); return -2; }
        net->b[nb][nbias++] = v;
      }
      c++;
    }
    if (weight) {
      if (c > cols) cols = c;
      rows++;
    }
  }
  fclose(f);
  net->n_layers = nw;
  if (nw == 0 || nb != nw) return -3;
  return 0;
}

int tor_net_save(const char* path, const tor_net* net) {
  /* format written by save_net_weights, DSP-WENO_network_trainer.py:197-213 */
  FILE* f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "Weights\n");
  for (int l = 0; l < net->n_layers; ++l) {
    fprintf(f, "Layer %d\nBEGIN\n", l + 1);
    for (int o = 0; o < net->width[l + 1]; ++o) {
      for (int i = 0; i < net->width[l]; ++i)
        fprintf(f, "%s%.18e", i ? " " : "", net->W[l][o * TOR_MAX_WIDTH + i]);
      fprintf(f, "\n");
    }
    fprintf(f, "END\n");
  }
  fprintf(f, "Biases\n");
  for (int l = 0; l < net->n_layers; ++l) {
    fprintf(f, "Layer %d\nBEGIN\n", l + 1);
    for (int o = 0; o < net->width[l + 1]; ++o) fprintf(f, "%.18e\n", net->b[l][o]);
    fprintf(f, "END\n");
  }
  fclose(f);
  return 0;
}

/* compute_vertices, vertex_mode 2, permutation 2: neural_network.cc:135-260 */
void tor_nn_vertices(double theta_m, double theta_p, double max_delta,
                     double xv[5], double yv[5]) {
  const double thr = 1e-8;
  double psi_p = (1 - theta_m) / (1 - theta_p);
  double psi_m = 1 / psi_p;
  double g1 = dmin(max_delta, 1.0 / 8.0);
  double g2 = -dmin(max_delta, 3.0 / 8.0);
  double x1 = 1.0 / 8.0 * (1 + psi_p) - g1 * psi_p;
  double y1 = 1.0 / 8.0 * (1 + psi_m) - g2 * psi_m;
  double x2 = 1.0 / 8.0 * (1 + psi_p) - g2 * psi_p;
  double y2 = 1.0 / 8.0 * (1 + psi_m) - g1 * psi_m;
  const double e = 1.0 / 8.0, t = -3.0 / 8.0, m = -1.0 / 8.0;
#define SETV(a0, a1, a2, a3, a4, b0, b1, b2, b3, b4) \
  do {                                               \
    xv[0] = a0; xv[1] = a1; xv[2] = a2; xv[3] = a3; xv[4] = a4; \
    yv[0] = b0; yv[1] = b1; yv[2] = b2; yv[3] = b3; yv[4] = b4; \
  } while (0)
  if ((theta_m - 1.0) > thr && (theta_p - 1.0) > thr) {
    SETV(e, e, e, e, e, e, e, e, e, e);
  } else if ((theta_m - 1.0) > thr && (theta_p - 1.0) < -1.0 * thr) {
    if (psi_p < -1.0) {
      if (x1 < g2) {
        double nx = ((psi_m * psi_m) + psi_m) / (8 * ((psi_m * psi_m) + 1));
        double ny = 1.0 / 8.0 * (1 + psi_m) - nx * psi_m;
        SETV(nx, nx, nx, nx, nx, ny, ny, ny, ny, ny);
      } else {
        double cx = (2 * g2 + x1) / 3.0;
        double cy = (2 * g1 + y2) / 3.0;
        SETV(g2, x1, g2, cx, cx, g1, g1, y2, cy, cy);
      }
    } else {
      if (y2 < g2)
        SETV(g2, g1, g2, g1, 0.0, g1, g1, g2, g2, 0.0);
      else
        SETV(g2, g1, g2, g1, x2, g1, g1, g2, y2, g2);
    }
  } else if ((theta_m - 1.0) < -1.0 * thr && (theta_p - 1.0) > thr) {
    if (psi_p < -1.0) {
      if (x1 < g2)
        SETV(g2, g1, g2, g1, 0.0, g1, g1, g2, g2, 0.0);
      else
        SETV(g1, g1, g2, x1, g2, g2, g1, g2, g1, y1);
    } else {
      if (y2 < g2) {
        double nx = ((psi_m * psi_m) + psi_m) / (8 * ((psi_m * psi_m) + 1));
        double ny = 1.0 / 8.0 * (1 + psi_m) - nx * psi_m;
        SETV(nx, nx, nx, nx, nx, ny, ny, ny, ny, ny);
      } else {
        double cx = (2 * g1 + x1) / 3.0;
        double cy = (2 * g2 + y2) / 3.0;
        SETV(g1, x1, g1, cx, cx, g2, g2, y2, cy, cy);
      }
    }
  } else if (fabs(theta_m - 1.0) < thr && (theta_p - 1.0) > thr) {
    SETV(e, e, e, e, e, t, e, e, t, m);
  } else if (fabs(theta_p - 1.0) < thr && (theta_m - 1.0) > thr) {
    SETV(e, e, t, t, m, e, e, e, e, e);
  } else {
    SETV(e, e, t, t, m, e, t, t, e, m);
  }
#undef SETV
}

/* scale + features: neural_network.cc:22-38, 300-313 */
static void nn_prep(const double in[4], double st[5], double* tm_out,
                    double* tp_out, double* maxd_out) {
  double c[4];
  for (int i = 0; i < 4; ++i) c[i] = in[i];
  double mv = 1.0;
  for (int i = 0; i < 4; ++i) mv = dmax(mv, fabs(c[i]));
  for (int i = 0; i < 4; ++i) c[i] /= mv;
  double tm = (c[3] - c[2]) / (c[2] - c[1]);
  double tp = (c[1] - c[0]) / (c[2] - c[1]);
  double d1 = fabs(c[1] - c[0]);
  double d2 = fabs(c[2] - c[1]);
  double d3 = fabs(c[3] - c[2]);
  double md = dmax(dmax(d1, d2), d3);
  st[0] = tanh(tm);
  st[1] = tanh(tp);
  st[2] = d1;
  st[3] = d2;
  st[4] = d3;
  *tm_out = tm;
  *tp_out = tp;
  *maxd_out = md;
}

void tor_nn_features(const double u[4], double feat[5], double xv[5],
                     double yv[5]) {
  double tm, tp, md;
  nn_prep(u, feat, &tm, &tp, &md);
  tor_nn_vertices(tm, tp, md, xv, yv);
}

/* forward: neural_network.cc:16-77 (+ relu 265-271, softmax 276-295,
 * compute_recon 320-325).  The state vector has 5 entries; a layer writes its
 * first width[l+1] entries and leaves the rest (the reference's behaviour for
 * widths <= 5). */
void tor_nn_forward(const tor_net* net, const double in[4], double* recl,
                    double* recr) {
  double st[TOR_MAX_WIDTH];
  double tm, tp, md, xv[5], yv[5];
  for (int i = 0; i < TOR_MAX_WIDTH; ++i) st[i] = 0.0;
  nn_prep(in, st, &tm, &tp, &md);
  tor_nn_vertices(tm, tp, md, xv, yv);
  for (int l = 0; l < net->n_layers; ++l) {
    double tmp[TOR_MAX_WIDTH];
    int no = net->width[l + 1], ni = net->width[l];
    for (int j = 0; j < no; ++j) {
      double dot = 0.0;
      for (int k = 0; k < ni; ++k) dot += net->W[l][j * TOR_MAX_WIDTH + k] * st[k];
      tmp[j] = dot;
    }
    for (int j = 0; j < no; ++j) st[j] = tmp[j] + net->b[l][j];
    if (l != net->n_layers - 1) {
      int nr = no > 5 ? no : 5; /* relu over the state vector (>= 5 entries) */
      for (int j = 0; j < nr; ++j) st[j] = dmax(0.0, st[j]);
    }
  }
  double mx = 0.0;
  for (int i = 0; i < 5; ++i) mx = dmax(mx, st[i]);
  double sum = 0;
  for (int i = 0; i < 5; ++i) {
    st[i] = exp(st[i] - mx);
    sum += st[i];
  }
  for (int i = 0; i < 5; ++i) st[i] /= sum;
  double C1 = 0.0, C2 = 0.0;
  for (int i = 0; i < 5; ++i) {
    C1 += st[i] * xv[i];
    C2 += st[i] * yv[i];
  }
  double w1 = 1.0 / 4.0 - 2 * C1;
  double tw0 = 1.0 / 4.0 - 2 * C2;
  double wt[4] = {1.0 - w1, w1, tw0, 1.0 - tw0};
  *recl = wt[0] * 0.5 * (in[1] + in[2]) + wt[1] * (-0.5 * in[0] + 1.5 * in[1]);
  *recr = wt[3] * 0.5 * (in[1] + in[2]) + wt[2] * (-0.5 * in[3] + 1.5 * in[2]);
}

/* minmod (limiter.h:20-35): min(|a|,|b|)*SIGN(a) if a*b > 0, else 0;
 * minabs (limiter.h:40-52): the argument of smaller magnitude, a on ties;
 * limited_slope (reconstruct.cc:162-186) halves either */
static double minmod(double a, double b) {
  if (a * b > 0.0) return dmin(fabs(a), fabs(b)) * ((a < 0) ? -1 : 1);
  return 0.0;
}
static double minabs(double a, double b) { return (fabs(a) > fabs(b)) ? b : a; }
double tor_limited_slope(double a, double b, int use_minabs) {
  return 0.5 * (use_minabs ? minabs(a, b) : minmod(a, b));
}

/* reconstruct dispatch: enflo/src/reconstruct.cc:115-157, 192-239 */
void tor_reconstruct(const tecno_cfg* cfg, const tor_net* net,
                     const double* st, int ng, double* l, double* r) {
  int s = 2 * ng;
  switch (cfg->recon) {
    case TECNO_REC_SP_WENO:
      tor_sp_weno(st + s / 2 - 2, cfg->spweno_corr, l, r);
      break;
    case TECNO_REC_ENO:
      tor_eno_rec(cfg->eno_level, st, ng, l, r);
      break;
    case TECNO_REC_SP_WENO_DL: {
      const double* c = st + s / 2 - 2;
      if (fabs(c[2] - c[1]) < 1e-15) {
        *l = 0.0;
        *r = 0.0;
      } else {
        tor_nn_forward(net, c, l, r);
      }
      break;
    }
    case TECNO_REC_SECOND: {
      /* reconstruct_second_muscl, reconstruct.cc:30-50; KKK is the macro
       * 1.0/3.0, so the weights are (1.0-1.0/3.0) and (1.0+1.0/3.0) */
      double d = st[s / 2] - st[s / 2 - 1];
      double dl = st[s / 2 - 1] - st[s / 2 - 2];
      double dr = st[s / 2 + 1] - st[s / 2];
      *l = st[s / 2 - 1] + (dl * (1.0 - 1.0 / 3.0) + d * (1.0 + 1.0 / 3.0)) * 0.25;
      *r = st[s / 2] - (dr * (1.0 - 1.0 / 3.0) + d * (1.0 + 1.0 / 3.0)) * 0.25;
      break;
    }
    case TECNO_REC_MINMOD:
    case TECNO_REC_MINABS: {
      /* reconstruct_muscl_limited, reconstruct.cc:55-80 */
      int mabs = cfg->recon == TECNO_REC_MINABS;
      double d = st[s / 2] - st[s / 2 - 1];
      double dl = st[s / 2 - 1] - st[s / 2 - 2];
      double dr = st[s / 2 + 1] - st[s / 2];
      *l = st[s / 2 - 1] + tor_limited_slope(dl, d, mabs);
      *r = st[s / 2] - tor_limited_slope(dr, d, mabs);
      break;
    }
    case TECNO_REC_FIRST:
    default:
      *l = st[s / 2 - 1];
      *r = st[s / 2];
      break;
  }
}

/* ------------------------------------------------------------------------ */
/* scalar fluxes: enflo_scalar/src_scalar/{lin,bur}_*.cc, num_flux.cc:12-57  */
typedef struct {
  double h[3];
  double area[3];
} geom_t;

static double sc_dot_axis(const double v[3], int axis) {
  /* Vector::operator* with a unit axis normal (vec.h:239-242) */
  double n[3] = {0.0, 0.0, 0.0};
  n[axis] = 1.0;
  return n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
}

static double lin_ec(const tecno_cfg* c, double ul, double ur, int axis) {
  double vel = sc_dot_axis(c->vel, axis);
  return 0.5 * vel * (ul + ur); /* lin_ec.cc:10-20 */
}
static double bur_ec(double ul, double ur) {
  return (ul * ul + ur * ur + ul * ur) / 6.0; /* bur_ec.cc:10-19 */
}

static double scalar_face_flux(const tecno_cfg* c, const tor_net* net,
                               const double* S, int ng, int axis,
                               const geom_t* g) {
  int s = 2 * ng;
  double ul = S[s / 2 - 1], ur = S[s / 2];
  double flux = 0.0, f0;
  switch (c->flux) {
    case TECNO_FLUX_LIN_EC:
    case TECNO_FLUX_LIN_TECNO:
      flux = lin_ec(c, ul, ur, axis);
      break;
    case TECNO_FLUX_LIN_EC4:
    case TECNO_FLUX_LIN_TECNO4:
      /* lin_ec4.cc:10-34 — note flux0*4.0/3.0 associates as (flux0*4)/3 */
      flux = 0.0;
      f0 = lin_ec(c, S[s / 2 - 1], S[s / 2], axis);
      flux += f0 * 4.0 / 3.0;
      f0 = lin_ec(c, S[s / 2 - 2], S[s / 2], axis);
      flux += f0 * (-1.0 / 6.0);
      f0 = lin_ec(c, S[s / 2 - 1], S[s / 2 + 1], axis);
      flux += f0 * (-1.0 / 6.0);
      break;
    case TECNO_FLUX_BUR_EC:
    case TECNO_FLUX_BUR_TECNO:
      flux = bur_ec(ul, ur);
      break;
    case TECNO_FLUX_BUR_EC4:
    case TECNO_FLUX_BUR_TECNO4:
      /* bur_ec4.cc:10-34 */
      flux = 0.0;
      f0 = bur_ec(S[s / 2 - 1], S[s / 2]);
      flux += f0 * (4.0 / 3.0);
      f0 = bur_ec(S[s / 2 - 2], S[s / 2]);
      flux += f0 * (-1.0 / 6.0);
      f0 = bur_ec(S[s / 2 - 1], S[s / 2 + 1]);
      flux += f0 * (-1.0 / 6.0);
      break;
  }
  int tecno = (c->flux == TECNO_FLUX_LIN_TECNO || c->flux == TECNO_FLUX_LIN_TECNO4 ||
               c->flux == TECNO_FLUX_BUR_TECNO || c->flux == TECNO_FLUX_BUR_TECNO4);
  if (tecno) {
    /* lin_tecno4.cc:10-31, bur_tecno4.cc:10-35 */
    double D;
    if (c->flux == TECNO_FLUX_LIN_TECNO || c->flux == TECNO_FLUX_LIN_TECNO4)
      D = fabs(sc_dot_axis(c->vel, axis));
    else
      D = 0.5 * (fabs(ul) + fabs(ur));
    double uij, uji;
    tor_reconstruct(c, net, S, ng, &uij, &uji);
    flux -= 0.5 * D * (uji - uij);
  }
  return flux * g->area[axis]; /* num_flux.cc:53 */
}

/* ------------------------------------------------------------------------ */
/* Euler fluxes                                                             */
typedef struct { double rho, u[3], p; } prim_t;
typedef struct { double f, v[3], l; } ent_t;

static double vdot(const double a[3], const double b[3]) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static double vsq(const double a[3]) { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }

/* kepec_flux: enflo/src/kepec.cc:10-48 (area included) */
static void kepec_flux(const tecno_cfg* c, const prim_t* Pl, const prim_t* Pr,
                       const double n[3], double area, double F[5]) {
  double R = c->gas_const, G = c->gamma;
  double rhol = Pl->rho, rhor = Pr->rho;
  double rho = tor_logavg(rhol, rhor);
  double templ = Pl->p / (R * rhol);
  double tempr = Pr->p / (R * rhor);
  double vel2 = 0.5 * (vsq(Pl->u) + vsq(Pr->u));
  double betal = 0.5 / (R * templ);
  double betar = 0.5 / (R * tempr);
  double beta = tor_logavg(betal, betar);
  double p = 0.5 * (rhol + rhor) / (betal + betar);
  double vel[3];
  for (int d = 0; d < 3; ++d) vel[d] = Pl->u[d] * 0.5 + Pr->u[d] * 0.5;
  double vn = vdot(vel, n);
  double mass = rho * vn;
  double mom[3];
  for (int d = 0; d < 3; ++d) mom[d] = n[d] * p + vel[d] * mass;
  double en = 0.5 * (1.0 / ((G - 1.0) * beta) - vel2) * mass + vdot(mom, vel);
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

/* roe_ec_flux: enflo/src/roe_ec.cc:10-69 */
static void roe_ec_flux(const tecno_cfg* c, const prim_t* Pl, const prim_t* Pr,
                        const double n[3], double area, double F[5]) {
  double G = c->gamma;
  double rhol = Pl->rho, rhor = Pr->rho, pl = Pl->p, pr = Pr->p;
  double bl = sqrt(rhol / pl);
  double br = sqrt(rhor / pr);
  double z1b = 0.5 * (bl + br);
  double zvb[3];
  for (int d = 0; d < 3; ++d) zvb[d] = Pl->u[d] * (0.5 * bl) + Pr->u[d] * (0.5 * br);
  double z5b = 0.5 * (bl * pl + br * pr);
  double zvn = vdot(zvb, n);
  double z1l = tor_logavg(bl, br);
  double z5l = tor_logavg(bl * pl, br * pr);
  double mass = z5l * zvn;
  double mom[3];
  for (int d = 0; d < 3; ++d) mom[d] = n[d] * (z5b / z1b) + zvb[d] * (mass / z1b);
  double en = 0.5 * ((G + 1.0) / (G - 1.0) * mass / z1l + vdot(mom, zvb)) / z1b;
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

/* kep_flux: enflo/src/kep.cc:10-44, Enthalpy fv.cc:565-569 */
static double enthalpy(const tecno_cfg* c, const prim_t* P) {
  double a = sqrt(c->gamma * P->p / P->rho);
  return a * a / (c->gamma - 1.0) + 0.5 * vsq(P->u);
}
static void kep_flux(const tecno_cfg* c, const prim_t* Pl, const prim_t* Pr,
                     const double n[3], double area, double F[5]) {
  double rho = 0.5 * (Pl->rho + Pr->rho);
  double p = 0.5 * (Pl->p + Pr->p);
  double vel[3];
  for (int d = 0; d < 3; ++d) vel[d] = Pl->u[d] * 0.5 + Pr->u[d] * 0.5;
  double H = 0.5 * (enthalpy(c, Pl) + enthalpy(c, Pr));
  double vn = vdot(vel, n);
  double mass = rho * vn;
  double mom[3];
  for (int d = 0; d < 3; ++d) mom[d] = n[d] * p + vel[d] * mass;
  double en = rho * H * vn;
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

/* keps_flux: enflo/src/keps.cc:10-48 (arithmetic density, mean beta) */
static void keps_flux(const tecno_cfg* c, const prim_t* Pl, const prim_t* Pr,
                      const double n[3], double area, double F[5]) {
  double R = c->gas_const, G = c->gamma;
  double rhol = Pl->rho, rhor = Pr->rho;
  double rho = 0.5 * (rhol + rhor);
  double templ = Pl->p / (R * rhol);
  double tempr = Pr->p / (R * rhor);
  double vel2 = 0.5 * (vsq(Pl->u) + vsq(Pr->u));
  double betal = 0.5 / (R * templ);
  double betar = 0.5 / (R * tempr);
  double beta = 0.5 * (betal + betar);
  double p = 0.5 * rho / beta;
  double vel[3];
  for (int d = 0; d < 3; ++d) vel[d] = Pl->u[d] * 0.5 + Pr->u[d] * 0.5;
  double vn = vdot(vel, n);
  double mass = rho * vn;
  double mom[3];
  for (int d = 0; d < 3; ++d) mom[d] = n[d] * p + vel[d] * mass;
  double en = 0.5 * (1.0 / ((G - 1.0) * beta) - vel2) * mass + vdot(mom, vel);
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

/* roe_flux: enflo/src/roe.cc:10-86, Roe's flux without entropy fix: the
 * upwind cell's physical flux plus the one acoustic wave that crosses */
static void roe_flux(const tecno_cfg* c, const prim_t* Pl, const prim_t* Pr,
                     const double n[3], double area, double F[5]) {
  double G = c->gamma;
  double rhol = Pl->rho, rhor = Pr->rho, pl = Pl->p, pr = Pr->p;
  double hl = G * pl / rhol / (G - 1.0) + 0.5 * vsq(Pl->u);
  double hr = G * pr / rhor / (G - 1.0) + 0.5 * vsq(Pr->u);
  double sl = sqrt(rhol), sr = sqrt(rhor);
  double fl = sl / (sl + sr);
  double fr = 1.0 - fl;
  double rho = sl * sr;
  double vel[3];
  for (int d = 0; d < 3; ++d) vel[d] = Pl->u[d] * fl + Pr->u[d] * fr;
  double h = hl * fl + hr * fr;
  double vn = vdot(vel, n);
  double a = sqrt((G - 1.0) * (h - 0.5 * vsq(vel)));
  double dp = pr - pl;
  double vln = vdot(Pl->u, n), vrn = vdot(Pr->u, n);
  double dV = vrn - vln;
  double mass, mom[3], en;
  if (vn >= 0.0) {
    double coeff = 0.5 * (dp - rho * a * dV) / (a * a);
    double factor = dmin(vn - a, 0.0) * coeff;
    mass = rhol * vln;
    for (int d = 0; d < 3; ++d) mom[d] = n[d] * pl + Pl->u[d] * (rhol * vln);
    en = hl * mass;
    mass += factor;
    for (int d = 0; d < 3; ++d) mom[d] += vel[d] * factor + n[d] * (-a * factor);
    en += (h - a * vn) * factor;
  } else {
    double coeff = 0.5 * (dp + rho * a * dV) / (a * a);
    double factor = dmax(vn + a, 0.0) * coeff;
    mass = rhor * vrn;
    for (int d = 0; d < 3; ++d) mom[d] = n[d] * pr + Pr->u[d] * (rhor * vrn);
    en = hr * mass;
    mass -= factor;
    for (int d = 0; d < 3; ++d) mom[d] += vel[d] * (-factor) + n[d] * (-a * factor);
    en -= (h + a * vn) * factor;
  }
  F[0] = mass * area;
  F[1] = mom[0] * area;
  F[2] = mom[1] * area;
  F[3] = mom[2] * area;
  F[4] = en * area;
}

typedef void (*ec_fn)(const tecno_cfg*, const prim_t*, const prim_t*,
                      const double*, double, double*);

/* 4th-order combination: kepec4.cc:10-38, roe_ec4.cc:10-34, kep4.cc:10-38,
 * keps4.cc:10-38 (Flux::sadd) */
static void ec4_flux(ec_fn fn, const tecno_cfg* c, const prim_t* P, int ng,
                     const double n[3], double area, double F[5]) {
  int s = 2 * ng;
  double f0[5];
  for (int v = 0; v < 5; ++v) F[v] = 0.0;
  fn(c, &P[s / 2 - 1], &P[s / 2], n, area, f0);
  for (int v = 0; v < 5; ++v) F[v] += f0[v] * (4.0 / 3.0);
  fn(c, &P[s / 2 - 2], &P[s / 2], n, area, f0);
  for (int v = 0; v < 5; ++v) F[v] += f0[v] * (-1.0 / 6.0);
  fn(c, &P[s / 2 - 1], &P[s / 2 + 1], n, area, f0);
  for (int v = 0; v < 5; ++v) F[v] += f0[v] * (-1.0 / 6.0);
}

/* TeCNO fluxes: kepes_tecno4_roe.cc:10-176 and its six variants
 * (kepes_tecno_roe.cc, kepes_tecno{,4}_rusanov.cc, roe_tecno{,4}_roe.cc,
 * kep_tecno_roe.cc), which differ only in the entropy-conservative flux and
 * in the eigenvalue matrix Lambda. */
static void euler_face_flux(const tecno_cfg* c, const tor_net* net,
                            const prim_t* P, const ent_t* V, int ng, int axis,
                            const geom_t* g, double F[5]) {
  double n[3] = {0.0, 0.0, 0.0};
  n[axis] = 1.0;
  double area = g->area[axis];
  int s = 2 * ng;
  int fl = c->flux;
  switch (fl) {
    case TECNO_FLUX_KEPEC: kepec_flux(c, &P[s / 2 - 1], &P[s / 2], n, area, F); return;
    case TECNO_FLUX_KEPEC4: ec4_flux(kepec_flux, c, P, ng, n, area, F); return;
    case TECNO_FLUX_ROE_EC: roe_ec_flux(c, &P[s / 2 - 1], &P[s / 2], n, area, F); return;
    case TECNO_FLUX_ROE_EC4: ec4_flux(roe_ec_flux, c, P, ng, n, area, F); return;
    case TECNO_FLUX_KEP: kep_flux(c, &P[s / 2 - 1], &P[s / 2], n, area, F); return;
    case TECNO_FLUX_ROE: roe_flux(c, &P[s / 2 - 1], &P[s / 2], n, area, F); return;
    case TECNO_FLUX_KEP4: ec4_flux(kep_flux, c, P, ng, n, area, F); return;
    case TECNO_FLUX_KEPS: keps_flux(c, &P[s / 2 - 1], &P[s / 2], n, area, F); return;
    case TECNO_FLUX_KEPS4: ec4_flux(keps_flux, c, P, ng, n, area, F); return;
    default: break;
  }
  double R = c->gas_const, G = c->gamma;
  const prim_t* Pl = &P[s / 2 - 1];
  const prim_t* Pr = &P[s / 2];
  double rhol = Pl->rho, rhor = Pr->rho;
  double rho = tor_logavg(rhol, rhor);
  double templ = Pl->p / (R * rhol);
  double tempr = Pr->p / (R * rhor);
  double betal = 0.5 / (R * templ);
  double betar = 0.5 / (R * tempr);
  double beta = tor_logavg(betal, betar);
  double p = 0.5 * (rhol + rhor) / (betal + betar);
  double vel[3];
  for (int d = 0; d < 3; ++d) vel[d] = Pl->u[d] * 0.5 + Pr->u[d] * 0.5;
  double vn = vdot(vel, n);

  switch (fl) {
    case TECNO_FLUX_KEPES_TECNO4_ROE:
    case TECNO_FLUX_KEPES_TECNO4_RUSANOV:
      ec4_flux(kepec_flux, c, P, ng, n, area, F);
      break;
    case TECNO_FLUX_KEPES_TECNO_ROE:
    case TECNO_FLUX_KEPES_TECNO_RUSANOV:
      kepec_flux(c, Pl, Pr, n, area, F);
      break;
    case TECNO_FLUX_ROE_TECNO4_ROE:
      ec4_flux(roe_ec_flux, c, P, ng, n, area, F);
      break;
    case TECNO_FLUX_ROE_TECNO_ROE:
      roe_ec_flux(c, Pl, Pr, n, area, F);
      break;
    case TECNO_FLUX_KEP_TECNO_ROE:
      kep_flux(c, Pl, Pr, n, area, F);
      break;
  }
  double inv_area = 1 / area;
  for (int v = 0; v < 5; ++v) F[v] *= inv_area;

  double a = sqrt(0.5 * G / beta);
  double H = a * a / (G - 1.0) + 0.5 * vsq(vel);
  double v1 = vel[0] * n[1] - vel[1] * n[0];
  double v2 = vel[2] * n[0] - vel[0] * n[2];
  double Rm[5][5] = {
      {1, 1, 0, 0, 1},
      {vel[0] - a * n[0], vel[0], n[1], -n[2], vel[0] + a * n[0]},
      {vel[1] - a * n[1], vel[1], -n[0], 0, vel[1] + a * n[1]},
      {vel[2] - a * n[2], vel[2], 0, n[0], vel[2] + a * n[2]},
      {H - a * vn, 0.5 * vsq(vel), v1, v2, H + a * vn}};
  double const1 = sqrt(rho / G);
  double psqrt = sqrt(p);
  double S[5] = {sqrt(0.5) * const1, sqrt(G - 1.0) * const1, psqrt, psqrt,
                 sqrt(0.5) * const1};
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) Rm[j][i] *= S[i];

  double Lam[5];
  int rusanov = (fl == TECNO_FLUX_KEPES_TECNO_RUSANOV || fl == TECNO_FLUX_KEPES_TECNO4_RUSANOV);
  if (rusanov) {
    double lr = fabs(vn) + fabs(a);
    for (int i = 0; i < 5; ++i) Lam[i] = lr;
  } else {
    double vnl = vdot(Pl->u, n);
    double vnr = vdot(Pr->u, n);
    double al = sqrt(0.5 * G / betal);
    double ar = sqrt(0.5 * G / betar);
    double LL0 = vnl - al, LR0 = vnr - ar, LL3 = vnl, LR3 = vnr;
    double l2 = fabs(vn);
    double eb = 0.0, ea = 0.0;
    if (c->ec1_fix && fl != TECNO_FLUX_KEP_TECNO_ROE) ea = 1.0 / 6.0;
    Lam[0] = (1 + eb) * fabs(vn - a) + ea * fabs(LL0 - LR0);
    Lam[1] = l2;
    Lam[2] = l2;
    Lam[3] = l2;
    Lam[4] = (1 + eb) * fabs(vn + a) + ea * fabs(LL3 - LR3);
  }

  /* scaled entropy variables Z[k] = R~^T V[k] (kepes_tecno4_roe.cc:125-137) */
  double Z[6][5];
  for (int k = 0; k < s; ++k)
    for (int i = 0; i < 5; ++i)
      Z[k][i] = 0.0 + (Rm[0][i] * V[k].f + Rm[1][i] * V[k].v[0] + Rm[2][i] * V[k].v[1] +
                       Rm[3][i] * V[k].v[2] + Rm[4][i] * V[k].l);
  double Zl[5], Zr[5];
  for (int i = 0; i < 5; ++i) {
    double st[6];
    for (int k = 0; k < s; ++k) st[k] = Z[k][i];
    tor_reconstruct(c, net, st, ng, &Zl[i], &Zr[i]);
  }
  double dZ[5], Diff[5];
  for (int i = 0; i < 5; ++i) {
    dZ[i] = Zr[i] - Zl[i];
    Diff[i] = 0;
  }
  for (int i = 0; i < 5; ++i)
    for (int j = 0; j < 5; ++j) Diff[i] += Rm[i][j] * Lam[j] * dZ[j];
  F[0] -= 0.5 * Diff[0];
  F[1] -= 0.5 * Diff[1];
  F[2] -= 0.5 * Diff[2];
  if (c->fix_zmom)
    F[3] -= 0.5 * Diff[3];
  else
    F[2] -= 0.5 * Diff[3]; /* kepes_tecno4_roe.cc:170 */
  F[4] -= 0.5 * Diff[4];
  for (int v = 0; v < 5; ++v) F[v] *= area;
}

/* prim_to_ent: enflo/src/fv.cc:421-431, Entropy 557-560 */
static void prim_to_ent(double G, const prim_t* P, ent_t* V) {
  double beta = P->rho / (2 * P->p);
  double s = log(P->p) - G * log(P->rho);
  V->f = (G - s) / (G - 1) - vsq(P->u) * beta;
  for (int d = 0; d < 3; ++d) V->v[d] = P->u[d] * (2 * beta);
  V->l = -2 * beta;
}

/* ------------------------------------------------------------------------ */
/* batched wrappers                                                          */
static void make_geom(const double h[3], geom_t* g) {
  for (int d = 0; d < 3; ++d) g->h[d] = h[d];
  g->area[0] = h[1] * h[2];
  g->area[1] = h[0] * h[2];
  g->area[2] = h[0] * h[1];
}

void tor_sp_weno_batch(int corr, const double* u4, int64_t n, double* l, double* r) {
  for (int64_t i = 0; i < n; ++i) tor_sp_weno(u4 + 4 * i, corr, l + i, r + i);
}
void tor_reconstruct_batch(const tecno_cfg* cfg, const tor_net* net,
                           const double* st, int ng, int64_t n, double* l, double* r) {
  for (int64_t i = 0; i < n; ++i) tor_reconstruct(cfg, net, st + 2 * ng * i, ng, l + i, r + i);
}
void tor_nn_forward_batch(const tor_net* net, const double* u4, int64_t n,
                          double* l, double* r) {
  for (int64_t i = 0; i < n; ++i) tor_nn_forward(net, u4 + 4 * i, l + i, r + i);
}
void tor_logavg_batch(const double* a, const double* b, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = tor_logavg(a[i], b[i]);
}
void tor_prim_to_ent_batch(const tecno_cfg* cfg, const double* P, int64_t n, double* V) {
  for (int64_t i = 0; i < n; ++i) {
    prim_t p = {P[5 * i], {P[5 * i + 1], P[5 * i + 2], P[5 * i + 3]}, P[5 * i + 4]};
    ent_t v;
    prim_to_ent(cfg->gamma, &p, &v);
    V[5 * i] = v.f;
    V[5 * i + 1] = v.v[0];
    V[5 * i + 2] = v.v[1];
    V[5 * i + 3] = v.v[2];
    V[5 * i + 4] = v.l;
  }
}
void tor_flux_batch(const tecno_cfg* cfg, const tor_net* net, const double* P,
                    const double* V, int ng, int64_t n, int axis,
                    const double h[3], double* F) {
  geom_t g;
  make_geom(h, &g);
  int s = 2 * ng;
  for (int64_t i = 0; i < n; ++i) {
    if (cfg->system == TECNO_SYS_SCALAR) {
      F[i] = scalar_face_flux(cfg, net, P + s * i, ng, axis, &g);
    } else {
      prim_t pp[6];
      ent_t vv[6];
      for (int k = 0; k < s; ++k) {
        const double* q = P + (i * s + k) * 5;
        const double* w = V + (i * s + k) * 5;
        pp[k].rho = q[0];
        pp[k].u[0] = q[1];
        pp[k].u[1] = q[2];
        pp[k].u[2] = q[3];
        pp[k].p = q[4];
        vv[k].f = w[0];
        vv[k].v[0] = w[1];
        vv[k].v[1] = w[2];
        vv[k].v[2] = w[3];
        vv[k].l = w[4];
      }
      euler_face_flux(cfg, net, pp, vv, ng, axis, &g, F + 5 * i);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* whole-grid solver (FV::run and friends)                                   */
struct tor_solver {
  tecno_cfg c;
  tor_net net;
  int has_net;
  int ng;
  int n[3];       /* cells per direction held */
  int active[3];  /* direction swept          */
  int N[3];       /* padded extents            */
  geom_t g;
  double hmin;
  int64_t ncell, npad;
  int nvar;       /* 1 scalar, 5 euler (always 3 velocity comps internally) */
  /* scalar: u (padded) ; euler: prim (padded, AoS prim_t), ent (padded),
   * U, Uold, res (interior, 5 comps) */
  double* u;
  double* uold;
  prim_t* P;
  ent_t* V;
  double* U;
  double* Uo;
  double* res;
  double* res2;   /* classical RK4: accumulated residual (residual2, fv.h:221) */
  double* Fface;  /* face flux scratch, (ncell + extra) x nvar per axis */
  double time;
  int iter;
  int first_stage_prim; /* euler: prim holds IC values (not con_to_prim)   */
};

static int64_t pidx(const tor_solver* s, int i, int j, int k) {
  /* padded index of interior-relative coordinates (may be negative) */
  return ((int64_t)(k + s->ng) * s->N[1] + (j + s->ng)) * s->N[0] + (i + s->ng);
}
static int64_t iidx(const tor_solver* s, int i, int j, int k) {
  return ((int64_t)k * s->n[1] + j) * s->n[0] + i;
}

static int flux_ng(int flux) {
  switch (flux) {
    case TECNO_FLUX_LIN_EC4: case TECNO_FLUX_LIN_TECNO4: case TECNO_FLUX_BUR_EC4:
    case TECNO_FLUX_BUR_TECNO4: case TECNO_FLUX_KEPES_TECNO4_ROE:
    case TECNO_FLUX_KEPES_TECNO4_RUSANOV: case TECNO_FLUX_ROE_TECNO4_ROE:
    case TECNO_FLUX_KEPEC4: case TECNO_FLUX_ROE_EC4: case TECNO_FLUX_KEP4:
    case TECNO_FLUX_KEPS4:
      return 2;
    default:
      return 1;
  }
}
static int rec_ng(const tecno_cfg* c) {
  switch (c->recon) {
    case TECNO_REC_ENO: return c->eno_level;
    case TECNO_REC_FIRST: return 1;
    default: return 2;
  }
}

void tor_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

tor_solver* tor_create(const tecno_cfg* cfg, const tor_net* net, int faithful) {
  tor_solver* s = (tor_solver*)calloc(1, sizeof(tor_solver));
  if (!s) return NULL;
  s->c = *cfg;
  if (net) {
    s->net = *net;
    s->has_net = 1;
  }
  s->ng = flux_ng(cfg->flux);
  if (rec_ng(cfg) > s->ng) s->ng = rec_ng(cfg);
  for (int d = 0; d < 3; ++d) {
    double h = (cfg->hi[d] - cfg->lo[d]) / cfg->n[d];
    s->g.h[d] = h;
    int act = faithful ? 1 : (d < cfg->dims);
    s->active[d] = act;
    s->n[d] = act ? cfg->n[d] : 1;
    s->N[d] = s->n[d] + 2 * s->ng;
  }
  make_geom(s->g.h, &s->g);
  s->hmin = dmin(s->g.h[0], dmin(s->g.h[1], s->g.h[2]));
  s->ncell = (int64_t)s->n[0] * s->n[1] * s->n[2];
  s->npad = (int64_t)s->N[0] * s->N[1] * s->N[2];
  s->nvar = (cfg->system == TECNO_SYS_SCALAR) ? 1 : 5;
  int64_t nf = s->ncell + (int64_t)s->n[1] * s->n[2] + (int64_t)s->n[0] * s->n[2] +
               (int64_t)s->n[0] * s->n[1] + 16;
  s->Fface = (double*)calloc(nf * 3 * s->nvar, sizeof(double));
  if (s->nvar == 1) {
    s->u = (double*)calloc(s->npad, sizeof(double));
    s->uold = (double*)calloc(s->npad, sizeof(double));
  } else {
    s->P = (prim_t*)calloc(s->npad, sizeof(prim_t));
    s->V = (ent_t*)calloc(s->npad, sizeof(ent_t));
    s->U = (double*)calloc(s->ncell * 5, sizeof(double));
    s->Uo = (double*)calloc(s->ncell * 5, sizeof(double));
  }
  s->res = (double*)calloc(s->ncell * s->nvar, sizeof(double));
  s->res2 = (double*)calloc(s->ncell * s->nvar, sizeof(double));
  return s;
}

void tor_destroy(tor_solver* s) {
  if (!s) return;
  free(s->u);
  free(s->uold);
  free(s->P);
  free(s->V);
  free(s->U);
  free(s->Uo);
  free(s->res);
  free(s->res2);
  free(s->Fface);
  free(s);
}

int64_t tor_num_cells(const tor_solver* s) { return s->ncell; }
int tor_num_vars(const tor_solver* s) { return s->nvar; }

/* Parameter::Viscosity parameter.cc:646-663 */
static double viscosity(const tecno_cfg* c, double T) {
  switch (c->mu_model) {
    case TECNO_MU_SUTHERLAND:
      return c->mu_ref * pow(T / c->T_ref, 1.5) * (c->T_ref + c->T_0) / (T + c->T_0);
    case TECNO_MU_POWER: return c->mu_ref * pow(T / c->T_ref, c->omega);
    default: return c->mu_ref;
  }
}
static int is_ns(const tecno_cfg* c) {
  return c->system == TECNO_SYS_EULER &&
         (c->model == TECNO_MODEL_NS_ES || c->model == TECNO_MODEL_NS_KEP);
}

/* ghost fill.  scalar: enflo_scalar fv.cc:110-281; euler: enflo fv.cc:121-361
 * (primitive variables; edges/corners are only needed by the viscous terms and
 * are filled periodically as the reference does). */
static void scalar_bc(tor_solver* s, double* u) {
  int g = s->ng;
  for (int d = 0; d < 3; ++d) {
    int nd = s->n[d];
    int e1 = (d == 0) ? 1 : 0, e2 = (d == 2) ? 1 : 2;
    int n1 = s->n[e1], n2 = s->n[e2];
    for (int side = 0; side < 2; ++side) {
      int bc = s->c.bc[2 * d + side];
      for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n2; ++b)
          for (int l = 0; l < g; ++l) {
            int gi = side ? nd + l : l - g; /* ghost coordinate */
            int src;
            double sign = 1.0;
            if (bc == TECNO_BC_PERIODIC)
              src = side ? l : nd - g + l;
            else if (bc == TECNO_BC_WALL) {
              src = side ? nd - 1 - l : g - 1 - l;
              sign = -1.0;
            } else
              src = side ? nd - 1 : 0;
            int cg[3], cs[3];
            cg[d] = gi;
            cs[d] = src;
            cg[e1] = cs[e1] = a;
            cg[e2] = cs[e2] = b;
            double v = u[pidx(s, cs[0], cs[1], cs[2])];
            u[pidx(s, cg[0], cg[1], cg[2])] = (sign < 0) ? -v : v;
          }
    }
  }
}

static void euler_bc(tor_solver* s) {
  int g = s->ng;
  for (int d = 0; d < 3; ++d) {
    int nd = s->n[d];
    int e1 = (d == 0) ? 1 : 0, e2 = (d == 2) ? 1 : 2;
    int n1 = s->n[e1], n2 = s->n[e2];
    for (int side = 0; side < 2; ++side) {
      int bc = s->c.bc[2 * d + side];
      for (int a = 0; a < n1; ++a)
        for (int b = 0; b < n2; ++b)
          for (int l = 0; l < g; ++l) {
            int gi = side ? nd + l : l - g;
            int src, flip = 0;
            if (bc == TECNO_BC_PERIODIC)
              src = side ? l : nd - g + l;
            else if (bc == TECNO_BC_WALL) {
              src = side ? nd - 1 - l : g - 1 - l;
              flip = 1;
            } else
              src = side ? nd - 1 : 0;
            int cg[3], cs[3];
            cg[d] = gi;
            cs[d] = src;
            cg[e1] = cs[e1] = a;
            cg[e2] = cs[e2] = b;
            prim_t p = s->P[pidx(s, cs[0], cs[1], cs[2])];
            if (flip) p.u[d] = -p.u[d];
            s->P[pidx(s, cg[0], cg[1], cg[2])] = p;
          }
    }
  }
  if (is_ns(&s->c)) {
    /* edges and corners: periodic images (fv.cc:317-359; the viscous node
     * operators read one layer of them).  With neumann or wall boundaries,
     * which the reference does not allow for Navier-Stokes, each out-of-range
     * coordinate is mapped as the faces map it (copy, or mirror with the
     * normal velocity negated), composed over the axes. */
    for (int k = -1; k <= s->n[2]; ++k)
      for (int j = -1; j <= s->n[1]; ++j)
        for (int i = -1; i <= s->n[0]; ++i) {
          int c[3] = {i, j, k};
          int out = 0, flip[3] = {0, 0, 0};
          for (int d = 0; d < 3; ++d) {
            int nd = s->n[d];
            if (c[d] >= 0 && c[d] < nd) continue;
            ++out;
            int bc = s->c.bc[2 * d + (c[d] >= nd)];
            if (bc == TECNO_BC_PERIODIC)
              c[d] = (c[d] + nd) % nd;
            else if (bc == TECNO_BC_WALL) {
              c[d] = c[d] < 0 ? -1 - c[d] : 2 * nd - 1 - c[d];
              flip[d] = 1;
            } else
              c[d] = c[d] < 0 ? 0 : nd - 1;
          }
          if (out < 2) continue;
          prim_t p = s->P[pidx(s, c[0], c[1], c[2])];
          for (int d = 0; d < 3; ++d)
            if (flip[d]) p.u[d] = -p.u[d];
          s->P[pidx(s, i, j, k)] = p;
        }
  }
}

int tor_set_prim(tor_solver* s, const double* prim) {
  int64_t N = s->ncell;
  for (int k = 0; k < s->n[2]; ++k)
    for (int j = 0; j < s->n[1]; ++j)
      for (int i = 0; i < s->n[0]; ++i) {
        int64_t q = iidx(s, i, j, k), p = pidx(s, i, j, k);
        if (s->nvar == 1) {
          s->u[p] = prim[q];
        } else {
          prim_t* P = &s->P[p];
          P->rho = prim[q];
          P->u[0] = prim[N + q];
          P->u[1] = prim[2 * N + q];
          P->u[2] = prim[3 * N + q];
          P->p = prim[4 * N + q];
          /* prim_to_con: fv.cc:405-416 */
          double G = s->c.gamma;
          double* U = s->U + 5 * q;
          U[0] = P->rho;
          for (int d = 0; d < 3; ++d) U[1 + d] = P->u[d] * P->rho;
          U[4] = 0.5 * P->rho * vsq(P->u) + P->p / (G - 1.0);
        }
      }
  if (s->nvar == 1)
    scalar_bc(s, s->u);
  else
    euler_bc(s);
  s->time = 0.0;
  s->iter = 0;
  return 0;
}

int tor_get_prim(const tor_solver* s, double* prim) {
  int64_t N = s->ncell;
  for (int k = 0; k < s->n[2]; ++k)
    for (int j = 0; j < s->n[1]; ++j)
      for (int i = 0; i < s->n[0]; ++i) {
        int64_t q = iidx(s, i, j, k), p = pidx(s, i, j, k);
        if (s->nvar == 1) {
          prim[q] = s->u[p];
        } else {
          const prim_t* P = &s->P[p];
          prim[q] = P->rho;
          prim[N + q] = P->u[0];
          prim[2 * N + q] = P->u[1];
          prim[3 * N + q] = P->u[2];
          prim[4 * N + q] = P->p;
        }
      }
  return 0;
}

/* compute_dt: enflo fv.cc:366-383, enflo_scalar fv.cc:286-311.  The minimum
 * of hmin/speed_i is accumulated in the reference's cell order. */
static double compute_dt(tor_solver* s) {
  double dt = 1.0e20;
  for (int i = 0; i < s->n[0]; ++i)
    for (int j = 0; j < s->n[1]; ++j)
      for (int k = 0; k < s->n[2]; ++k) {
        double speed;
        if (s->nvar == 1) {
          if (s->c.model == TECNO_MODEL_LINADV)
            speed = sqrt(s->c.vel[0] * s->c.vel[0] + s->c.vel[1] * s->c.vel[1] +
                         s->c.vel[2] * s->c.vel[2]);
          else
            speed = fabs(s->u[pidx(s, i, j, k)]);
        } else {
          const prim_t* P = &s->P[pidx(s, i, j, k)];
          speed = sqrt(vsq(P->u)) + sqrt(s->c.gamma * P->p / P->rho);
          if (is_ns(&s->c)) /* fv.cc:375-377 */
            speed += viscosity(&s->c, P->p / (s->c.gas_const * P->rho)) / P->rho /
                     dmin(s->g.h[0], dmin(s->g.h[1], s->g.h[2]));
        }
        dt = dmin(dt, s->hmin / speed);
      }
  return dt * s->c.cfl;
}

/* one face stencil along axis d ending at padded offset; gathers 2ng cells */
static void scalar_residual(tor_solver* s) {
  int g = s->ng;
  int64_t N = s->ncell;
  for (int64_t q = 0; q < N; ++q) s->res[q] = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (!s->active[d]) continue;
    int fn[3] = {s->n[0], s->n[1], s->n[2]};
    fn[d] += 1;
    int64_t nf = (int64_t)fn[0] * fn[1] * fn[2];
    double* F = s->Fface;
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nf; ++f) {
      int i = (int)(f % fn[0]);
      int j = (int)((f / fn[0]) % fn[1]);
      int k = (int)(f / ((int64_t)fn[0] * fn[1]));
      double st[6];
      for (int l = 0; l < 2 * g; ++l) {
        int c[3] = {i, j, k};
        c[d] += l - g; /* face f-1/2 of cell c: cells c-g .. c+g-1 */
        st[l] = s->u[pidx(s, c[0], c[1], c[2])];
      }
      F[f] = scalar_face_flux(&s->c, &s->net, st, g, d, &s->g);
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < s->n[2]; ++k)
      for (int j = 0; j < s->n[1]; ++j)
        for (int i = 0; i < s->n[0]; ++i) {
          int c[3] = {i, j, k};
          int64_t fl = ((int64_t)c[2] * fn[1] + c[1]) * fn[0] + c[0];
          c[d] += 1;
          int64_t fr = ((int64_t)c[2] * fn[1] + c[1]) * fn[0] + c[0];
          double* r = &s->res[iidx(s, i, j, k)];
          *r -= F[fl];
          *r += F[fr];
        }
  }
}

static void euler_prim_to_ent_all(tor_solver* s) {
  double G = s->c.gamma;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < s->npad; ++p) prim_to_ent(G, &s->P[p], &s->V[p]);
}

static void euler_residual(tor_solver* s) {
  int g = s->ng;
  int64_t N = s->ncell;
  for (int64_t q = 0; q < 5 * N; ++q) s->res[q] = 0.0;
  for (int d = 0; d < 3; ++d) {
    if (!s->active[d]) continue;
    int fn[3] = {s->n[0], s->n[1], s->n[2]};
    fn[d] += 1;
    int64_t nf = (int64_t)fn[0] * fn[1] * fn[2];
    double* F = s->Fface;
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nf; ++f) {
      int i = (int)(f % fn[0]);
      int j = (int)((f / fn[0]) % fn[1]);
      int k = (int)(f / ((int64_t)fn[0] * fn[1]));
      prim_t pp[6];
      ent_t vv[6];
      for (int l = 0; l < 2 * g; ++l) {
        int c[3] = {i, j, k};
        c[d] += l - g;
        int64_t p = pidx(s, c[0], c[1], c[2]);
        pp[l] = s->P[p];
        vv[l] = s->V[p];
      }
      euler_face_flux(&s->c, &s->net, pp, vv, g, d, &s->g, F + 5 * f);
    }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < s->n[2]; ++k)
      for (int j = 0; j < s->n[1]; ++j)
        for (int i = 0; i < s->n[0]; ++i) {
          int c[3] = {i, j, k};
          int64_t fl = ((int64_t)c[2] * fn[1] + c[1]) * fn[0] + c[0];
          c[d] += 1;
          int64_t fr = ((int64_t)c[2] * fn[1] + c[1]) * fn[0] + c[0];
          double* r = &s->res[5 * iidx(s, i, j, k)];
          for (int v = 0; v < 5; ++v) {
            r[v] -= F[5 * fl + v];
            r[v] += F[5 * fr + v];
          }
        }
  }
}

/* ---- Navier-Stokes viscous terms (periodic grids) --------------------- */
typedef struct { double m[5]; } flux5; /* mass, mom x/y/z, energy */

/* the ES viscous node fluxes from the node average and differences of V
 * (vflux_ES.cc:74-145), before the face areas: X, Y, Z = momentum x, y, z
 * and energy of the x, y and z fluxes (m[0], the mass flux, is 0).  The
 * reference adds the K21 V_x part of Y and the K31 V_x, K32 V_y parts of Z
 * and then overwrites them with the K22 / K33 assignments (vflux_ES.cc:
 * 103-145), so y = K22 V_y + K23 V_z and z = K33 V_z; cfg->fix_visc keeps
 * them (y = K21 V_x + K22 V_y + K23 V_z, z = K31 V_x + K32 V_y + K33 V_z). */
static void es_flux_core(const tecno_cfg* c, const double avg[5], const double Tx[5],
                         const double Ty[5], const double Tz[5], flux5* X, flux5* Y, flux5* Z) {
  const double T_avg = -1 / avg[4] / c->gas_const;
  const double mu = viscosity(c, T_avg);
  const double Cp = c->gamma * c->gas_const / (c->gamma - 1.0); /* parameter.cc:286 */
  const double kappa = mu * Cp / c->prandtl;
  const double c1 = 4.0 * mu / 3.0, c2 = -2.0 * mu / 3.0, c3 = kappa / c->gas_const,
               e4 = 1.0 / avg[4], e1 = avg[1] * e4, e2 = avg[2] * e4, e3 = avg[3] * e4;
  /* X: K11 V_x + K12 V_y + K13 V_z */
  X->m[0] = 0.0;
  X->m[1] = c1 * e4 * (-Tx[1] + e1 * Tx[4]);
  X->m[2] = mu * e4 * (-Tx[2] + e2 * Tx[4]);
  X->m[3] = mu * e4 * (-Tx[3] + e3 * Tx[4]);
  X->m[4] = e4 * (c1 * e1 * Tx[1] + mu * e2 * Tx[2] + mu * e3 * Tx[3] +
                  (-c1 * e1 * e1 - mu * (e2 * e2 + e3 * e3) + c3 * e4) * Tx[4]);
  X->m[1] += c2 * e4 * (-Ty[2] + e2 * Ty[4]);
  X->m[2] += mu * e4 * (-Ty[1] + e1 * Ty[4]);
  X->m[3] += 0;
  X->m[4] += e4 * (mu * e2 * Ty[1] + c2 * e1 * Ty[2] - (mu + c2) * e1 * e2 * Ty[4]);
  X->m[1] += c2 * e4 * (-Tz[3] + e3 * Tz[4]);
  X->m[2] += 0;
  X->m[3] += mu * e4 * (-Tz[1] + e1 * Tz[4]);
  X->m[4] += e4 * (mu * e3 * Tz[1] + c2 * e1 * Tz[3] - (mu + c2) * e1 * e3 * Tz[4]);
  /* Y: K22 V_y (assigned), K23 V_z */
  Y->m[0] = 0.0;
  Y->m[1] = mu * e4 * (-Ty[1] + e1 * Ty[4]);
  Y->m[2] = c1 * e4 * (-Ty[2] + e2 * Ty[4]);
  Y->m[3] = mu * e4 * (-Ty[3] + e3 * Ty[4]);
  Y->m[4] = e4 * (mu * e1 * Ty[1] + c1 * e2 * Ty[2] + mu * e3 * Ty[3] +
                  (-c1 * e2 * e2 - mu * (e1 * e1 + e3 * e3) + c3 * e4) * Ty[4]);
  Y->m[1] += 0;
  Y->m[2] += c2 * e4 * (-Tz[3] + e3 * Tz[4]);
  Y->m[3] += mu * e4 * (-Tz[2] + e2 * Tz[4]);
  Y->m[4] += e4 * (mu * e3 * Tz[2] + c2 * e2 * Tz[3] - (mu + c2) * e2 * e3 * Tz[4]);
  if (c->fix_visc) { /* K21 V_x */
    Y->m[1] += mu * e4 * (-Tx[2] + e2 * Tx[4]);
    Y->m[2] += c2 * e4 * (-Tx[1] + e1 * Tx[4]);
    Y->m[4] += e4 * (c2 * e2 * Tx[1] + mu * e1 * Tx[2] - (mu + c2) * e1 * e2 * Tx[4]);
  }
  /* Z: K33 V_z (assigned) */
  Z->m[0] = 0.0;
  Z->m[1] = mu * e4 * (-Tz[1] + e1 * Tz[4]);
  Z->m[2] = mu * e4 * (-Tz[2] + e2 * Tz[4]);
  Z->m[3] = c1 * e4 * (-Tz[3] + e3 * Tz[4]);
  Z->m[4] = e4 * (mu * e1 * Tz[1] + mu * e2 * Tz[2] + c1 * e3 * Tz[3] +
                  (-c1 * e3 * e3 - mu * (e1 * e1 + e2 * e2) + c3 * e4) * Tz[4]);
  if (c->fix_visc) { /* K31 V_x, K32 V_y */
    Z->m[1] += mu * e4 * (-Tx[3] + e3 * Tx[4]);
    Z->m[3] += c2 * e4 * (-Tx[1] + e1 * Tx[4]);
    Z->m[4] += e4 * (c2 * e3 * Tx[1] + mu * e1 * Tx[3] - (mu + c2) * e1 * e3 * Tx[4]);
    Z->m[2] += mu * e4 * (-Ty[3] + e3 * Ty[4]);
    Z->m[3] += c2 * e4 * (-Ty[2] + e2 * Ty[4]);
    Z->m[4] += e4 * (c2 * e3 * Ty[2] + mu * e2 * Ty[3] - (mu + c2) * e2 * e3 * Ty[4]);
  }
}

void tor_es_node_flux(const tecno_cfg* cfg, const double avg[5], const double Tx[5],
                      const double Ty[5], const double Tz[5], double out[12]) {
  flux5 X, Y, Z;
  es_flux_core(cfg, avg, Tx, Ty, Tz, &X, &Y, &Z);
  for (int v = 0; v < 4; ++v) {
    out[v] = X.m[1 + v];
    out[4 + v] = Y.m[1 + v];
    out[8 + v] = Z.m[1 + v];
  }
}

/* nodal_vlux_ES vflux_ES.cc:15-151: node average and differences of the
 * entropy variables of the 8 cells, es_flux_core, face areas */
static void node_flux_es(const tor_solver* s, int i, int j, int k, flux5* X, flux5* Y, flux5* Z) {
  const tecno_cfg* c = &s->c;
  const double dx = s->g.h[0], dy = s->g.h[1], dz = s->g.h[2];
  /* cells: index (a,b,c) in {0,1}^3 -> V[pidx(i+a, j+b, k+c)] */
#define VV(a, b, cc) (&s->V[pidx(s, i + (a), j + (b), k + (cc))])
  const ent_t* E[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int q = 0; q < 2; ++q) E[a][b][q] = VV(a, b, q);
#undef VV
  double avg[5], Tx[5], Ty[5], Tz[5];
  const double f1 = 1.0 / 8.0, fx = 1.0 / 4.0 / dx, fy = 1.0 / 4.0 / dy, fz = 1.0 / 4.0 / dz;
  for (int comp = 0; comp < 5; ++comp) {
#define C(e) ((comp == 0) ? (e)->f : (comp == 4) ? (e)->l : (e)->v[comp - 1])
    /* ent_avg: (111)(011)(101)(110) then (001)(010)(100)(000) */
    avg[comp] = C(E[1][1][1]) * f1 + C(E[0][1][1]) * f1 + C(E[1][0][1]) * f1 + C(E[1][1][0]) * f1;
    avg[comp] += C(E[0][0][1]) * f1 + C(E[0][1][0]) * f1 + C(E[1][0][0]) * f1 + C(E[0][0][0]) * f1;
    Tx[comp] = C(E[1][1][1]) * fx + C(E[1][1][0]) * fx + C(E[1][0][1]) * fx + C(E[1][0][0]) * fx;
    Tx[comp] += C(E[0][1][1]) * -fx + C(E[0][1][0]) * -fx + C(E[0][0][1]) * -fx + C(E[0][0][0]) * -fx;
    Ty[comp] = C(E[1][1][1]) * fy + C(E[1][1][0]) * fy + C(E[0][1][1]) * fy + C(E[0][1][0]) * fy;
    Ty[comp] += C(E[1][0][1]) * -fy + C(E[1][0][0]) * -fy + C(E[0][0][1]) * -fy + C(E[0][0][0]) * -fy;
    Tz[comp] = C(E[1][1][1]) * fz + C(E[1][0][1]) * fz + C(E[0][1][1]) * fz + C(E[0][0][1]) * fz;
    Tz[comp] += C(E[1][1][0]) * -fz + C(E[1][0][0]) * -fz + C(E[0][1][0]) * -fz + C(E[0][0][0]) * -fz;
#undef C
  }
  es_flux_core(c, avg, Tx, Ty, Tz, X, Y, Z);
  const double ax = -(dy * dz), ay = -(dx * dz), az = -(dx * dy);
  for (int v = 0; v < 5; ++v) {
    X->m[v] *= ax;
    Y->m[v] *= ay;
    Z->m[v] *= az;
  }
}

/* nodal_vlux_KE vflux_KE.cc:15-96 with nodal_avg_and_vel_derivatives
 * nodal_avg_and_der.cc:16-82 */
static void node_flux_ke(const tor_solver* s, int i, int j, int k, flux5* X, flux5* Y, flux5* Z) {
  const tecno_cfg* c = &s->c;
  const double dx = s->g.h[0], dy = s->g.h[1], dz = s->g.h[2], R = c->gas_const;
  const prim_t* P[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int q = 0; q < 2; ++q) P[a][b][q] = &s->P[pidx(s, i + a, j + b, k + q)];
  const double f1 = 1.0 / 8.0, fx = 1.0 / (4.0 * dx), fy = 1.0 / (4.0 * dy), fz = 1.0 / (4.0 * dz);
  double avg[5];
  for (int comp = 0; comp < 5; ++comp) {
#define C(q_) ((comp == 0) ? (q_)->rho : (comp == 4) ? (q_)->p : (q_)->u[comp - 1])
    avg[comp] = C(P[1][1][1]) * f1 + C(P[0][1][1]) * f1 + C(P[1][0][1]) * f1 + C(P[1][1][0]) * f1;
    avg[comp] += C(P[0][0][1]) * f1 + C(P[0][1][0]) * f1 + C(P[1][0][0]) * f1 + C(P[0][0][0]) * f1;
#undef C
  }
  double Tx[3], Ty[3], Tz[3];
  for (int d = 0; d < 3; ++d) {
    Tx[d] = P[1][1][1]->u[d] * fx + P[1][1][0]->u[d] * fx + P[1][0][1]->u[d] * fx + P[1][0][0]->u[d] * fx;
    Tx[d] += P[0][1][1]->u[d] * -fx + P[0][1][0]->u[d] * -fx + P[0][0][1]->u[d] * -fx + P[0][0][0]->u[d] * -fx;
    Ty[d] = P[1][1][1]->u[d] * fy + P[1][1][0]->u[d] * fy + P[0][1][1]->u[d] * fy + P[0][1][0]->u[d] * fy;
    Ty[d] += P[1][0][1]->u[d] * -fy + P[1][0][0]->u[d] * -fy + P[0][0][1]->u[d] * -fy + P[0][0][0]->u[d] * -fy;
    Tz[d] = P[1][1][1]->u[d] * fz + P[1][0][1]->u[d] * fz + P[0][1][1]->u[d] * fz + P[0][0][1]->u[d] * fz;
    Tz[d] += P[1][1][0]->u[d] * -fz + P[1][0][0]->u[d] * -fz + P[0][1][0]->u[d] * -fz + P[0][0][0]->u[d] * -fz;
  }
#define TT(c_) ((c_)->p / (R * (c_)->rho))
  const double T_avg = avg[4] / (R * avg[0]);
  const double mu = viscosity(c, T_avg);
  const double Cp = c->gamma * R / (c->gamma - 1.0);
  const double kappa = mu * Cp / c->prandtl;
  const double Txt = fx * (TT(P[1][1][1]) + TT(P[1][1][0]) + TT(P[1][0][1]) + TT(P[1][0][0]) -
                           TT(P[0][1][1]) - TT(P[0][1][0]) - TT(P[0][0][1]) - TT(P[0][0][0]));
  const double Tyt = fy * (TT(P[1][1][1]) + TT(P[1][1][0]) + TT(P[0][1][1]) + TT(P[0][1][0]) -
                           TT(P[1][0][1]) - TT(P[1][0][0]) - TT(P[0][0][1]) - TT(P[0][0][0]));
  const double Tzt = fz * (TT(P[1][1][1]) + TT(P[1][0][1]) + TT(P[0][1][1]) + TT(P[0][0][1]) -
                           TT(P[1][1][0]) - TT(P[1][0][0]) - TT(P[0][1][0]) - TT(P[0][0][0]));
#undef TT
  const double gv = Tx[0] + Ty[1] + Tz[2];
  const double txx = mu * (2 * Tx[0] - (2.0 / 3.0) * gv), tyy = mu * (2 * Ty[1] - (2.0 / 3.0) * gv),
               tzz = mu * (2 * Tz[2] - (2.0 / 3.0) * gv), txy = mu * (Ty[0] + Tx[1]),
               txz = mu * (Tx[2] + Tz[0]), tyz = mu * (Ty[2] + Tz[1]);
  const double* u = &avg[1];
  X->m[0] = 0.0; X->m[1] = txx; X->m[2] = txy; X->m[3] = txz;
  X->m[4] = u[0] * txx + u[1] * txy + u[2] * txz + kappa * Txt;
  Y->m[0] = 0.0; Y->m[1] = txy; Y->m[2] = tyy; Y->m[3] = tyz;
  Y->m[4] = u[0] * txy + u[1] * tyy + u[2] * tyz + kappa * Tyt;
  Z->m[0] = 0.0; Z->m[1] = txz; Z->m[2] = tyz; Z->m[3] = tzz;
  Z->m[4] = u[0] * txz + u[1] * tyz + u[2] * tzz + kappa * Tzt;
  const double ax = -(dy * dz), ay = -(dx * dz), az = -(dx * dy);
  for (int v = 0; v < 5; ++v) {
    X->m[v] *= ax;
    Y->m[v] *= ay;
    Z->m[v] *= az;
  }
}

/* compute_visc_residual fv.cc:786-1157 on a periodic grid: every node
 * (i+1/2, j+1/2, k+1/2), i = -1 .. n-1, scatters its three fluxes to the
 * (up to 8) interior cells around it with weights (-1)^l/4 per direction
 * (Flux::sadd of 3 fluxes).  The reference visits interior, face and corner
 * nodes once, but every edge node (two coordinates on the boundary) twice:
 * once in each of the two boundary groups that share the edge (e.g. node
 * (Nx-1, j, -1) in the "X-Right" and the "Z-Left" loops), so edge nodes add
 * their fluxes twice. */
static void visc_residual(tor_solver* s) {
  const int es = s->c.model == TECNO_MODEL_NS_ES;
  for (int k = -1; k < s->n[2]; ++k)
    for (int j = -1; j < s->n[1]; ++j)
      for (int i = -1; i < s->n[0]; ++i) {
        flux5 X, Y, Z;
        if (es)
          node_flux_es(s, i, j, k, &X, &Y, &Z);
        else
          node_flux_ke(s, i, j, k, &X, &Y, &Z);
        /* a node on an edge of two periodic boundaries is added twice by the
         * reference's boundary loops (fv.cc:821-1156; fix_edges: once);
         * nodes on neumann / wall boundaries once */
        const int ox = i == -1 || i == s->n[0] - 1, oy = j == -1 || j == s->n[1] - 1,
                  oz = k == -1 || k == s->n[2] - 1;
        const int nb = ox + oy + oz;
        const int np = (ox && s->c.bc[0] == TECNO_BC_PERIODIC) +
                       (oy && s->c.bc[2] == TECNO_BC_PERIODIC) +
                       (oz && s->c.bc[4] == TECNO_BC_PERIODIC);
        const int twice = nb == 2 && np == 2 && !s->c.fix_edges;
        for (int rep = 0; rep < (twice ? 2 : 1); ++rep)
        for (int l = 0; l < 2; ++l)
          for (int m = 0; m < 2; ++m)
            for (int q = 0; q < 2; ++q) {
              int ci = i + l, cj = j + m, ck = k + q;
              if (ci < 0 || cj < 0 || ck < 0 || ci >= s->n[0] || cj >= s->n[1] || ck >= s->n[2])
                continue;
              double a = (l ? -1.0 : 1.0) / 4.0, b = (m ? -1.0 : 1.0) / 4.0, cc = (q ? -1.0 : 1.0) / 4.0;
              double* r = &s->res[5 * iidx(s, ci, cj, ck)];
              for (int v = 0; v < 5; ++v) r[v] += X.m[v] * a + Y.m[v] * b + Z.m[v] * cc;
            }
      }
}

static const double ARK[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0}; /* ext_constants.h:7 */
static const double BRK[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0}; /* ext_constants.h:8 */
static const double JRK[4] = {1.0 / 4.0, 1.0 / 3.0, 1.0 / 2.0, 1.0}; /* ext_constants.h:9 */
static const double RK4C[4] = {1.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};      /* ext_constants.h:10 */

int tor_step(tor_solver* s, double t_stamp, double* dt_out) {
  double dt = compute_dt(s);
  if (s->time + dt > t_stamp) dt = t_stamp - s->time; /* fv.cc:1266 */
  double factor = dt / (s->g.h[0] * s->g.h[1] * s->g.h[2]);
  int ts = s->c.time_scheme;
  /* n_rks: parameter.cc:111-132 */
  int nrk = ts == TECNO_TS_RK1 ? 1 : (ts == TECNO_TS_SSPRK3 ? 3 : 4);
  int64_t N = s->ncell;
  if (ts == TECNO_TS_RK4) /* residual2 zeroed per step, fv.cc:1267-1274 */
    memset(s->res2, 0, (size_t)s->nvar * N * sizeof(double));
  if (s->nvar == 1) {
    memcpy(s->uold, s->u, s->npad * sizeof(double));
    for (int rk = 0; rk < nrk; ++rk) {
      scalar_residual(s);
      /* update_solution: enflo_scalar fv.cc:459-499 */
      for (int k = 0; k < s->n[2]; ++k)
        for (int j = 0; j < s->n[1]; ++j)
          for (int i = 0; i < s->n[0]; ++i) {
            int64_t p = pidx(s, i, j, k);
            int64_t q = iidx(s, i, j, k);
            if (ts == TECNO_TS_JAMESON_RK4) {
              s->u[p] = (s->uold[p] - factor * JRK[rk] * s->res[q]);
            } else if (ts == TECNO_TS_RK4) {
              s->res2[q] += (s->res[q] / RK4C[rk]);
              if (rk < 3)
                s->u[p] = (s->uold[p] - factor * RK4C[rk + 1] * s->res[q]);
              else
                s->u[p] = (s->uold[p] - factor * s->res2[q] / 6.0);
            } else {
              s->u[p] *= BRK[rk];
              s->u[p] += (s->uold[p] * ARK[rk] - factor * BRK[rk] * s->res[q]);
            }
          }
      scalar_bc(s, s->u);
    }
  } else {
    memcpy(s->Uo, s->U, 5 * N * sizeof(double));
    double G = s->c.gamma;
    for (int rk = 0; rk < nrk; ++rk) {
      euler_prim_to_ent_all(s);
      euler_residual(s);
      if (is_ns(&s->c)) visc_residual(s); /* compute_residual fv.cc:591-603 */
      /* update_solution: enflo fv.cc:1160-1198 (ConVar::sadd convar.h:118,
       * ConVar::equ convar.h:103, Flux::sadd flux.h:127) */
      if (ts == TECNO_TS_JAMESON_RK4) {
        double c2 = -factor * JRK[rk];
        for (int64_t q = 0; q < 5 * N; ++q) s->U[q] = s->Uo[q] * 1.0 + s->res[q] * c2;
      } else if (ts == TECNO_TS_RK4) {
        double w = 1.0 / RK4C[rk];
        for (int64_t q = 0; q < 5 * N; ++q) {
          s->res2[q] += s->res[q] * w;
          if (rk < 3)
            s->U[q] = s->Uo[q] * 1.0 + s->res[q] * (-factor * RK4C[rk + 1]);
          else
            s->U[q] = s->Uo[q] * 1.0 + s->res2[q] * (-factor / 6.0);
        }
      } else {
        double c2 = -factor * BRK[rk];
        for (int64_t q = 0; q < N; ++q)
          for (int v = 0; v < 5; ++v) {
            s->U[5 * q + v] *= BRK[rk];
            s->U[5 * q + v] += s->Uo[5 * q + v] * ARK[rk] + s->res[5 * q + v] * c2;
          }
      }
      /* con_to_prim: fv.cc:388-400 */
      for (int k = 0; k < s->n[2]; ++k)
        for (int j = 0; j < s->n[1]; ++j)
          for (int i = 0; i < s->n[0]; ++i) {
            const double* U = s->U + 5 * iidx(s, i, j, k);
            prim_t* P = &s->P[pidx(s, i, j, k)];
            P->rho = U[0];
            double inv = 1.0 / U[0];
            for (int d = 0; d < 3; ++d) P->u[d] = U[1 + d] * inv;
            double m2 = U[1] * U[1] + U[2] * U[2] + U[3] * U[3];
            P->p = (G - 1.0) * (U[4] - 0.5 * m2 / U[0]);
          }
      euler_bc(s);
    }
  }
  s->time += dt;
  s->iter++;
  if (dt_out) *dt_out = dt;
  return 0;
}

int tor_advance(tor_solver* s, double final_time, int t_stamps, int max_iter,
                int* iters_out, double* time_out) {
  double t_int = final_time / t_stamps;
  double t_save = t_int;
  /* stamps already passed (when called on a running solver) */
  while (s->time >= t_save && t_save < final_time) t_save += t_int;
  while (s->time < final_time && s->iter < max_iter) {
    double dt;
    tor_step(s, t_save, &dt);
    if (s->time == t_save) t_save += t_int;
  }
  if (iters_out) *iters_out = s->iter;
  if (time_out) *time_out = s->time;
  return 0;
}

/* compute_global compute_global.cc:18-27,72-73 (scalar
 * enflo_scalar/src_scalar/compute_global.cc:12-20): sums over the
 * reference's cells; a collapsed axis (not swept outside faithful mode)
 * holds one cell for the cfg->n[d] the reference stores */
void tor_audit(const tor_solver* s, double out[8]) {
  double dV = s->g.h[0] * s->g.h[1] * s->g.h[2];
  for (int d = 0; d < 3; ++d)
    if (!s->active[d]) dV *= s->c.n[d];
  for (int v = 0; v < 8; ++v) out[v] = 0.0;
  double ent = 0.0, ke = 0.0, m = 0.0, mo[3] = {0, 0, 0}, e = 0.0, smax = 0.0;
  for (int i = 0; i < s->n[0]; ++i)
    for (int j = 0; j < s->n[1]; ++j)
      for (int k = 0; k < s->n[2]; ++k) {
        int64_t p = pidx(s, i, j, k);
        if (s->nvar == 1) {
          double u = s->u[p];
          ent += 0.5 * pow(u, 2);
          m += u;
          double sp = (s->c.model == TECNO_MODEL_LINADV)
                          ? sqrt(s->c.vel[0] * s->c.vel[0] + s->c.vel[1] * s->c.vel[1] +
                                 s->c.vel[2] * s->c.vel[2])
                          : fabs(u);
          smax = dmax(smax, sp);
        } else {
          const prim_t* P = &s->P[p];
          double G = s->c.gamma;
          double sv = log(P->p) - G * log(P->rho);
          ent += -P->rho * sv / (G - 1.0);
          ke += P->rho * vsq(P->u);
          const double* U = s->U + 5 * iidx(s, i, j, k);
          m += U[0];
          for (int d = 0; d < 3; ++d) mo[d] += U[1 + d];
          e += U[4];
          smax = dmax(smax, sqrt(vsq(P->u)) + sqrt(G * P->p / P->rho));
        }
      }
  out[0] = ent * dV;
  out[1] = 0.5 * ke * dV;
  out[2] = m * dV;
  out[3] = mo[0] * dV;
  out[4] = mo[1] * dV;
  out[5] = mo[2] * dV;
  out[6] = e * dV;
  out[7] = smax;
}

/* ------------------------------------------------------------------------ */
/* trainer semantics (DSP-WENO_network_trainer.py)                          */
/* Rows: features/vertices as compute_vertices + the forward's inputs
 * (neural_network.cc:16-77, 135-260), i.e. what the trainer reads from its
 * data file columns (trainer:281-333).                                      */
void tor_train_rows(const double* u4, const double* target, int64_t n,
                    double* rows) {
  for (int64_t i = 0; i < n; ++i) {
    double* r = rows + i * TOR_ROW;
    tor_nn_features(u4 + 4 * i, r, r + 5, r + 10);
    for (int k = 0; k < 4; ++k) r[15 + k] = u4[4 * i + k];
    r[19] = target[2 * i];
    r[20] = target[2 * i + 1];
  }
}

static double bf16_round(double x) {
  float f = (float)x;
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  memcpy(&f, &u, 4);
  return (double)f;
}

/* MLP.forward (trainer:63-91): Linear layers with ReLU between them, softmax,
 * C1 = s.xv, C2 = s.yv, omega1 = 1/4 - 2 C1, omega0~ = 1/4 - 2 C2;
 * inter_evaluation (trainer:94-106): left = (1-omega1)(u1+u2)/2 +
 * omega1 (3u1-u0)/2, right = omega0~ (3u2-u3)/2 + (1-omega0~)(u1+u2)/2;
 * loss (trainer:355-356) = (MSE(left,tL) + MSE(right,tR)) / 2.  The gradient
 * is the analytic backward of exactly these expressions. */
double tor_train_loss_grad(const tor_net* net, const double* rows, int64_t n,
                           int bf16, tor_net* grad) {
  const int L = net->n_layers;
  tor_net Wq = *net;
  if (bf16)
    for (int l = 0; l < L; ++l) {
      for (int i = 0; i < TOR_MAX_WIDTH * TOR_MAX_WIDTH; ++i)
        Wq.W[l][i] = bf16_round(net->W[l][i]);
      for (int i = 0; i < TOR_MAX_WIDTH; ++i) Wq.b[l][i] = bf16_round(net->b[l][i]);
    }
  if (grad) {
    memset(grad, 0, sizeof(*grad));
    grad->n_layers = L;
    for (int l = 0; l <= L; ++l) grad->width[l] = net->width[l];
  }
  double loss = 0.0;
  for (int64_t r = 0; r < n; ++r) {
    const double* row = rows + r * TOR_ROW;
    double h[TOR_MAX_LAYERS + 1][TOR_MAX_WIDTH];
    for (int k = 0; k < 5; ++k) h[0][k] = bf16 ? bf16_round(row[k]) : row[k];
    for (int l = 0; l < L; ++l) {
      const int no = net->width[l + 1], ni = net->width[l];
      for (int j = 0; j < no; ++j) {
        double z = Wq.b[l][j];
        for (int k = 0; k < ni; ++k) z += Wq.W[l][j * TOR_MAX_WIDTH + k] * h[l][k];
        if (l != L - 1) z = dmax(0.0, z);
        h[l + 1][j] = (bf16 && l != L - 1) ? bf16_round(z) : z;
      }
    }
    const double* lg = h[L];
    double mx = lg[0];
    for (int i = 1; i < 5; ++i) mx = dmax(mx, lg[i]);
    double s[5], sum = 0.0;
    for (int i = 0; i < 5; ++i) {
      s[i] = exp(lg[i] - mx);
      sum += s[i];
    }
    double C1 = 0.0, C2 = 0.0;
    for (int i = 0; i < 5; ++i) {
      s[i] /= sum;
      C1 += s[i] * row[5 + i];
      C2 += s[i] * row[10 + i];
    }
    const double u0 = row[15], u1 = row[16], u2 = row[17], u3 = row[18];
    const double w1 = 0.25 - 2.0 * C1, tw0 = 0.25 - 2.0 * C2;
    const double left = (1.0 - w1) * 0.5 * (u1 + u2) + w1 * (1.5 * u1 - 0.5 * u0);
    const double right = tw0 * (1.5 * u2 - 0.5 * u3) + (1.0 - tw0) * 0.5 * (u1 + u2);
    const double eL = left - row[19], eR = right - row[20];
    loss += 0.5 * (eL * eL + eR * eR);
    if (!grad) continue;
    /* d loss / d left = eL / n; d left / d C1 = -2 (1.5u1 - 0.5u0 - 0.5(u1+u2)) */
    const double gC1 = (eL / (double)n) * (-2.0) * (u1 - 0.5 * u0 - 0.5 * u2);
    const double gC2 = (eR / (double)n) * (-2.0) * (u2 - 0.5 * u3 - 0.5 * u1);
    double gz[TOR_MAX_WIDTH] = {0};
    for (int i = 0; i < 5; ++i)
      gz[i] = s[i] * ((row[5 + i] - C1) * gC1 + (row[10 + i] - C2) * gC2);
    for (int l = L - 1; l >= 0; --l) {
      const int no = net->width[l + 1], ni = net->width[l];
      double gh[TOR_MAX_WIDTH] = {0};
      for (int j = 0; j < no; ++j) {
        grad->b[l][j] += gz[j];
        for (int k = 0; k < ni; ++k) {
          grad->W[l][j * TOR_MAX_WIDTH + k] += gz[j] * h[l][k];
          gh[k] += Wq.W[l][j * TOR_MAX_WIDTH + k] * gz[j];
        }
      }
      if (l > 0)
        for (int k = 0; k < ni; ++k) gz[k] = (h[l][k] > 0.0) ? gh[k] : 0.0;
    }
  }
  return loss / (double)n;
}

/* torch.optim.Adam (weight_decay as L2 on the gradient, no amsgrad):
 * g += wd p; m = b1 m + (1-b1) g; v = b2 v + (1-b2) g^2;
 * p -= (lr / (1-b1^t)) m / (sqrt(v) / sqrt(1-b2^t) + eps)                  */
static void adam_one(double* p, double g, double* m, double* v, double bc1,
                     double bc2, double lr, double b1, double b2, double eps,
                     double wd) {
  g += wd * (*p);
  *m = b1 * (*m) + (1.0 - b1) * g;
  *v = b2 * (*v) + (1.0 - b2) * g * g;
  *p -= (lr / bc1) * (*m) / (sqrt(*v) / sqrt(bc2) + eps);
}

void tor_adam_step(tor_net* net, const tor_net* grad, tor_net* m, tor_net* v,
                   int64_t t, double lr, double beta1, double beta2, double eps,
                   double weight_decay) {
  const double bc1 = 1.0 - pow(beta1, (double)t), bc2 = 1.0 - pow(beta2, (double)t);
  for (int l = 0; l < net->n_layers; ++l) {
    const int no = net->width[l + 1], ni = net->width[l];
    for (int j = 0; j < no; ++j) {
      for (int k = 0; k < ni; ++k) {
        const int q = j * TOR_MAX_WIDTH + k;
        adam_one(&net->W[l][q], grad->W[l][q], &m->W[l][q], &v->W[l][q], bc1, bc2,
                 lr, beta1, beta2, eps, weight_decay);
      }
      adam_one(&net->b[l][j], grad->b[l][j], &m->b[l][j], &v->b[l][j], bc1, bc2, lr,
               beta1, beta2, eps, weight_decay);
    }
  }
}
