/*
 * lbm_oracle.c — CPU oracle (FP64, serial) for the D2Q13 thermal FD-LBM.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the
 * reported CPU baseline. It restates the reference numerics function by
 * function, keeping the reference's operation order (including pow() for
 * small integer powers) so that its FP64 trajectories follow the reference
 * as closely as plain C allows. Every function names the reference lines it
 * follows; the prefix CPP/ stands for CODE/C++/VS2010/simLB/simLB/.
 *
 * Built with -O2 -ffp-contract=off and no fast-math (see Makefile).
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define NQ 13

/* Mirror remaps of direction indices: about the x axis (north/south walls)
 * and about the y axis (east/west walls). CPP/cpuLBM.cpp:1156-1167 */
static const int k_mirror_ns[NQ] = {0, 1, 4, 3, 2, 8, 7, 6, 5, 9, 12, 11, 10};
static const int k_mirror_ew[NQ] = {0, 3, 2, 1, 4, 6, 5, 8, 7, 11, 10, 9, 12};

/* Wall closure tables: neighbour offsets, velocities pointing into the wall
 * from that neighbour, and the velocities leaving the wall back towards it.
 * CPP/cpuLBM.cpp:1518-1521 */
static const int k_wall_dx[4] = {-1, 1, 0, 0};
static const int k_wall_dy[4] = {0, 0, -1, 1};
static const int k_wall_in[4][4] = {{1, 5, 8, 9}, {3, 6, 7, 11}, {2, 5, 6, 10}, {4, 7, 8, 12}};
static const int k_wall_out[4][4] = {{3, 6, 7, 11}, {1, 5, 8, 9}, {4, 8, 7, 12}, {2, 6, 5, 10}};

struct orc {
  lbm_params p;
  int nx, ny;
  int* bnd;               /* [x][y] property id */
  int nprops;
  double *dens, *pres, *temp, *velx, *vely;
  int* ctype;

  double ex[NQ], ey[NQ];  /* lattice velocities times c */
  double b, K, Tc, c, dt;
  double tau_ref, u_ref, t_ref;
  int nsteps, step;
  double tstep;
  int failed;

  double *f, *h;          /* [x][y][i] */
  double *rho, *ux, *uy, *pr, *T; /* [x][y] */
  double* res_copy;

  /* RK3 temporaries, allocated once (the reference allocates per step,
   * CPP/cpuLBM.cpp:1742-1793; the arithmetic is the same) */
  double *fr1, *fr2, *fr3, *hr1, *hr2, *hr3;
  double *feq1, *heq1, *feq3, *heq3;
  double *rho1, *ux1, *uy1, *p1, *T1, *tf1, *th1, *thf1;
  double *rho3, *ux3, *uy3, *p3, *T3, *tf3, *th3, *thf3;
  double *flx_f2, *flx_h2, *flx_f3, *flx_h3; /* [x][y][i][2] */
  double *dF2, *dH2;                          /* [x][y][i] */

  double step_seconds;

  /* viscosity model of the OpenCL front end (PY/codeCL.py:59-69, 987-990):
   * 0 constant (deck mu), 1 VHS mu_ref (T/T_ref)^(0.5+param),
   * 2 Sutherland mu_ref (T/T_ref)^1.5 (T_ref+S)/(T+S) with S = param */
  int visc_model;
  double visc_mu_ref, visc_T_ref, visc_param;

  /* grid spacing per column and per row (uniform: the deck's dx, dy; the
   * OpenCL front end's non-uniform grids, PY/codeCL.py:611-618, 676-679) */
  double *dxs, *dys;
};

#define CELL(o, x, y) ((size_t)(x) * (size_t)(o)->ny + (size_t)(y))
#define DIST(o, x, y, i) (CELL(o, x, y) * NQ + (size_t)(i))

static int cell_type_at(const orc* o, int x, int y) {
  return o->ctype[o->bnd[CELL(o, x, y)]];
}

/* ------------------------------------------------------------------ */
/* Equilibrium polynomials. CPP/inlineFunctions.cpp:10-61 (the closed
 * forms are the model's definition and are kept term for term).       */
/* ------------------------------------------------------------------ */

static double eq_f0(double rho, double u, double v, double T) {
  return (rho / 4.0) * (4 + 10 * pow(T, 2) + pow(u, 4) - 5 * pow(v, 2) + pow(v, 4) +
                        10 * T * (-1 + pow(u, 2) + pow(v, 2)) + pow(u, 2) * (-5 + 4 * pow(v, 2)));
}

static double eq_f1(double rho, double u, double v, double T) {
  return (rho / 6.0) * (-6 * pow(T, 2) - u * (1 + u) * (-4 + pow(u, 2) + 3 * pow(v, 2)) -
                        T * (-4 + 6 * u + 9 * pow(u, 2) + 3 * pow(v, 2)));
}

static double eq_f5(double rho, double u, double v, double T) {
  return (rho / 4.0) * ((T + u + pow(u, 2)) * (T + v + pow(v, 2)));
}

static double eq_f9(double rho, double u, double v, double T) {
  return (rho / 24.0) * (3 * pow(T, 2) + (-1 + u) * u * (1 + u) * (2 + u) + T * (-1 + 6 * u * (1 + u)));
}

static double eq_h0(double K, double T, double u, double v) {
  return (10 * (16 + 3 * K) * pow(T, 3) +
          3 * T * (8 + 4 * K - 40 * pow(u, 2) - 5 * K * pow(u, 2) + 20 * pow(u, 4) + K * pow(u, 4) +
                   (-5 * (8 + K) + 4 * (15 + K) * pow(u, 2)) * pow(v, 2) + (20 + K) * pow(v, 4)) +
          30 * pow(T, 2) * (-4 + 9 * pow(u, 2) + 9 * pow(v, 2) + K * (-1 + pow(u, 2) + pow(v, 2))) +
          3 * (pow(u, 2) + pow(v, 2)) *
              (4 + pow(u, 4) - 5 * pow(v, 2) + pow(v, 4) + pow(u, 2) * (-5 + 4 * pow(v, 2)))) /
         24.0;
}

static double eq_h1(double K, double T, double u, double v) {
  return (-2 * (16 + 3 * K) * pow(T, 3) -
          u * (1 + u) * (pow(u, 2) + pow(v, 2)) * (-4 + pow(u, 2) + 3 * pow(v, 2)) -
          T * (u * (-4 * (4 + K) - 4 * (7 + K) * u + (14 + K) * pow(u, 2) + (19 + K) * pow(u, 3)) +
               (-4 + 3 * u * (10 + K + (14 + K) * u)) * pow(v, 2) + 3 * pow(v, 4)) -
          pow(T, 2) * (-16 + 6 * u * (6 + 13 * u) + 30 * pow(v, 2) +
                       K * (-4 + 6 * u + 9 * pow(u, 2) + 3 * pow(v, 2)))) /
         12.0;
}

static double eq_h5(double K, double T, double u, double v) {
  return ((16 + 3 * K) * pow(T, 3) + 3 * u * (1 + u) * v * (1 + v) * (pow(u, 2) + pow(v, 2)) +
          3 * pow(T, 2) * ((6 + K) * u + (9 + K) * pow(u, 2) + v * (6 + K + (9 + K) * v)) +
          3 * T *
              (pow(u, 3) + pow(u, 4) + pow(v, 3) * (1 + v) + u * v * (6 + K + (9 + K) * v) +
               pow(u, 2) * v * (9 + K + (12 + K) * v))) /
         24.0;
}

static double eq_h9(double K, double T, double u, double v) {
  return ((16 + 3 * K) * pow(T, 3) +
          T * u * (-8 + K * (-1 + u) * (1 + u) * (2 + u) + u * (-7 + 2 * u * (11 + 8 * u))) +
          T * (-1 + 6 * u * (1 + u)) * pow(v, 2) + (-1 + u) * u * (1 + u) * (2 + u) * (pow(u, 2) + pow(v, 2)) +
          pow(T, 2) * (-4 + 36 * u + 51 * pow(u, 2) + K * (-1 + 6 * u * (1 + u)) + 3 * pow(v, 2))) /
         48.0;
}

/* equilibrium2D: CPP/cpuLBM.cpp:785-844 */
static void equilibrium(orc* o, double rho, double ux, double uy, double Tin, double* eqf, double* eqh) {
  const double s = sqrt(o->p.R * o->Tc);
  const double u = ux / s;
  const double v = uy / s;
  const double T = Tin / o->Tc;
  const double K = o->K;

  eqf[0] = eq_f0(rho, u, v, T);
  eqf[1] = eq_f1(rho, u, v, T);
  eqf[2] = eq_f1(rho, v, u, T);
  eqf[3] = eq_f1(rho, -u, v, T);
  eqf[4] = eq_f1(rho, -v, u, T);
  eqf[5] = eq_f5(rho, u, v, T);
  eqf[6] = eq_f5(rho, -u, v, T);
  eqf[7] = eq_f5(rho, -u, -v, T);
  eqf[8] = eq_f5(rho, u, -v, T);
  eqf[9] = eq_f9(rho, u, v, T);
  eqf[10] = eq_f9(rho, v, u, T);
  eqf[11] = eq_f9(rho, -u, v, T);
  eqf[12] = eq_f9(rho, -v, u, T);

  const double scale = rho * o->p.R * o->Tc;
  eqh[0] = scale * eq_h0(K, T, u, v);
  eqh[1] = scale * eq_h1(K, T, u, v);
  eqh[2] = scale * eq_h1(K, T, v, u);
  eqh[3] = scale * eq_h1(K, T, -u, v);
  eqh[4] = scale * eq_h1(K, T, -v, u);
  eqh[5] = scale * eq_h5(K, T, u, v);
  eqh[6] = scale * eq_h5(K, T, -u, v);
  eqh[7] = scale * eq_h5(K, T, -u, -v);
  eqh[8] = scale * eq_h5(K, T, u, -v);
  eqh[9] = scale * eq_h9(K, T, u, v);
  eqh[10] = scale * eq_h9(K, T, v, u);
  eqh[11] = scale * eq_h9(K, T, -u, v);
  eqh[12] = scale * eq_h9(K, T, -v, u);

  /* mass self-check; a NaN never trips it (fabs(NaN) > x is false) */
  double sum = 0;
  for (int i = 0; i < NQ; i++) sum += eqf[i];
  if (fabs(sum - rho) > 1e-6) o->failed = 1;
}

/* ------------------------------------------------------------------ */
/* Moments. CPP/cpuLBM.cpp:877-1032                                    */
/* ------------------------------------------------------------------ */

typedef struct {
  double rho, ux, uy, T, p, tf, th, thf;
} macro_t;

/* temperature and pressure from the raw sums, CPP/cpuLBM.cpp:894-906 */
static void finish_macro(const orc* o, double rho, double rux, double ruy, double sh, macro_t* m) {
  const double ux = rux / rho;
  const double uy = ruy / rho;
  const double ax = fabs(ux), ay = fabs(uy);
  const double usq = ax * ax + ay * ay;
  const double T = 2.0 * (sh / rho - usq / 2.0) / (o->b * o->p.R);
  m->rho = rho;
  m->ux = ux;
  m->uy = uy;
  m->T = T;
  m->p = rho * o->p.R * T;
}

/* viscosity at temperature T (constant for the C++ solver; VHS and
 * Sutherland laws of PY/codeCL.py:987-990) */
static double viscosity(const orc* o, double T) {
  if (o->visc_model == 1) return o->visc_mu_ref * pow(T / o->visc_T_ref, 0.5 + o->visc_param);
  if (o->visc_model == 2)
    return o->visc_mu_ref * pow(T / o->visc_T_ref, 3.0 / 2.0) *
           ((o->visc_T_ref + o->visc_param) / (T + o->visc_param));
  return o->p.mu;
}

/* relaxation times, CPP/cpuLBM.cpp:952-955 (PY/codeCL.py:992-994) */
static void finish_tau(const orc* o, macro_t* m) {
  m->tf = viscosity(o, m->T) / m->p;
  m->th = m->tf / o->p.Pr;
  m->thf = (m->th * m->tf) / (m->tf - m->th);
}

/* macroProp / macroPropShort sums: CPP/cpuLBM.cpp:880-892, 920-932 */
static void moments(const orc* o, const double* f, const double* h, int x, int y, macro_t* m) {
  double rho = 0, rux = 0, ruy = 0, sh = 0;
  for (int i = 0; i < NQ; i++) {
    const double fi = f[DIST(o, x, y, i)];
    rho += fi;
    rux += fi * o->ex[i];
    ruy += fi * o->ey[i];
    sh += h[DIST(o, x, y, i)];
  }
  finish_macro(o, rho, rux, ruy, sh, m);
}

/* macroPropShort: CPP/cpuLBM.cpp:877-914 (writes the public fields) */
static void macro_short(orc* o, int x, int y) {
  macro_t m;
  moments(o, o->f, o->h, x, y, &m);
  const size_t c = CELL(o, x, y);
  o->rho[c] = m.rho;
  o->ux[c] = m.ux;
  o->uy[c] = m.uy;
  o->T[c] = m.T;
  o->pr[c] = m.p;
}

/* ------------------------------------------------------------------ */
/* Index mapping, mirror remap, sign. CPP/cpuLBM.cpp:1033-1169          */
/* ------------------------------------------------------------------ */

static int map_index(const orc* o, int i, int pm, int d, int* mir) {
  i += pm;
  *mir = 0;
  if (d == 0) {
    if (i < 0) {
      if (o->p.periodic_x > 0) i = o->nx + i;
      else if (o->p.mirror_west > 0) { i = -i - 1; *mir = 1; }
      else i = 0;
    } else if (i > o->nx - 1) {
      if (o->p.periodic_x > 0) i = i - o->nx;
      else if (o->p.mirror_east > 0) { i = 2 * o->nx - i - 1; *mir = 1; }
      else i = o->nx - 1;
    }
  } else {
    if (i < 0) {
      if (o->p.periodic_y > 0) i = o->ny + i;
      else if (o->p.mirror_south > 0) { i = -i - 1; *mir = 1; }
      else i = 0;
    } else if (i > o->ny - 1) {
      if (o->p.periodic_y > 0) i = i - o->ny;
      else if (o->p.mirror_north > 0) { i = 2 * o->ny - i - 1; *mir = 1; }
      else i = o->ny - 1;
    }
  }
  return i;
}

static int sgn(double a) { return a > 0 ? 1 : (a < 0 ? -1 : 0); }

static void mirror_dirs(int i, int mir_x, int mir_y, int* i_x, int* i_y) {
  *i_x = mir_x == 1 ? k_mirror_ns[i] : i;
  *i_y = mir_y == 1 ? k_mirror_ew[i] : i;
}

/* ------------------------------------------------------------------ */
/* Stencil with solid-point extrapolation. CPP/cpuLBM.cpp:1171-1326     */
/* ------------------------------------------------------------------ */

/* Replace the first solid run of one stencil line by the linear
 * extrapolation 2*f(m1) - f(m2) of the two nearest fluid points, and the
 * point beyond it by the same value. `d` is the axis, `e` the integer step
 * along it, `start` the first stencil index, `cx,cy` the fixed coordinate. */
static void extrapolate_line(const orc* o, const double* f, int i, int len, int d, int e, int start,
                             int fixed, const int* solid, double* S) {
  int diff = 0, jj = 0, pos = 0, mir;
  for (int j = 0; j < len - 1; j++) {
    diff = solid[j] - solid[j + 1];
    if (diff > 0) {
      jj = j;
      pos = map_index(o, start, e * j, d, &mir);
      break;
    } else if (diff < 0) {
      jj = j + 1;
      pos = map_index(o, start, e * (j + 1), d, &mir);
      break;
    }
  }
  const int m1 = map_index(o, pos, e * diff, d, &mir);
  const int m2 = map_index(o, pos, e * 2 * diff, d, &mir);
  if (d == 0) S[jj] = 2 * f[DIST(o, m1, fixed, i)] - f[DIST(o, m2, fixed, i)];
  else S[jj] = 2 * f[DIST(o, fixed, m1, i)] - f[DIST(o, fixed, m2, i)];
  const int next = jj - diff;
  if (next >= 0 && next < len) S[next] = S[jj];
}

static void stencil(const orc* o, const double* f, int x, int y, int i, int len, double* Sx, double* Sy) {
  if (i == 0) return;
  const int sx = sgn(o->ex[i]);
  const int sy = sgn(o->ey[i]);
  const int low = (len - 1) / 2;
  const int x1 = x - sx * low;
  const int y1 = y - sy * low;
  int solid_x[8], solid_y[8];
  int any_x = 0, any_y = 0;
  for (int j = 0; j < len; j++) {
    int mir_x, mir_y, i_x, i_y;
    const int jx = map_index(o, x1, sx * j, 0, &mir_y);
    const int jy = map_index(o, y1, sy * j, 1, &mir_x);
    mirror_dirs(i, mir_x, mir_y, &i_x, &i_y);
    Sx[j] = f[DIST(o, jx, y, i_y)];
    solid_x[j] = cell_type_at(o, jx, y) == 2;
    any_x |= solid_x[j];
    Sy[j] = f[DIST(o, x, jy, i_x)];
    solid_y[j] = cell_type_at(o, x, jy) == 2;
    any_y |= solid_y[j];
  }
  if (any_x) extrapolate_line(o, f, i, len, 0, sx, x1, y, solid_x, Sx);
  if (any_y) extrapolate_line(o, f, i, len, 1, sy, y1, x, solid_y, Sy);
}

/* ------------------------------------------------------------------ */
/* Face reconstructions. CPP/cpuLBM.cpp:1130-1145, 1328-1403            */
/* ------------------------------------------------------------------ */

static double minmod(double a, double b) {
  if ((fabs(a) < fabs(b)) && (a * b > 0)) return a;
  if ((fabs(b) < fabs(a)) && (a * b > 0)) return b;
  if (a == b) return a;
  if (a * b <= 0) return 0;
  return a + b; /* only reachable with NaN input */
}

/* NND: second-order TVD upwind face value, CPP/cpuLBM.cpp:1328-1356 */
static double nnd(double ee, const double* S) {
  const double v = fabs(ee);
  if (v == 0) return 0;
  const double fi = v * S[1], fp = v * S[2], fm = v * S[0];
  return fi + 0.5 * minmod(fp - fi, fi - fm);
}

/* WENO5 (Jiang-Shu) upwind face value of |e|*S, CPP/cpuLBM.cpp:1358-1403 */
static double weno5(double ee, const double* S) {
  const double eps = 1e-6;
  const double v = fabs(ee);
  if (v == 0) return 0;
  double q[5];
  for (int k = 0; k < 5; k++) q[k] = S[k] * v;
  const double b0 = (13.0 / 12.0) * pow(q[0] - 2 * q[1] + q[2], 2) + (1.0 / 4.0) * pow(q[0] - 4 * q[1] + 3 * q[2], 2);
  const double b1 = (13.0 / 12.0) * pow(q[1] - 2 * q[2] + q[3], 2) + (1.0 / 4.0) * pow(q[1] - q[3], 2);
  const double b2 = (13.0 / 12.0) * pow(q[2] - 2 * q[3] + q[4], 2) + (1.0 / 4.0) * pow(3 * q[2] - 4 * q[3] + q[4], 2);
  const double a0 = (1.0 / 10.0) * pow(1.0 / (eps + b0), 2);
  const double a1 = (6.0 / 10.0) * pow(1.0 / (eps + b1), 2);
  const double a2 = (3.0 / 10.0) * pow(1.0 / (eps + b2), 2);
  const double w0 = a0 / (a0 + a1 + a2);
  const double w1 = a1 / (a0 + a1 + a2);
  const double w2 = a2 / (a0 + a1 + a2);
  const double r0 = (2.0 / 6.0) * q[0] - (7.0 / 6.0) * q[1] + (11.0 / 6.0) * q[2];
  const double r1 = -(1.0 / 6.0) * q[1] + (5.0 / 6.0) * q[2] + (2.0 / 6.0) * q[3];
  const double r2 = (2.0 / 6.0) * q[2] + (5.0 / 6.0) * q[3] - (1.0 / 6.0) * q[4];
  return w0 * r0 + w1 * r1 + w2 * r2;
}

/* computeFlux: CPP/cpuLBM.cpp:1405-1434 */
static void cell_flux(const orc* o, const double* f, int x, int y, int i, double* fx, double* fy) {
  if (o->p.flux_method == 0) {
    double Sx[3] = {0, 0, 0}, Sy[3] = {0, 0, 0};
    stencil(o, f, x, y, i, 3, Sx, Sy);
    *fx = nnd(o->ex[i], Sx);
    *fy = nnd(o->ey[i], Sy);
  } else {
    double Sx[5] = {0, 0, 0, 0, 0}, Sy[5] = {0, 0, 0, 0, 0};
    stencil(o, f, x, y, i, 5, Sx, Sy);
    *fx = weno5(o->ex[i], Sx);
    *fy = weno5(o->ey[i], Sy);
  }
}

/* cFLUX: net flux out of (x,y) along direction i, CPP/cpuLBM.cpp:1436-1470 */
static double net_flux(const orc* o, const double* flux, int x, int y, int i) {
  if (i == 0) return 0;
  const int sx = sgn(o->ex[i]), sy = sgn(o->ey[i]);
  int mir_x, mir_y, i_x, i_y;
  const int xu = map_index(o, x, -sx, 0, &mir_y);
  const int yu = map_index(o, y, -sy, 1, &mir_x);
  mirror_dirs(i, mir_x, mir_y, &i_x, &i_y);
#if defined(ORC_PLANT_SPACING) && ORC_PLANT_SPACING == 1
  /* planted mistake, built only by the pin test: the upwind column's dx */
  const double rx = o->dt / o->dxs[xu], ry = o->dt / o->dys[y];
#elif defined(ORC_PLANT_SPACING) && ORC_PLANT_SPACING == 2
  /* planted mistake, built only by the pin test: dx and dy indices swapped
   * (valid on square grids only) */
  const double rx = o->dt / o->dxs[y], ry = o->dt / o->dys[x];
#else
  const double rx = o->dt / o->dxs[x], ry = o->dt / o->dys[y];
#endif
  const double out = rx * flux[2 * DIST(o, x, y, i)] + ry * flux[2 * DIST(o, x, y, i) + 1];
  const double in = rx * flux[2 * DIST(o, xu, y, i_y)] + ry * flux[2 * DIST(o, x, yu, i_x) + 1];
  return out - in;
}

/* posFluxes: CPP/cpuLBM.cpp:1471-1500 */
static void fluid_fluxes(orc* o, const double* f, const double* h, double* ff, double* hf) {
  for (int x = 0; x < o->nx; x++)
    for (int y = 0; y < o->ny; y++) {
      if (cell_type_at(o, x, y) == 2) continue;
      for (int i = 0; i < NQ; i++) {
        const size_t k = DIST(o, x, y, i);
        cell_flux(o, f, x, y, i, &ff[2 * k], &ff[2 * k + 1]);
        cell_flux(o, h, x, y, i, &hf[2 * k], &hf[2 * k + 1]);
      }
    }
}

/* wallFluxes: mass-balanced diffuse wall, CPP/cpuLBM.cpp:1502-1590. On a
 * non-uniform grid the fluxes through the wall cell's faces are weighted by
 * their face lengths as the OpenCL front end weights every face flux
 * (dy for x faces, dx for y faces, PY/codeCL.py:676-679, 722-737); a wall
 * cell's axis neighbours share its row (x faces) or column (y faces), so the
 * balance is the C++ one with the wall cell's own dx[x] and dy[y], which is
 * the deck's dx, dy on a uniform grid. */
static void wall_fluxes(orc* o, const double* f, double* ff, const double* h, double* hf) {
  for (int x = 0; x < o->nx; x++)
    for (int y = 0; y < o->ny; y++) {
      if (cell_type_at(o, x, y) != 2) continue;
#if defined(ORC_PLANT_SPACING) && ORC_PLANT_SPACING == 3
      /* planted mistake, built only by the pin test: the deck's uniform
       * spacing in the wall balance on a non-uniform grid */
      const double dx = o->p.dx, dy = o->p.dy;
#else
      const double dx = o->dxs[x], dy = o->dys[y];
#endif
      double in_f = 0, out_f = 0, in_h = 0, out_h = 0;
      for (int j = 0; j < 4; j++) {
        int mir_x, mir_y;
        const int xn = map_index(o, x, k_wall_dx[j], 0, &mir_y);
        const int yn = map_index(o, y, k_wall_dy[j], 1, &mir_x);
        if (cell_type_at(o, xn, yn) >= 2) continue;
        const double on_x = abs(k_wall_dx[j]);
        const double on_y = abs(k_wall_dy[j]);
        for (int k = 0; k < 4; k++) {
          const int i = k_wall_in[j][k];
          const int io = k_wall_out[j][k];
          const size_t kn = DIST(o, xn, yn, i);
          in_f += on_x * ff[2 * kn] / dx + on_y * ff[2 * kn + 1] / dy;
          in_h += on_x * hf[2 * kn] / dx + on_y * hf[2 * kn + 1] / dy;
          const double vx = fabs(o->ex[io]), vy = fabs(o->ey[io]);
          const double fw = f[DIST(o, x, y, io)], hw = h[DIST(o, x, y, io)];
          out_f += on_x * fw * vx / dx + on_y * fw * vy / dy;
          out_h += on_x * hw * vx / dx + on_y * hw * vy / dy;
        }
      }
      (void)in_h;
      (void)out_h; /* computed by the reference, never used */
      const double alpha = in_f / out_f;
      for (int i = 0; i < NQ; i++) {
        const size_t k = DIST(o, x, y, i);
        const double vx = fabs(o->ex[i]), vy = fabs(o->ey[i]);
        ff[2 * k] = alpha * vx * f[k];
        ff[2 * k + 1] = alpha * vy * f[k];
        hf[2 * k] = alpha * vx * h[k];
        hf[2 * k + 1] = alpha * vy * h[k];
      }
    }
}

/* boundaryFunctions: CPP/cpuLBM.cpp:1591-1596 */
static void boundary_fluxes(orc* o, const double* f, const double* h, double* ff, double* hf) {
  fluid_fluxes(o, f, h, ff, hf);
  wall_fluxes(o, f, ff, h, hf);
}

static double dot2(double a1, double a2, double b1, double b2) { return a1 * b1 + a2 * b2; }

/* ------------------------------------------------------------------ */
/* Steppers                                                             */
/* ------------------------------------------------------------------ */

/* RK3: IMEX-SSP2(3,2,2), CPP/cpuLBM.cpp:1738-2103 */
static void step_rk3(orc* o) {
  const int nx = o->nx, ny = o->ny;
  const double dt = o->dt;
  double eqf[NQ], eqh[NQ];

  /* stage 1: implicit relaxation over dt/2, CPP/cpuLBM.cpp:1819-1865 */
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++) {
      const size_t c = CELL(o, x, y);
      if (cell_type_at(o, x, y) > 0) {
        for (int i = 0; i < NQ; i++) {
          o->fr1[c * NQ + i] = o->f[c * NQ + i];
          o->hr1[c * NQ + i] = o->h[c * NQ + i];
        }
        continue;
      }
      macro_t m;
      moments(o, o->f, o->h, x, y, &m);
      finish_tau(o, &m);
      o->rho1[c] = m.rho; o->ux1[c] = m.ux; o->uy1[c] = m.uy; o->T1[c] = m.T; o->p1[c] = m.p;
      o->tf1[c] = m.tf; o->th1[c] = m.th; o->thf1[c] = m.thf;
      equilibrium(o, m.rho, m.ux, m.uy, m.T, eqf, eqh);
      for (int i = 0; i < NQ; i++) {
        const size_t k = c * NQ + i;
        o->feq1[k] = eqf[i];
        o->heq1[k] = eqh[i];
        const double fv = o->f[k], tf = m.tf, fe = eqf[i];
        o->fr1[k] = (fv + (dt / (2 * tf)) * fe) / (1 + dt / (2 * tf));
        const double hv = o->h[k], f1 = o->fr1[k], th = m.th, thf = m.thf, he = eqh[i];
        const double eu = dot2(o->ex[i], o->ey[i], m.ux, m.uy);
        o->hr1[k] = (hv - (dt / (2 * thf)) * eu * (fe - f1) + (dt / (2 * th)) * he) / (1 + dt / (2 * th));
      }
    }

  /* stage 2: feq2 = feq1, tau2 = tau1, u2 = u1, CPP/cpuLBM.cpp:1869-1924 */
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++) {
      const size_t c = CELL(o, x, y);
      if (cell_type_at(o, x, y) > 0) {
        for (int i = 0; i < NQ; i++) {
          o->fr2[c * NQ + i] = o->fr1[c * NQ + i];
          o->hr2[c * NQ + i] = o->hr1[c * NQ + i];
        }
        continue;
      }
      const double tf1 = o->tf1[c], tf2 = o->tf1[c];
      const double th1 = o->th1[c], th2 = o->th1[c];
      const double thf1 = o->thf1[c], thf2 = o->thf1[c];
      for (int i = 0; i < NQ; i++) {
        const size_t k = c * NQ + i;
        const double fv = o->f[k], fe1 = o->feq1[k], fe2 = o->feq1[k], f1 = o->fr1[k];
        o->fr2[k] = (fv - (dt / (2 * tf1)) * (fe1 - f1) + (dt / (2 * tf2)) * fe2) / (1 + dt / (2 * tf2));
        const double hv = o->h[k], he1 = o->heq1[k], he2 = o->heq1[k], h1 = o->hr1[k];
        const double eu1 = dot2(o->ex[i], o->ey[i], o->ux1[c], o->uy1[c]);
        const double eu2 = eu1;
        const double f2 = o->fr2[k];
        o->hr2[k] = (hv - (dt / 2) * ((eu2 / thf2) * (fe2 - f2) - (eu1 / thf1) * (fe1 - f1)) -
                     (dt / (2 * th1)) * (he1 - h1) + (dt / (2 * th2)) * he2) /
                    (1 + dt / (2 * th2));
      }
    }

  /* stage 3: moments of f - F(f2), implicit relaxation, CPP/cpuLBM.cpp:1926-1993 */
  boundary_fluxes(o, o->fr2, o->hr2, o->flx_f2, o->flx_h2);
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++) {
      const size_t c = CELL(o, x, y);
      if (cell_type_at(o, x, y) > 0) {
        for (int i = 0; i < NQ; i++) {
          o->fr3[c * NQ + i] = o->fr2[c * NQ + i];
          o->hr3[c * NQ + i] = o->hr2[c * NQ + i];
        }
        continue;
      }
      /* macroProp3, CPP/cpuLBM.cpp:973-1032 */
      double rho = 0, rux = 0, ruy = 0, sh = 0;
      for (int i = 0; i < NQ; i++) {
        const size_t k = c * NQ + i;
        const double dfl = net_flux(o, o->flx_f2, x, y, i);
        const double dhl = net_flux(o, o->flx_h2, x, y, i);
        rho += o->f[k] - dfl;
        rux += o->f[k] * o->ex[i] - dfl * o->ex[i];
        ruy += o->f[k] * o->ey[i] - dfl * o->ey[i];
        sh += o->h[k] - dhl;
      }
      macro_t m;
      finish_macro(o, rho, rux, ruy, sh, &m);
      finish_tau(o, &m);
      o->rho3[c] = m.rho; o->ux3[c] = m.ux; o->uy3[c] = m.uy; o->T3[c] = m.T; o->p3[c] = m.p;
      o->tf3[c] = m.tf; o->th3[c] = m.th; o->thf3[c] = m.thf;
      equilibrium(o, m.rho, m.ux, m.uy, m.T, eqf, eqh);
      const double tf2 = o->tf1[c], th2 = o->th1[c], thf2 = o->thf1[c];
      for (int i = 0; i < NQ; i++) {
        const size_t k = c * NQ + i;
        o->feq3[k] = eqf[i];
        o->heq3[k] = eqh[i];
        const double fv = o->f[k], f2 = o->fr2[k], fe2 = o->feq1[k], fe3 = eqf[i];
        const double dF2 = net_flux(o, o->flx_f2, x, y, i);
#if defined(ORC_PLANT_RK) && ORC_PLANT_RK == 2
        /* planted mistake, built only by the pin test: stage 3 without the
         * stage-2 flux */
        o->fr3[k] = (fv + (dt / (2 * tf2)) * (fe2 - f2) + (dt / (2 * m.tf)) * fe3) / (1 + dt / (2 * m.tf));
#else
        o->fr3[k] = (fv - dF2 + (dt / (2 * tf2)) * (fe2 - f2) + (dt / (2 * m.tf)) * fe3) / (1 + dt / (2 * m.tf));
#endif
        const double hv = o->h[k], h2 = o->hr2[k], he2 = o->heq1[k], he3 = eqh[i], f3 = o->fr3[k];
        const double eu2 = dot2(o->ex[i], o->ey[i], o->ux1[c], o->uy1[c]);
        const double eu3 = dot2(o->ex[i], o->ey[i], m.ux, m.uy);
        const double dH2 = net_flux(o, o->flx_h2, x, y, i);
        o->hr3[k] = (hv - dH2 - (dt / 2) * ((eu2 / thf2) * (fe2 - f2) + (eu3 / m.thf) * (fe3 - f3)) +
                     (dt / (2 * th2)) * (he2 - h2) + (dt / (2 * m.th)) * he3) /
                    (1 + dt / (2 * m.th));
        o->dF2[k] = dF2;
        o->dH2[k] = dH2;
      }
    }

  /* final combination, CPP/cpuLBM.cpp:1995-2042 */
  boundary_fluxes(o, o->fr3, o->hr3, o->flx_f3, o->flx_h3);
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++) {
      const size_t c = CELL(o, x, y);
      if (cell_type_at(o, x, y) > 0) continue;
      const double tf2 = o->tf1[c], th2 = o->th1[c], thf2 = o->thf1[c];
      const double tf3 = o->tf3[c], th3 = o->th3[c], thf3 = o->thf3[c];
      for (int i = 0; i < NQ; i++) {
        const size_t k = c * NQ + i;
        const double fv = o->f[k], dF2 = o->dF2[k], fe2 = o->feq1[k], f2 = o->fr2[k];
        const double fe3 = o->feq3[k], f3 = o->fr3[k];
        const double dF3 = net_flux(o, o->flx_f3, x, y, i);
#if defined(ORC_PLANT_RK) && ORC_PLANT_RK == 1
        /* planted mistake, built only by the pin test: the final combination
         * takes the whole stage-2 flux instead of the mean of the two */
        o->f[k] = fv - dF2 + (dt / 2.0) * ((fe2 - f2) / tf2 + (fe3 - f3) / tf3);
#else
        o->f[k] = fv - 0.5 * (dF2 + dF3) + (dt / 2.0) * ((fe2 - f2) / tf2 + (fe3 - f3) / tf3);
#endif
        const double hv = o->h[k], dH2 = o->dH2[k], he2 = o->heq1[k], h2 = o->hr2[k];
        const double he3 = o->heq3[k], h3 = o->hr3[k];
        const double eu2 = dot2(o->ex[i], o->ey[i], o->ux1[c], o->uy1[c]);
        const double eu3 = dot2(o->ex[i], o->ey[i], o->ux3[c], o->uy3[c]);
        const double dH3 = net_flux(o, o->flx_h3, x, y, i);
#if defined(ORC_PLANT_RK) && ORC_PLANT_RK == 1
        o->h[k] = hv - dH2 + (dt / 2.0) * ((he2 - h2) / th2 + (he3 - h3) / th3) -
#else
        o->h[k] = hv - 0.5 * (dH2 + dH3) + (dt / 2.0) * ((he2 - h2) / th2 + (he3 - h3) / th3) -
#endif
#if defined(ORC_PLANT_RK) && ORC_PLANT_RK == 3
                  /* planted mistake, built only by the pin test: the f-h
                   * coupling term of the final combination with its sign flipped */
                  -(dt / 2.0) * ((eu2 / thf2) * (fe2 - f2) + (eu3 / thf3) * (fe3 - f3));
#else
                  (dt / 2.0) * ((eu2 / thf2) * (fe2 - f2) + (eu3 / thf3) * (fe3 - f3));
#endif
      }
    }

  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
      if (cell_type_at(o, x, y) != 2) macro_short(o, x, y);
}

/* RK1: IMEX one-stage, CPP/cpuLBM.cpp:1597-1737 */
static void step_rk1(orc* o) {
  const int nx = o->nx, ny = o->ny;
  const double dt = o->dt;
  double eqf[NQ], eqh[NQ];
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++) {
      const size_t c = CELL(o, x, y);
      if (cell_type_at(o, x, y) > 0) {
        for (int i = 0; i < NQ; i++) {
          o->fr1[c * NQ + i] = o->f[c * NQ + i];
          o->hr1[c * NQ + i] = o->h[c * NQ + i];
        }
        continue;
      }
      macro_t m;
      moments(o, o->f, o->h, x, y, &m);
      finish_tau(o, &m);
      /* RK1 writes these into the public fields, CPP/cpuLBM.cpp:1635 */
      o->rho[c] = m.rho; o->ux[c] = m.ux; o->uy[c] = m.uy; o->T[c] = m.T; o->pr[c] = m.p;
      o->tf1[c] = m.tf; o->th1[c] = m.th; o->thf1[c] = m.thf;
      equilibrium(o, m.rho, m.ux, m.uy, m.T, eqf, eqh);
      for (int i = 0; i < NQ; i++) {
        const size_t k = c * NQ + i;
        o->feq1[k] = eqf[i];
        o->heq1[k] = eqh[i];
        const double fv = o->f[k], tf = m.tf, fe = eqf[i];
        o->fr1[k] = (fv + (dt / tf) * fe) / (1.0 + dt / tf);
        const double hv = o->h[k], f1 = o->fr1[k], th = m.th, thf = m.thf, he = eqh[i];
        const double eu = dot2(o->ex[i], o->ey[i], m.ux, m.uy);
        o->hr1[k] = (hv - dt * eu * ((fe - f1) / thf) + (dt / th) * he) / (1 + dt / th);
      }
    }
  boundary_fluxes(o, o->fr1, o->hr1, o->flx_f2, o->flx_h2);
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++) {
      const size_t c = CELL(o, x, y);
      if (cell_type_at(o, x, y) > 0) continue;
      for (int i = 0; i < NQ; i++) {
        const size_t k = c * NQ + i;
        const double fv = o->f[k], f1 = o->fr1[k], tf = o->tf1[c], fe = o->feq1[k];
        const double dF = net_flux(o, o->flx_f2, x, y, i);
        o->f[k] = fv - dF + (dt / tf) * (fe - f1);
        const double hv = o->h[k], h1 = o->hr1[k], th = o->th1[c], thf = o->thf1[c], he = o->heq1[k];
        const double eu = dot2(o->ex[i], o->ey[i], o->ux[c], o->uy[c]);
        const double dH = net_flux(o, o->flx_h2, x, y, i);
        o->h[k] = hv - dH + (dt / th) * (he - h1) - ((dt * eu) / thf) * (fe - f1);
      }
    }
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
      if (cell_type_at(o, x, y) != 2) macro_short(o, x, y);
}

/* EULER (equilibrium_only = 0), CPP/cpuLBM.cpp:2104-2193, and
 * EULER_EQ (equilibrium_only = 1), CPP/cpuLBM.cpp:2194-2283 */
static void step_euler(orc* o, int equilibrium_only) {
  const int nx = o->nx, ny = o->ny;
  const double dt = o->dt;
  double eqf[NQ], eqh[NQ];
  boundary_fluxes(o, o->f, o->h, o->flx_f2, o->flx_h2);
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++) {
      const size_t c = CELL(o, x, y);
      if (cell_type_at(o, x, y) > 0) continue;
      macro_t m;
      moments(o, o->f, o->h, x, y, &m);
      finish_tau(o, &m);
      o->rho[c] = m.rho; o->ux[c] = m.ux; o->uy[c] = m.uy; o->T[c] = m.T; o->pr[c] = m.p;
      equilibrium(o, m.rho, m.ux, m.uy, m.T, eqf, eqh);
      for (int i = 0; i < NQ; i++) {
        const size_t k = c * NQ + i;
        const double fv = o->f[k], fe = eqf[i];
        const double dF = net_flux(o, o->flx_f2, x, y, i);
        const double dH = net_flux(o, o->flx_h2, x, y, i);
        if (equilibrium_only) {
          o->f[k] = fe - dF;
          o->h[k] = eqh[i] - dH;
        } else {
          o->f[k] = fv - dF + (dt / m.tf) * (fe - fv);
          const double hv = o->h[k], he = eqh[i];
          const double eu = dot2(o->ex[i], o->ey[i], m.ux, m.uy);
          o->h[k] = hv - dH + (dt / m.th) * (he - hv) - ((dt * eu) / m.thf) * (fe - fv);
        }
      }
    }
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
      if (cell_type_at(o, x, y) != 2) macro_short(o, x, y);
}

/* runOne: CPP/cpuLBM.cpp:2284-2335 */
static int run_one(orc* o, int want_res, double* res) {
  o->failed = 0;
  const size_t n = (size_t)o->nx * o->ny;
  if (want_res) memcpy(o->res_copy, o->T, n * sizeof(double));
  switch (o->p.rk_method) {
    case 0: step_rk1(o); break;
    case 1: step_rk3(o); break;
    case 2: step_euler(o, 0); break;
    case 3: step_euler(o, 1); break;
    default: break;
  }
  if (want_res) {
    double s = 0;
    for (size_t k = 0; k < n; k++) s += (o->T[k] - o->res_copy[k]) * (o->T[k] - o->res_copy[k]);
#if defined(ORC_PLANT_RES) && ORC_PLANT_RES == 1
    /* planted mistake, built only by the pin test: the mean over one row */
    *res = sqrt(s / (double)o->nx);
#else
    *res = sqrt(s / (double)n);
#endif
  }
  return o->failed;
}

/* ------------------------------------------------------------------ */
/* Set-up: refineInput, setupVars, initialiseFunctions                  */
/* ------------------------------------------------------------------ */

/* CPP/cpuLBM.cpp:283-358 */
static void refine_input(orc* o) {
  lbm_params* p = &o->p;
  if (p->dx == 0) p->dx = p->Lx / ((double)o->nx - 1);
  else p->Lx = p->dx * (o->nx - 1);
  if (p->dy == 0 && p->Ly != 0) p->dy = p->Ly / ((double)o->ny - 1);
  else if (p->dy == 0 && p->Ly == 0) { p->dy = p->dx; p->Ly = p->dy * (o->ny - 1); }
  else p->Ly = p->dy * (o->ny - 1);

  if (p->rho_ref == 0) p->rho_ref = p->p_ref / (p->R * p->T_ref);
  else if (p->p_ref == 0) p->p_ref = p->rho_ref * p->R * p->T_ref;
  else if (p->T_ref == 0) p->T_ref = p->p_ref / (p->rho_ref * p->R);

  for (int r = 0; r < o->nprops; r++) {
    const double rho = o->dens[r], pr = o->pres[r], T = o->temp[r];
    if (rho == 0) o->dens[r] = pr / (p->R * T);
    else if (pr == 0) o->pres[r] = rho * p->R * T;
    else if (T == 0) o->temp[r] = pr / (rho * p->R);
  }
}

/* CPP/cpuLBM.cpp:360-450 */
static void steps_from_dt(orc* o);

static void setup_vars(orc* o) {
  static const int lat_x[NQ] = {0, 1, 0, -1, 0, 1, -1, -1, 1, 2, 0, -2, 0};
  static const int lat_y[NQ] = {0, 0, 1, 0, -1, 1, 1, -1, -1, 0, 2, 0, -2};
  lbm_params* p = &o->p;
  o->b = 2 / (p->gamma - 1);
  o->K = o->b - 2;
  o->Tc = p->Tc_Tref * p->T_ref;
  o->tau_ref = p->mu / p->p_ref;
  o->u_ref = sqrt(p->R * p->T_ref);
  o->t_ref = (p->Lx > p->Ly ? p->Lx : p->Ly) / o->u_ref;
  o->c = sqrt(p->R * o->Tc);
  for (int i = 0; i < NQ; i++) {
    o->ex[i] = lat_x[i] * o->c;
    o->ey[i] = lat_y[i] * o->c;
  }
  if (p->dtau == 0) o->dt = (p->cfl * (p->dx < p->dy ? p->dx : p->dy)) / o->ex[9];
  else o->dt = p->dtau * o->tau_ref;
  steps_from_dt(o);
}

/* nsteps from tt_tref, tt_time or steps (CPP/cpuLBM.cpp:423-448) */
static void steps_from_dt(orc* o) {
  const lbm_params* p = &o->p;
  double total = 0;
  int need = 0;
  o->nsteps = 0;
  if (p->tt_tref > 0) { total = p->tt_tref * o->t_ref; need = 1; }
  else if (p->tt_time > 0) { total = p->tt_time; need = 1; }
  else if (p->steps > 0) o->nsteps = p->steps;
  else if (p->tol > 0) o->nsteps = 0;
  if (need) o->nsteps = (int)ceil(total / o->dt);
}

/* CPP/cpuLBM.cpp:846-875 */
static void init_functions(orc* o) {
  double eqf[NQ], eqh[NQ];
  for (int y = 0; y < o->ny; y++)
    for (int x = 0; x < o->nx; x++) {
      const int id = o->bnd[CELL(o, x, y)];
      equilibrium(o, o->dens[id], o->velx[id], o->vely[id], o->temp[id], eqf, eqh);
      for (int i = 0; i < NQ; i++) {
        o->f[DIST(o, x, y, i)] = eqf[i];
        o->h[DIST(o, x, y, i)] = eqh[i];
      }
      macro_short(o, x, y);
    }
}

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

orc* orc_create(const lbm_params* p, int32_t nx, int32_t ny, const int32_t* bnd, const lbm_props* props) {
  if (!p || !bnd || !props || nx < 1 || ny < 1 || props->num_props < 1) return NULL;
  orc* o = (orc*)calloc(1, sizeof(orc));
  o->p = *p;
  o->nx = nx;
  o->ny = ny;
  const size_t n = (size_t)nx * ny, nq = n * NQ;
  o->nprops = props->num_props;
  o->bnd = (int*)malloc(n * sizeof(int));
  for (size_t k = 0; k < n; k++) {
    if (bnd[k] < 0 || bnd[k] >= o->nprops) { free(o->bnd); free(o); return NULL; }
    o->bnd[k] = bnd[k];
  }
  o->dens = dalloc(o->nprops); o->pres = dalloc(o->nprops); o->temp = dalloc(o->nprops);
  o->velx = dalloc(o->nprops); o->vely = dalloc(o->nprops);
  o->ctype = (int*)malloc(o->nprops * sizeof(int));
  for (int r = 0; r < o->nprops; r++) {
    o->dens[r] = props->density[r];
    o->pres[r] = props->pressure[r];
    o->temp[r] = props->temp[r];
    o->velx[r] = props->velocity_x[r];
    o->vely[r] = props->velocity_y[r];
    o->ctype[r] = props->cell_type[r];
  }
  o->f = dalloc(nq); o->h = dalloc(nq);
  o->rho = dalloc(n); o->ux = dalloc(n); o->uy = dalloc(n); o->pr = dalloc(n); o->T = dalloc(n);
  o->res_copy = dalloc(n);
  o->fr1 = dalloc(nq); o->fr2 = dalloc(nq); o->fr3 = dalloc(nq);
  o->hr1 = dalloc(nq); o->hr2 = dalloc(nq); o->hr3 = dalloc(nq);
  o->feq1 = dalloc(nq); o->heq1 = dalloc(nq); o->feq3 = dalloc(nq); o->heq3 = dalloc(nq);
  o->rho1 = dalloc(n); o->ux1 = dalloc(n); o->uy1 = dalloc(n); o->p1 = dalloc(n); o->T1 = dalloc(n);
  o->tf1 = dalloc(n); o->th1 = dalloc(n); o->thf1 = dalloc(n);
  o->rho3 = dalloc(n); o->ux3 = dalloc(n); o->uy3 = dalloc(n); o->p3 = dalloc(n); o->T3 = dalloc(n);
  o->tf3 = dalloc(n); o->th3 = dalloc(n); o->thf3 = dalloc(n);
  o->flx_f2 = dalloc(2 * nq); o->flx_h2 = dalloc(2 * nq);
  o->flx_f3 = dalloc(2 * nq); o->flx_h3 = dalloc(2 * nq);
  o->dF2 = dalloc(nq); o->dH2 = dalloc(nq);
  refine_input(o);
  setup_vars(o);
  o->dxs = dalloc(nx); o->dys = dalloc(ny);
  for (int x = 0; x < nx; x++) o->dxs[x] = o->p.dx;
  for (int y = 0; y < ny; y++) o->dys[y] = o->p.dy;
  init_functions(o);
  o->step = 0;
  o->tstep = 0;
  return o;
}

void orc_destroy(orc* o) {
  if (!o) return;
  double** bufs[] = {&o->dens, &o->pres, &o->temp, &o->velx, &o->vely, &o->f, &o->h, &o->rho, &o->ux,
                     &o->uy, &o->pr, &o->T, &o->res_copy, &o->fr1, &o->fr2, &o->fr3, &o->hr1, &o->hr2,
                     &o->hr3, &o->feq1, &o->heq1, &o->feq3, &o->heq3, &o->rho1, &o->ux1, &o->uy1, &o->p1,
                     &o->T1, &o->tf1, &o->th1, &o->thf1, &o->rho3, &o->ux3, &o->uy3, &o->p3, &o->T3,
                     &o->tf3, &o->th3, &o->thf3, &o->flx_f2, &o->flx_h2, &o->flx_f3, &o->flx_h3,
                     &o->dF2, &o->dH2, &o->dxs, &o->dys};
  for (size_t k = 0; k < sizeof(bufs) / sizeof(bufs[0]); k++) free(*bufs[k]);
  free(o->bnd);
  free(o->ctype);
  free(o);
}

int orc_step(orc* o, int32_t nsteps, int32_t want_res, double* res_out) {
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int failed = 0;
  double res = 0;
  for (int s = 0; s < nsteps; s++) {
    o->step += 1; /* updateStepCount, CPP/cpuLBM.cpp:771-775 */
    o->tstep = o->step * o->dt;
    failed = run_one(o, want_res, &res);
    if (failed) break;
  }
  clock_gettime(CLOCK_MONOTONIC, &t1);
  o->step_seconds += (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  if (res_out) *res_out = res;
  return failed;
}

void orc_get_dist(const orc* o, double* f, double* h) {
  const size_t nq = (size_t)o->nx * o->ny * NQ;
  if (f) memcpy(f, o->f, nq * sizeof(double));
  if (h) memcpy(h, o->h, nq * sizeof(double));
}

void orc_set_dist(orc* o, const double* f, const double* h) {
  const size_t nq = (size_t)o->nx * o->ny * NQ;
  if (f) memcpy(o->f, f, nq * sizeof(double));
  if (h) memcpy(o->h, h, nq * sizeof(double));
  for (int x = 0; x < o->nx; x++)
    for (int y = 0; y < o->ny; y++)
      if (cell_type_at(o, x, y) != 2) macro_short(o, x, y);
}

void orc_get_macro(const orc* o, double* rho, double* ux, double* uy, double* p, double* T) {
  const size_t n = (size_t)o->nx * o->ny;
  if (rho) memcpy(rho, o->rho, n * sizeof(double));
  if (ux) memcpy(ux, o->ux, n * sizeof(double));
  if (uy) memcpy(uy, o->uy, n * sizeof(double));
  if (p) memcpy(p, o->pr, n * sizeof(double));
  if (T) memcpy(T, o->T, n * sizeof(double));
}

void orc_get_consts(const orc* o, double* out) {
  out[0] = o->dt; out[1] = o->c; out[2] = o->Tc; out[3] = o->b; out[4] = o->K;
  out[5] = o->p.dx; out[6] = o->p.dy; out[7] = o->nsteps; out[8] = o->step; out[9] = o->tstep;
}

void orc_equilibrium(orc* o, double rho, double ux, double uy, double T, double* eqf, double* eqh,
                     int32_t* failed) {
  const int saved = o->failed;
  o->failed = 0;
  equilibrium(o, rho, ux, uy, T, eqf, eqh);
  if (failed) *failed = o->failed;
  o->failed = saved;
}

double orc_weno5(double ee, const double* S) { return weno5(ee, S); }
double orc_nnd(double ee, const double* S) { return nnd(ee, S); }
double orc_minmod(double a, double b) { return minmod(a, b); }

int32_t orc_index(const orc* o, int32_t i, int32_t pm, int32_t d, int32_t* mir) {
  int m;
  const int r = map_index(o, i, pm, d, &m);
  if (mir) *mir = m;
  return r;
}

void orc_mirror_dirs(int32_t i, int32_t mir_x, int32_t mir_y, int32_t* i_x, int32_t* i_y) {
  int a, b;
  mirror_dirs(i, mir_x, mir_y, &a, &b);
  *i_x = a;
  *i_y = b;
}

void orc_stencil(const orc* o, const double* f, int32_t x, int32_t y, int32_t i, int32_t len, double* Sx,
                 double* Sy) {
  stencil(o, f, x, y, i, len, Sx, Sy);
}

void orc_boundary_fluxes(orc* o, const double* f, const double* h, double* f_flux, double* h_flux) {
  boundary_fluxes(o, f, h, f_flux, h_flux);
}

double orc_cflux(const orc* o, const double* flux, int32_t x, int32_t y, int32_t i) {
  return net_flux(o, flux, x, y, i);
}

double orc_step_seconds(const orc* o) { return o->step_seconds; }

int orc_set_viscosity(orc* o, int32_t model, double mu_ref, double T_ref, double param) {
  if (model < 0 || model > 2) return -1;
  if (model > 0 && !(mu_ref > 0 && T_ref > 0)) return -1;
  o->visc_model = model;
  o->visc_mu_ref = mu_ref;
  o->visc_T_ref = T_ref;
  o->visc_param = param;
  return 0;
}

/* Non-uniform spacing, after the OpenCL front end: face fluxes weighted by
 * their face length (dy[y] for x faces, dx[x] for y faces) and the net
 * divided by the cell area (PY/codeCL.py:611-618, 676-679). The two cells
 * of a face share its length, so cFLUX becomes dt/dx[x] and dt/dy[y] of the
 * cell; dt comes from the smallest spacing (PY/LBM.py:216). Walls keep the
 * C++ closure with face-length weights (wall_fluxes). Refused (-1) for
 * non-positive spacings. */
int orc_set_spacing(orc* o, const double* dx, const double* dy) {
  if (!o || !dx || !dy) return -1;
  for (int x = 0; x < o->nx; x++)
    if (!(dx[x] > 0) || !isfinite(dx[x])) return -1;
  for (int y = 0; y < o->ny; y++)
    if (!(dy[y] > 0) || !isfinite(dy[y])) return -1;
  double mn = dx[0];
  for (int x = 0; x < o->nx; x++) {
    o->dxs[x] = dx[x];
    if (dx[x] < mn) mn = dx[x];
  }
  for (int y = 0; y < o->ny; y++) {
    o->dys[y] = dy[y];
    if (dy[y] < mn) mn = dy[y];
  }
  if (o->p.dtau == 0) {
    o->dt = (o->p.cfl * mn) / o->ex[9];
    steps_from_dt(o);
  }
  return 0;
}
