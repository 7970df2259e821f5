/*
 * sanitize_check.c — drives every entry point of the CPU oracle on small
 * decks (periodic, diffuse walls, mirrors, clamped edges, permanent cells;
 * WENO5 and NND; every RKMETHOD; viscosity laws; non-uniform spacing) so
 * that a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle sanitize`) checks the oracle's memory use and arithmetic.
 * TEST INFRASTRUCTURE ONLY, like the oracle itself. Exit status 0 when every
 * call behaved.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lbm_oracle.h"

static lbm_params deck(int rk, int fm) {
  lbm_params p;
  memset(&p, 0, sizeof(p));
  p.rk_method = rk;
  p.flux_method = fm;
  p.cfl = 0.3;
  p.steps = 4;
  p.gamma = 1.4;
  p.R = 1.0;
  p.mu = 0.05;
  p.Pr = 0.71;
  p.dx = p.dy = 1.0 / 14.0;
  p.rho_ref = 1.0;
  p.p_ref = 1.0;
  p.T_ref = 1.0;
  p.Tc_Tref = 2.5;
  return p;
}

/* Property rows: 0 fluid at rest, 1 fluid moving, 2 permanent inflow, 3 solid
 * wall at rest, 4 moving solid wall. */
static const double k_rho[5] = {1.0, 1.2, 1.0, 1.0, 1.0};
static const double k_pres[5] = {1.0, 0.0, 1.0, 1.0, 1.0};
static const double k_temp[5] = {1.0, 1.0, 0.0, 1.0, 1.0};
static const double k_ux[5] = {0.0, 0.1, 0.3, 0.0, 0.12};
static const double k_uy[5] = {0.0, -0.05, 0.0, 0.0, 0.0};
static const int32_t k_type[5] = {0, 0, 1, 2, 2};

static int run(const char* name, lbm_params p, int nx, int ny, int geometry, int visc, int spacing) {
  int32_t* bnd = (int32_t*)malloc(sizeof(int32_t) * (size_t)nx * ny);
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++) {
      int id = ((x + y) % 3 == 0) ? 1 : 0;
      if (geometry == 1 && y == 0) id = 3;                 /* bottom wall */
      if (geometry == 1 && y == ny - 1) id = 4;            /* moving lid */
      if (geometry == 2 && x == 0) id = 2;                 /* inflow column */
      if (geometry == 2 && y == 0 && x > nx / 3) id = 3;   /* plate */
      bnd[(size_t)x * ny + y] = id;
    }
  lbm_props pr = {5, k_rho, k_pres, k_temp, k_ux, k_uy, k_type};
  orc* o = orc_create(&p, nx, ny, bnd, &pr);
  if (!o) {
    fprintf(stderr, "%s: orc_create failed\n", name);
    free(bnd);
    return 1;
  }
  if (visc) orc_set_viscosity(o, visc, 0.05, 1.0, visc == 1 ? 0.25 : 0.4);
  const size_t nq = (size_t)nx * ny * 13;
  if (spacing) {
    double* dx = (double*)malloc(sizeof(double) * nx);
    double* dy = (double*)malloc(sizeof(double) * ny);
    for (int x = 0; x < nx; x++) dx[x] = p.dx * (1.0 + 0.3 * sin(6.283185307179586 * x / nx));
    for (int y = 0; y < ny; y++) dy[y] = p.dy * (1.0 + 0.3 * cos(6.283185307179586 * y / ny));
    if (orc_set_spacing(o, dx, dy) != 0) fprintf(stderr, "%s: spacing refused\n", name);
    free(dx);
    free(dy);
  }
  double res = 0;
  int st = orc_step(o, 3, 1, &res);
  double* f = (double*)malloc(sizeof(double) * nq);
  double* h = (double*)malloc(sizeof(double) * nq);
  double* ff = (double*)malloc(sizeof(double) * 2 * nq);
  double* hf = (double*)malloc(sizeof(double) * 2 * nq);
  double* m = (double*)malloc(sizeof(double) * (size_t)nx * ny * 5);
  orc_get_dist(o, f, h);
  orc_set_dist(o, f, h);
  orc_get_macro(o, m, m + nx * ny, m + 2 * nx * ny, m + 3 * nx * ny, m + 4 * nx * ny);
  orc_boundary_fluxes(o, f, h, ff, hf);
  double sum = 0;
  for (int x = 0; x < nx; x++)
    for (int y = 0; y < ny; y++)
      for (int i = 0; i < 13; i++) sum += orc_cflux(o, ff, x, y, i);
  double Sx[5], Sy[5], eqf[13], eqh[13], consts[10];
  int32_t failed = 0, mir = 0;
  orc_stencil(o, f, 0, 0, 9, 5, Sx, Sy);
  orc_stencil(o, f, nx - 1, ny - 1, 7, 3, Sx, Sy);
  orc_equilibrium(o, 1.0, 0.1, -0.1, 1.0, eqf, eqh, &failed);
  orc_get_consts(o, consts);
  (void)orc_index(o, -2, 0, 0, &mir);
  (void)orc_index(o, ny + 1, 0, 1, &mir);
  st |= orc_step(o, 1, 0, NULL);
  printf("%s: status %d residual %.3e net-flux sum %.3e step seconds %.3f\n", name, st, res, sum,
         orc_step_seconds(o));
  free(f);
  free(h);
  free(ff);
  free(hf);
  free(m);
  free(bnd);
  orc_destroy(o);
  return 0;
}

int main(void) {
  int bad = 0;
  const double S[5] = {1.0, 1.5, 0.7, 2.0, -1.0};
  printf("weno5 %.6f nnd %.6f minmod %.3f\n", orc_weno5(2.0, S), orc_nnd(-1.0, S), orc_minmod(-1.0, 2.0));
  for (int rk = 0; rk < 4; rk++) {
    lbm_params p = deck(rk, 1);
    p.periodic_x = p.periodic_y = 1;
    char name[64];
    snprintf(name, sizeof(name), "periodic rk%d", rk);
    bad |= run(name, p, 14, 12, 0, 0, rk == 1);
  }
  lbm_params w = deck(1, 1);
  w.periodic_x = 1;
  bad |= run("walls", w, 16, 14, 1, 1, 0);
  lbm_params m = deck(1, 0);
  m.mirror_north = m.mirror_south = 1;
  bad |= run("mirrors nnd", m, 15, 13, 0, 2, 0);
  lbm_params c = deck(1, 1);
  c.mirror_north = 1;
  bad |= run("plate clamp", c, 18, 12, 2, 0, 0);
  return bad;
}
