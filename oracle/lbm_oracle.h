/*
 * lbm_oracle.h — CPU oracle for the D2Q13 thermal FD-LBM hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C, FP64, serial restatement of
 * the reference solver's numerics (reference CODE/C++/VS2010/simLB/simLB/
 * cpuLBM.cpp:283-450, 771-2335 and inlineFunctions.cpp:10-61), used as the
 * parity checker for the HIP path and as the reported CPU baseline. Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it. The product library never links or calls it.
 *
 * Layouts follow the reference: maps [x][y], distributions [x][y][i].
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include "../include/uqlbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* refineInput + setupVars + initialiseFunctions. NULL on bad input. */
orc* orc_create(const lbm_params* p, int32_t nx, int32_t ny,
                const int32_t* bnd, const lbm_props* props);
void orc_destroy(orc* o);

/* updateStepCount + runOne(res) per step; stops early when a step fails.
 * Returns the failure flag of the last step taken. */
int orc_step(orc* o, int32_t nsteps, int32_t want_res, double* res_out);

void orc_get_dist(const orc* o, double* f, double* h);
/* Overwrite f and h, then recompute the macroscopic fields of non-solid
 * cells (macroPropShort). */
void orc_set_dist(orc* o, const double* f, const double* h);
void orc_get_macro(const orc* o, double* rho, double* ux, double* uy,
                   double* p, double* T);
/* out[0..9] = dt, c, Tc, b, K, dx, dy, nsteps, step, tstep */
void orc_get_consts(const orc* o, double* out);

/* --- unit entry points (for pinning tests) --- */
void orc_equilibrium(orc* o, double rho, double ux, double uy, double T,
                     double* eqf, double* eqh, int32_t* failed);
double orc_weno5(double ee, const double* S);
double orc_nnd(double ee, const double* S);
double orc_minmod(double a, double b);
int32_t orc_index(const orc* o, int32_t i, int32_t pm, int32_t d, int32_t* mir);
void orc_mirror_dirs(int32_t i, int32_t mir_x, int32_t mir_y, int32_t* i_x,
                     int32_t* i_y);
void orc_stencil(const orc* o, const double* f, int32_t x, int32_t y,
                 int32_t i, int32_t len, double* Sx, double* Sy);
/* boundaryFunctions(f, h): face fluxes of every cell, [x][y][i][2]
 * (component 0 = x face, 1 = y face). */
void orc_boundary_fluxes(orc* o, const double* f, const double* h,
                         double* f_flux, double* h_flux);
/* cFLUX: net flux of direction i at (x,y) from a [x][y][i][2] flux field. */
double orc_cflux(const orc* o, const double* flux, int32_t x, int32_t y,
                 int32_t i);
/* Viscosity law (PY/codeCL.py:59-69, 987-990): 0 constant deck mu, 1 VHS
 * mu_ref (T/T_ref)^(0.5+param), 2 Sutherland mu_ref (T/T_ref)^1.5
 * (T_ref+S)/(T+S), S = param. Returns -1 on bad arguments. */
int orc_set_viscosity(orc* o, int32_t model, double mu_ref, double T_ref, double param);
/* Per-column dx[nx] and per-row dy[ny] (PY/codeCL.py:611-618, 676-679):
 * cFLUX with the cell's own dt/dx, dt/dy; dt from the smallest spacing.
 * Before the first step. Returns -1 on bad spacings or with solid cells. */
int orc_set_spacing(orc* o, const double* dx, const double* dy);
/* Seconds spent inside orc_step (steady clock around the step loop only). */
double orc_step_seconds(const orc* o);

#ifdef __cplusplus
}
#endif
#endif
