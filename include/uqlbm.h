/*
 * uqlbm.h — C ABI of the MI355X (gfx950) D2Q13 thermal finite-difference
 * lattice Boltzmann solver (double distribution f/h, BGK, WENO5/NND fluxes,
 * IMEX Runge-Kutta stepping).
 *
 * Plain C, POD structs, caller-owned host buffers, library-owned device state
 * behind opaque handles. Every entry point returns an lbm_status; on a
 * negative status lbm_last_error() describes the problem.
 *
 * Array layouts at this boundary are the reference solver's own:
 *   cell maps and macroscopic fields  [x][y]     index x*ny + y
 *   distributions f, h                [x][y][i]  index (x*ny + y)*13 + i
 * (reference CODE/C++/VS2010/simLB/simLB/macros.h:7-14).  The geometry map
 * `bnd` is expected already y-flipped, as cpuLBM::loadInput leaves it
 * (cpuLBM.cpp:254-266).
 */
#ifndef UQLBM_H
#define UQLBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_Q 13 /* D2Q13 velocities; f and h give 26 values per cell */

typedef enum lbm_status {
  LBM_OK = 0,
  LBM_FAILED = 1,          /* equilibrium mass check or non-finite state */
  LBM_ERR_ARG = -1,        /* bad argument */
  LBM_ERR_HIP = -2,        /* HIP runtime error */
  LBM_ERR_GEOMETRY = -3,   /* reserved: no geometry the reference accepts is refused
                              (thin solids and solids near mirror edges take
                              the physical-addressing path) */
  LBM_ERR_NOMEM = -4,      /* device allocation failed or budget exceeded */
  LBM_ERR_STATE = -5,      /* call out of order (e.g. step before init) */
  LBM_ERR_NOGPU = -6       /* no usable gfx950 device */
} lbm_status;

typedef enum lbm_precision {
  LBM_PREC_FP64 = 0,  /* FP64 state and arithmetic */
  LBM_PREC_MIXED = 1  /* FP32 state stored as deviations from a per-instance
                         reference equilibrium (at rest, mean fluid density
                         and temperature) with FP64 conserved-moment planes;
                         FP32 stencil reconstructions and equilibria, FP64
                         moments, stage updates, conservation sums and WENO
                         weight normalisation. Density, velocity and
                         temperature within 1e-6 relative L2 of the FP64
                         path over 5000 steps of the UQ decks. */
} lbm_precision;

/*
 * Deck scalars. Names and meaning follow the variables cpuLBM::loadInput
 * reads from the MAT deck (cpuLBM.cpp:113-199, cpuLBM.h:38-54); derived
 * values are filled in by the reference's refineInput/setupVars rules
 * (cpuLBM.cpp:283-450): dx from Lx when dx == 0, dy from Ly or dx,
 * rho_ref/p_ref/T_ref from the ideal-gas law, dt from CFL or dtau,
 * the step count from tt_tref, tt_time or steps.
 */
typedef struct lbm_params {
  int32_t rk_method;     /* RKMETHOD 0 RK1, 1 RK3 (IMEX-SSP2(3,2,2)), 2 EULER, 3 EULER_EQ */
  int32_t flux_method;   /* FMETHOD 0 NND, 1 WENO5 */
  double cfl;            /* CFL */
  double dtau;           /* dtau: dt = dtau*mu/p_ref when non-zero */
  double tt_tref;        /* run time in units of t_ref */
  double tt_time;        /* run time in seconds */
  int32_t steps;         /* step limit */
  double tol;            /* residual tolerance (0: off) */
  int32_t periodic_x, periodic_y;
  int32_t mirror_north, mirror_south, mirror_east, mirror_west;
  double gamma, R, mu, Pr;
  double dx, dy, Lx, Ly;
  double rho_ref, p_ref, T_ref;
  double Tc_Tref;
} lbm_params;

/* Property table: numProps rows of {density, pressure, temp, velocityX,
 * velocityY, cellType} (cpuLBM.h:101-108). cellType 0 = fluid, 1 = permanent
 * (frozen, still emits fluxes), 2 = solid wall. A zero in density, pressure
 * or temp is filled in from the other two (cpuLBM.cpp:333-357). */
typedef struct lbm_props {
  int32_t num_props;
  const double* density;
  const double* pressure;
  const double* temp;
  const double* velocity_x;
  const double* velocity_y;
  const int32_t* cell_type;
} lbm_props;

/* Derived run constants, read back after lbm_create. */
typedef struct lbm_info {
  int32_t nx, ny;
  int32_t nsteps;        /* step count from the deck (0: run to tolerance) */
  int32_t step;          /* steps taken so far */
  double dt;             /* time step */
  double c;              /* lattice speed sqrt(R*Tc) */
  double Tc;             /* characteristic temperature */
  double b, K;           /* b = 2/(gamma-1), K = b - 2 */
  double dx, dy;
  double tstep;          /* step*dt */
  int32_t precision;     /* lbm_precision */
  int32_t halo;          /* tile halo width the kernels use */
  size_t device_bytes;   /* device memory held by this solver */
} lbm_info;

typedef struct lbm_solver lbm_solver;
typedef struct lbm_batch lbm_batch;
typedef struct lbm_server lbm_server;

/* Last error message of the calling thread ("" when none). */
const char* lbm_last_error(void);

/* Number of usable gfx950 devices (0 when none); never fails. */
int32_t lbm_device_count(void);

/*
 * Create a solver: refineInput + setupVars semantics, geometry validation,
 * device allocation. Replaces cpuLBM::cpuLBM -> loadInput -> refineInput ->
 * setupVars (cpuLBM.cpp:7-20, 46-281, 283-450) with the deck handed over
 * as plain arrays instead of a MAT file.
 *   bnd : nx*ny property ids, [x][y], already y-flipped.
 */
int lbm_create(const lbm_params* params, int32_t nx, int32_t ny,
               const int32_t* bnd, const lbm_props* props,
               int32_t precision, lbm_solver** out);

/* Equilibrium initialisation from the property table plus macroscopic
 * fields; replaces cpuLBM::initialiseFunctions (cpuLBM.cpp:846-875). */
int lbm_init(lbm_solver* s);

/*
 * Advance nsteps steps of the deck's RK method; replaces the driver's
 * updateStepCount + runOne(res) pair (main.cpp:51-73, cpuLBM.cpp:771-775,
 * 2284-2335). Returns LBM_OK, or LBM_FAILED when a step trips the
 * equilibrium mass check |sum feq - rho| > 1e-6 (cpuLBM.cpp:825-843) or
 * produces a non-finite density; stepping stops after the failing step.
 * When want_residual != 0, *residual_out receives the RMS change of the
 * temperature field over the last step (cpuLBM.cpp:2322-2331).
 */
int lbm_step(lbm_solver* s, int32_t nsteps, int32_t want_residual,
             double* residual_out);

/*
 * lbm_step without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. Same arguments, results and
 * status as lbm_step, bit for bit; the kernels are launched directly
 * (no graph capture or instantiation, no events) and the residual's
 * per-row sums live in a fixed device array of the library instead of a
 * buffer the solver allocates on first use. Limit: with want_residual != 0
 * the grid may have at most 32768 rows (ny); larger grids get LBM_ERR_ARG.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (other host threads, other streams) race and must be serialised by
 * the caller.
 */
int lbm_step_noalloc(lbm_solver* s, int32_t nsteps, int32_t want_residual,
                     double* residual_out);

/* Macroscopic output in [x][y] layout (the public rho_G, u_G, p_G, T_G of
 * cpuLBM.h:93-96, split into ux/uy). Any pointer may be NULL. */
int lbm_get_macro(lbm_solver* s, double* rho, double* ux, double* uy,
                  double* p, double* T);

/* Distributions in [x][y][i] layout, FP64 regardless of precision
 * (the public f_G, h_G of cpuLBM.h:90-91). Either pointer may be NULL. */
int lbm_get_dist(lbm_solver* s, double* f, double* h);

/* Overwrite the distributions ([x][y][i]) and recompute the macroscopic
 * fields. No counterpart in the reference (which has no resume path);
 * used for checkpoint/resume and for injecting states in tests. */
int lbm_set_dist(lbm_solver* s, const double* f, const double* h);

/* Temperature-dependent viscosity of the reference's OpenCL front end
 * (CODE/Python/simLB/codeCL.py:59-69, 987-990), replacing the deck's
 * constant mu in the relaxation times tf = mu(T)/p of every later step:
 * model 0 constant (deck mu), 1 VHS mu_ref (T/T_ref)^(0.5 + param),
 * 2 Sutherland mu_ref (T/T_ref)^1.5 (T_ref + S)/(T + S) with S = param.
 * T is the cell temperature in deck units. */
int lbm_set_viscosity(lbm_solver* s, int32_t model, double mu_ref, double T_ref, double param);

/* Non-uniform grid of the reference's OpenCL front end
 * (CODE/Python/simLB/codeCL.py:611-618, 676-679): dx[nx] per column and
 * dy[ny] per row. Face fluxes are weighted by their face length and the net
 * flux divided by the cell area, i.e. the cell's own dt/dx[x] and dt/dy[y]
 * in cFLUX; dt is taken from the smallest spacing
 * (CODE/Python/simLB/LBM.py:216) and the step count re-derived. Walls keep
 * the C++ closure (CODE/C++/VS2010/simLB/simLB/cpuLBM.cpp:1502-1590) with
 * the front end's face-length weights, which reduces to the wall cell's
 * own dx, dy in its mass balance (the OpenCL closure's equilibrium-based
 * outgoing flux, codeCL.py:685-753, is not adopted). Before the first step
 * (LBM_ERR_STATE after); LBM_ERR_ARG for non-positive spacings. Totals
 * (lbm_get_totals) stay unweighted cell sums; the conserved quantities are
 * the sums weighted by dx[x] dy[y]. */
int lbm_set_spacing(lbm_solver* s, const double* dx, const double* dy);

/* Conserved totals over the fluid cells (cellType 0), FP64 sums:
 * out[0] = sum f (mass), out[1] = sum f e_x, out[2] = sum f e_y (momentum),
 * out[3] = sum h (total energy), e in physical units (lattice x c). The
 * reference checks conservation only through its per-cell equilibrium mass
 * check (cpuLBM.cpp:825-843); these totals let a caller verify that mass,
 * momentum and energy are conserved on closed (periodic) domains. */
int lbm_get_totals(lbm_solver* s, double* out);

/* Grid, derived constants (dt, c, Tc, b, K; cpuLBM's setupVars values,
 * CPP/cpuLBM.cpp:360-450), step count and device footprint of a solver. */
int lbm_get_info(const lbm_solver* s, lbm_info* out);

/* Frees the solver; one still held by a batch is freed when its last batch
 * is destroyed (the batch keeps stepping it until then). */
void lbm_destroy(lbm_solver* s);

/*
 * Batched stepping: every member advances together, one launch per RK stage
 * for the whole set (members may differ in size and parameters; all share
 * rk_method and precision). Per-member arithmetic does not depend on the
 * other members: a member stepped in a batch matches the same member stepped
 * alone bit for bit. The reference has no batched path (it steps one
 * cpuLBM); this is the ensemble form of lbm_step.
 */
int lbm_batch_create(lbm_solver* const* members, int32_t n, lbm_batch** out);
int lbm_batch_step(lbm_batch* b, int32_t nsteps);
void lbm_batch_destroy(lbm_batch* b);

/*
 * Device time of the last lbm_step / lbm_batch_step call: total elapsed
 * milliseconds between HIP events recorded on the solver's stream, and the
 * per-kernel average durations when timing was enabled with
 * lbm_set_kernel_timing (events around every launch).
 */
typedef struct lbm_timing {
  double total_ms;
  int32_t n_kernels;          /* entries used below */
  char names[8][32];
  double avg_ms[8];
  int64_t launches[8];
} lbm_timing;

/*
 * Flux tile shapes and the autotuner. The stencil kernels work on tiles of
 * own cells plus a 3-cell halo cross; shape 0 = 15 x 16 (256 threads),
 * 1 = 31 x 15 (512), 2 = 15 x 32 (512), 3 = 39 x 12 (512), 4 = 23 x 20 (512). lbm_batch_autotune times `steps`
 * steps of the batch with each shape, keeps the fastest (the lowest-numbered
 * shape within 0.5 % of the fastest, so that near-ties resolve the same way
 * from run to run) for the batch and
 * as the default of solvers, batches and servers created afterwards with
 * the batch's precision, restores the members' states (tuning changes no
 * result) and reports the shape and each shape's milliseconds per step
 * (ms_per_step_out: LBM_FLUX_SHAPES entries, may be NULL). Results are the
 * same for every shape. No counterpart in the reference (its stencil loops
 * are per cell).
 */
#define LBM_FLUX_SHAPES 5
int lbm_set_flux_shape(int32_t precision, int32_t shape);
/* Own-cell width and height of a flux tile shape. */
int lbm_flux_shape_dims(int32_t shape, int32_t* w, int32_t* h);
int lbm_batch_autotune(lbm_batch* b, int32_t steps, int32_t* shape_out, double* ms_per_step_out);

/* Per-kernel timing of later step calls (process-wide; 0 turns it off).
 * While on, steps launch kernel by kernel instead of replaying graphs.
 * lbm_get_timing copies the record of the last step call. No counterpart
 * in the reference. */
int lbm_set_kernel_timing(int32_t enable);
int lbm_get_timing(lbm_timing* out);

/*
 * Step sequences (lbm_step, lbm_batch_step, server ticks) replay captured
 * HIP graphs of one and two steps, recaptured when a group's work lists
 * change; enable = 0 launches every kernel directly instead (also the case
 * while kernel timing is on). Results are identical either way. No
 * counterpart in the reference (whose step is a host loop); on by default.
 */
int lbm_set_graphs(int32_t enable);

/*
 * Flux stages of instances whose cells all take the logical stencil
 * addressing can run on the strip kernels (one wavefront per 58-column,
 * 32-row strip, marched row by row) instead of the flux tile kernels, except
 * the moment-summing stage-3 flux of LBM_PREC_MIXED. mode 0: every flux
 * stage on the tile kernels; 1 (default): strips for launch groups with at
 * least 2048 strip items (enough to fill the GPU; e.g. a 4096^2 instance or
 * the 256-member ensemble), tiles for smaller groups; 2: strips wherever an
 * instance can take them. Results are bitwise equal either way (FP64, and
 * LBM_PREC_MIXED instances near their reference equilibrium; mixed
 * instances far from it compute their strip stencils in FP64, the tile
 * kernels in FP32). Applies to step calls made afterwards; LBM_ERR_ARG for
 * another mode. No counterpart in the reference.
 */
int lbm_set_strips(int32_t mode);

/*
 * Device buffers of destroyed solvers are kept (up to 32 GiB per process)
 * for later solvers of the same sizes, so that a served request stream
 * creates and destroys solvers without hipMalloc and hipFree; a solver
 * whose allocation finds the device full first returns the kept buffers of
 * its device. lbm_release_cached returns every kept buffer now. No
 * counterpart in the reference.
 */
void lbm_release_cached(void);

/*
 * Continuous-batching server. Requests are queued host-side; a request is
 * admitted (device state allocated and initialised) when a slot is free and
 * its device footprint fits within hbm_budget_bytes. Each tick advances
 * every admitted solve by up to tick_steps steps in batched launches and
 * retires the ones that reached their own step count. A request whose deck
 * has a residual tolerance (lbm_params.tol > 0) also retires once it meets
 * the reference driver's stop rule (main.cpp:54-110: from step 21 on, the
 * residual normalised by the largest residual of steps 2..20 is <= tol),
 * tested at the end of each batched round; with nsteps = 0 it runs until
 * then. Results of a request do not depend on which other requests shared
 * its ticks.
 */
typedef struct lbm_request_result {
  int64_t tag;
  int32_t status;       /* lbm_status of the solve */
  int32_t steps_done;
  double submit_s;      /* host clock (seconds) at submission */
  double admit_s;       /* host clock at admission */
  double finish_s;      /* host clock at completion */
  int64_t cells;        /* nx*ny */
  double residual;      /* normalised residual at completion (requests
                           with a tolerance; 0 otherwise) */
} lbm_request_result;

int lbm_server_create(int32_t precision, int32_t rk_method,
                      size_t hbm_budget_bytes, int32_t max_slots,
                      lbm_server** out);
int lbm_server_submit(lbm_server* sv, const lbm_params* params, int32_t nx,
                      int32_t ny, const int32_t* bnd, const lbm_props* props,
                      int32_t nsteps, int64_t tag);
/* Admit, advance one tick, retire. *active_out receives the number of
 * solves resident after the tick, *queued_out the number still waiting. */
int lbm_server_tick(lbm_server* sv, int32_t tick_steps, int32_t* active_out,
                    int32_t* queued_out);
/* Completed requests since the last poll (at most max). Returns the count
 * written, or a negative status. */
int lbm_server_poll(lbm_server* sv, lbm_request_result* out, int32_t max);
/* With enable != 0, requests that complete from now on keep their final
 * macroscopic fields in host memory (until the server is destroyed). */
int lbm_server_keep_results(lbm_server* sv, int32_t enable);
/* Copy the kept final fields of completed request `tag`, [x][y]; any
 * pointer may be NULL. LBM_ERR_ARG when nothing was kept for the tag. */
int lbm_server_result_macro(lbm_server* sv, int64_t tag, double* rho,
                            double* ux, double* uy, double* p, double* T);
/* Lattice updates done so far and device bytes held by resident solves;
 * either pointer may be NULL. */
int lbm_server_stats(lbm_server* sv, int64_t* lups_done, size_t* hbm_in_use);
/* Current host clock in seconds (the server's time base). */
double lbm_clock_s(void);
void lbm_server_destroy(lbm_server* sv);

#ifdef __cplusplus
}
#endif

#endif /* UQLBM_H */
