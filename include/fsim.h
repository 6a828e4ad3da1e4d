/*
 * fsim.h -- C ABI of libfsim_hip.so, the MI355X (gfx950) implementation of
 * fluidsim's plasma-transport hot path.
 *
 * The boundary replaces the Fortran module procedures that main.F90 calls.
 * Each entry point names the reference routine it stands in for.  Plain
 * pointers and sizes only; every function returns an int status:
 *     0  ok
 *     1  recoverable (Poisson solve not converged: caller takes the lerr=1
 *        reject path of main.F90:103-121)
 *    <0  fatal (see FS_E_* below); the reference stops on these.
 * One handle = one discharge instance, one HIP stream, one host thread.
 * Fields cross the boundary only in fs_set_field / fs_get_field.
 *
 * Array layout at the boundary is the reference's: (nx+2, ny+2) FP64,
 * column-major with i (x) fastest, one ghost layer (props.F90:16-21,
 * ptcl_lib.F90:17-20).  `ld` is the leading dimension of the host array
 * (>= nx+2).
 */
#ifndef FSIM_H
#define FSIM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_OK 0
#define FS_NOT_CONVERGED 1
#define FS_E_ARG (-1)        /* bad argument / configuration */
#define FS_E_NEGDENS (-2)    /* get_Te on a negative density (ptcl_props.F90:171-177) */
#define FS_E_NAN (-3)        /* NaN in a stage slope (ptcl_lib.F90:530-549) */
#define FS_E_DTMIN (-4)      /* dt <= 1.1e-12 (ptcl_lib.F90:68-72) */
#define FS_E_NANJ (-5)       /* NaN in dFdPh (lapl_lib.F90:632-635) */
#define FS_E_HIP (-6)        /* HIP runtime error */
#define FS_E_REPLAY (-7)     /* replay table exhausted */

/* Field ids for fs_set_field / fs_get_field (module arrays of ptcl_lib.F90:7-10
 * and lapl_lib.F90:14). */
#define FS_NI 0
#define FS_NE 1
#define FS_NT 2
#define FS_NM 3
#define FS_PH 4
#define FS_KI 5 /* stage slopes; only kept when debug_keep_k is set */
#define FS_KE 6
#define FS_KT 7
#define FS_KM 8
#define FS_SIG 9 /* dielectric wall charge sig(nx+2): level 1 sig_pl, 2 sig_mi, 3 sig_or
                    (sfc_lib.F90:7); handles with rwall only; ld is not used */

/* Grid description: the arguments of g_init (props.F90:52) plus the build's
 * geometric spacing.  Lengths are nondimensional (units of x0 = 1e-4 m). */
typedef struct fs_grid_desc {
  int32_t nx, ny;          /* interior cells (nx >= 3, ny >= 1; ny = 1: the reference's
                              one-dimensional mode, main.F90:29) */
  int32_t grid_x, grid_y;  /* 0 uniform (unif=T), 1 tanh (unif=F), 2 geometric */
  double ratio_x, ratio_y; /* geometric spacing ratio (grid kind 2) */
  double l, w, ew, gap;    /* as g_init */
  int32_t cyl;             /* r-z axisymmetric when 1 (props.F90:47) */
  int32_t rwall;           /* 1: the top wall (y = w) is a dielectric that collects
                              surface charge (props.F90:47, 192-198; sfc_lib.F90);
                              not for ensemble members */
} fs_grid_desc;

/* Run parameters: compile-time constants of the reference made runtime
 * (props.F90:41 p, ptcl_props.F90:25 gam), the circ_init/read_in inputs and
 * the solver controls. */
typedef struct fs_params {
  double p_torr;     /* gas pressure [Torr] */
  double gam;        /* secondary emission coefficient */
  double vset;       /* source amplitude, nondimensional (volts / ph0): circ_init(vl) */
  double res_ohm;    /* ballast resistor [ohm] (main.F90:286) */
  int32_t rf;        /* waveform: 0 DC, 1 AC, 2 sin^3, 3 real pulse (circ_lib.F90:26-50), 4 trapezoid */
  int32_t circuit;   /* 0 reference-literal circuit, 1 active RC ballast */
  double t_rise, t_width, t_fall; /* trapezoid pulse, t0 units */
  double dt0;        /* initial dt (main.F90:37) */
  double lapl_rtol;  /* Poisson: ||b - A d|| / ||b|| target (reference form) */
  int32_t lapl_maxit;
  int32_t lapl_chunk;     /* PCG iterations launched between convergence polls */
  int32_t debug_keep_k;   /* keep all five stage slopes for fs_get_field(FS_K*) */
  int32_t lapl_solver;    /* 0 Jacobi-PCG, 1 multigrid-preconditioned PCG, 2 learned
                             two-grid flexible CG (fs_learn_attach) */
  int32_t use_graphs;     /* replay the PCG iterations from a captured HIP graph */
  int32_t lapl_warm;      /* 1: a multigrid-PCG solve starts from the last two converged
                             corrections extrapolated to the new step, instead of zero
                             (the reference's PETSc KSP starts from zero); 2: from the
                             combination of the last three closest to the solution in
                             the operator's energy norm (Galerkin projection);
                             fs_set_field and fs_set_solver reset it */
  int32_t stage_tile;     /* tile height of the stage kernel (tiles of 64 x stage_tile
                             cells): 4, 8 or 16, or 0 to take the fastest of the three,
                             timed on the grid at creation (cached per grid shape).
                             Results do not depend on it. */
} fs_params;

/* Per-attempt controller record (replay and log), 10 doubles:
 * ts, t, dt_stage, err_cur, dt_next, accept_internal, lerr, iters, resid, t_m */
#define FS_REC_LEN 10

typedef struct fs_step_info {
  double err_cur, dt_stage, dt_next, t_m;
  double Vd_pl, Ida, Idc, Vsrc;
  int32_t attempts;      /* ptcl_step attempts incl. internal restarts */
  int32_t status;
} fs_step_info;

typedef struct fs_solve_stats {
  int32_t iters;
  int32_t lerr;
  double resid;          /* final relative residual, reference form */
  double bnorm;          /* ||b|| of the Newton residual */
  double seconds;        /* wall time of the solve (assembly + PCG) */
} fs_solve_stats;

typedef struct fs_handle fs_handle;

/* g_init + lapl_init + ptcl_init + circ_init (props.F90:52, lapl_lib.F90:22,
 * ptcl_lib.F90:15, circ_lib.F90:213).  Allocates all device state; fields
 * start at the ptcl_init uniform state. */
int fs_create(const fs_grid_desc* g, const fs_params* p, fs_handle** out);
int fs_destroy(fs_handle* h);

/* Grid metrics as g_init computes them (props.F90:96-218): x(nx+2), y(ny+2),
 * dx(nx+1), dlx(nx), dy(ny+1), dly(ny), r(ny+2); type_x/type_y (nx*ny, i fastest).
 * Any pointer may be NULL. */
int fs_get_grid(const fs_handle* h, double* x, double* y, double* dx, double* dlx, double* dy,
                double* dly, double* r, int32_t* type_x, int32_t* type_y);

/* Host <-> device copies of the module arrays. level 1..3 for densities
 * (1 new, 2 predictor, 3 base: ptcl_lib.F90:7-10), 1..5 for slopes, ignored for FS_PH.
 * Setting a density level writes that level only. */
int fs_set_field(fs_handle* h, int32_t field, int32_t level, const double* host, int32_t ld);
int fs_get_field(const fs_handle* h, int32_t field, int32_t level, double* host, int32_t ld);

/* ptcl_init uniform state (ptcl_lib.F90:22-29). */
int fs_init_uniform(fs_handle* h);
/* Gaussian blob + seeded multiplicative noise on ne/ni (all levels, ghosts
 * included): n = peak*exp(-r^2/(2 sigma^2))*(1+noise*u) + n_zero. */
int fs_init_gaussian(fs_handle* h, uint64_t seed, double peak, double sigma, double noise);

/* circ_init(Vset) (circ_lib.F90:213-223): the source amplitude Vmax and the
 * voltages Vsrc, Vd_pl, Vd_mi, Vd_or all take Vset (nondimensional). */
int fs_circ_init(fs_handle* h, double vset);

/* The stage kernel's tile height in use (4, 8 or 16; fs_params.stage_tile). */
int fs_stage_tile(const fs_handle* h);

/* Scalars: t, dt, err_prev, t_m, Vd_pl, Vd_mi, Vd_or, Ida, Idc, Vsrc, ts. */
int fs_get_scalars(const fs_handle* h, double* out11);
int fs_set_scalars(fs_handle* h, const double* in11);

/* ptcl_step(g, ph, Em) (ptcl_lib.F90:32-97): five RK-Merson stages with the
 * circuit update (circ_lib.F90:19-210), embedded error, dt controller and
 * internal restarts.  On return g%dt holds the next dt. */
int fs_ptcl_step(fs_handle* h, fs_step_info* info);

/* lapl_solve(g, ne1, ni1, nt1, sig, lerr) (lapl_lib.F90:57-121): assemble the
 * semi-implicit Newton system, solve it with matrix-free PCG on the
 * volume-symmetrised operator (Jacobi or aggregation-multigrid V-cycle
 * preconditioner, lapl_solver), and add the correction to ph.  *lerr = 1 when
 * the solve did not reach lapl_rtol within lapl_maxit (ph unchanged). */
int fs_lapl_solve(fs_handle* h, int32_t* lerr, fs_solve_stats* stats);

/* Accept / reject branches of the main loop (main.F90:83-121): accept
 * advances t by dt, rotates n3<-n2<-n1 (a pointer cycle) and the circuit
 * voltages; reject halves dt and restores n1 = n2 = n3. */
int fs_accept(fs_handle* h);
int fs_reject(fs_handle* h);

/* One pass of the main loop body (main.F90:74-121): [ptcl_step if ts>1],
 * lapl_solve, accept or reject.  n_iters passes; stops at the first fatal status. */
int fs_run(fs_handle* h, int32_t n_iters);

/* Replay table: records in the FS_REC_LEN layout; ptcl_step attempt k uses
 * dt_next = rec[k][4] and accept = rec[k][5] instead of its own controller
 * output.  n = 0 clears it. */
int fs_set_replay(fs_handle* h, const double* recs, int32_t n);
/* Attempt log (one record per ptcl_step attempt since the last clear). */
int fs_log_count(const fs_handle* h);
int fs_get_log(const fs_handle* h, double* out, int32_t max);
int fs_clear_log(fs_handle* h);

/* --- sub-steps exposed for kernel-level parity tests --- */
/* ptclEqn sweep for one stage (ptcl_lib.F90:44-48): stage slopes from level 2
 * and ph, no RK update.  Requires debug_keep_k. */
int fs_rhs_sweep(fs_handle* h, int32_t stage);
/* circ_step for one stage (circ_lib.F90:19-210) from level 2. */
int fs_circ_step(fs_handle* h, int32_t stage);
/* Poisson assembly from level 1 (lapl_lib.F90:65-87) in reference row form:
 * b[nx*ny] (laplEqn) and A5[nx*ny][5] = W,E,S,N,C (laplJ, scaled back from the
 * symmetrised storage). */
int fs_lapl_assemble(fs_handle* h, double* b, double* A5);

/* Constant table (40 entries, the order of the CPU oracle's table). */
int fs_constants(double p_torr, double gam, double* out40);
/* Device evaluation of the fits for known-answer tests: out[12*n]:
 * mue, mut, De, Dt (the Poisson assembly's forms), k_ex, k_ir, nu, k_sc, k_si,
 * beta, mue, De (the rates and the last two as the stage kernel evaluates them) */
int fs_eval_fits(double p_torr, int32_t n, const double* T, double* out);
/* Device SG flux / dFdPh: args7 = {E, dh, q, mu, D, nl, nr} per entry. */
int fs_eval_flux(int32_t n, const double* args7, double* out);
int fs_eval_dfdph(int32_t n, const double* args7, double* out);

/* fs_eval_flux (dfdph = 0) or fs_eval_dfdph (dfdph = 1) without any device
 * allocation.  No counterpart in the reference: this is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The device staging lives in a
 * fixed 2 MiB array per GPU, held until the process ends, so n is at most
 * 32768 (larger n: FS_E_ARG, nothing launched).  Every call on one GPU
 * shares that array: calls from threads of one process are serialised, and
 * results equal fs_eval_flux / fs_eval_dfdph bitwise. */
int fs_eval_flux_noalloc(int32_t n, const double* args7, double* out, int32_t dfdph);

/* Select the Poisson preconditioner of later fs_lapl_solve calls (0 Jacobi,
 * 1 multigrid), as fs_params.lapl_solver does at creation.  The reference
 * takes its Krylov method and preconditioner from the runtime options database
 * (KSPSetFromOptions, lapl_lib.F90:50-54); this is the equivalent switch. */
int fs_set_solver(fs_handle* h, int32_t solver);

/* Device timing of the last fs_run/fs_ptcl_step/fs_lapl_solve phases in ms:
 * [stage kernels, circuit+controller, assembly, pcg]. */
int fs_get_timers(const fs_handle* h, double* out4);
int fs_set_timing(fs_handle* h, int32_t on);

/* Average duration (ms) of n back-to-back launches of one hot-path kernel on
 * the handle's stream, measured with HIP events on the current state.  The
 * state is overwritten, so call it after the timed run.  which: 1..5 = the
 * fused stage kernel of that RK stage, 10 = Poisson assembly, 11/12 = the
 * Jacobi-PCG iteration and initialisation kernels, 13/14 = the multigrid down
 * and up kernels of the finest level (with the CG update and w = M u that the
 * multigrid-PCG fuses into them; the down kernel in its first-iteration form),
 * 15 = the attached network on the learned
 * level, 16 = one learned flexible-CG iteration, 17 = one multigrid-PCG
 * iteration (sequences of kernels, timed as a whole), 18 = the one-workgroup
 * coarse multigrid kernel, 20 + l / 40 + l = the down / up kernel of tile
 * level l of the multigrid-PCG iteration. */
int fs_time_kernel(fs_handle* h, int32_t which, int32_t n, double* avg_ms);

/* Device copy bandwidth: `reps` launches of a plain copy kernel moving
 * `bytes` (read) + `bytes` (write); *gbps = 2*bytes*reps / time.  The measured
 * HBM reference next to the kernel rooflines, and the calibration kernel for
 * the FETCH_SIZE / WRITE_SIZE counters. */
int fs_copy_bench(int64_t bytes, int32_t reps, double* gbps);

/* ---- ensembles --------------------------------------------------------------
 * No counterpart in the reference (one discharge per run).  n independent
 * discharges (grid_desc g[i], params p[i]; multigrid solver) advanced by
 * launches that each cover every member: a round gives each member one
 * ptcl_step attempt or, after an accepted attempt, its Poisson solve and the
 * accept/reject of the pass.  Each member has its own dt, controller,
 * replay table and solver convergence, and its fields equal those of the same
 * discharge run alone with fs_run, bitwise.  Members are ordinary handles
 * (initial state, fields, scalars and logs through the fs_* calls above),
 * owned by the ensemble. */
typedef struct fs_ensemble fs_ensemble;

int fs_ensemble_create(int32_t n, const fs_grid_desc* g, const fs_params* p, fs_ensemble** out);
int fs_ensemble_destroy(fs_ensemble* e);
fs_handle* fs_ensemble_member(fs_ensemble* e, int32_t i);
/* Every member advances `passes` main-loop passes (fs_run semantics). */
int fs_ensemble_run(fs_ensemble* e, int32_t passes);
/* The tile height of the ensemble's batched stage launches: its members'
 * common fs_params.stage_tile, or (0) the fastest of 4, 8, 16 timed over all
 * members at creation. */
int fs_ensemble_stage_tile(const fs_ensemble* e);

/* ---- learned Poisson preconditioner ------------------------------------
 * No counterpart in the reference (its Poisson solve is PETSc's KSP with the
 * default preconditioner, lapl_lib.F90:50-54, 104).  A 4-level convolutional
 * U-Net (16/32/48/64 channels, 3x3 kernels, BF16 weights and activations,
 * FP32 accumulation) is the coarse solver of the multigrid preconditioner
 * (red-black Gauss-Seidel smoothing, 2x2 aggregation) on the first level of at
 * most 64 x 64 cells, inside flexible CG: lapl_solver = 2.  Its first
 * application, to the right-hand side, is the solve's warm start.  It is
 * trained on the device with the A-norm energy loss over that level's
 * residuals harvested from the simulator's own solves.  Without an attached
 * network the level gets a zero correction (the method without the network). */
typedef struct fs_learn fs_learn;

/* Training-set capacity (samples) and the grid of the samples (H x W,
 * multiples of 8, normally 64 x 64; smaller grids are zero-padded). */
int fs_learn_create(int32_t capacity, int32_t H, int32_t W, uint64_t seed, fs_learn** out);
int fs_learn_destroy(fs_learn* n);
int64_t fs_learn_nparams(const fs_learn* n);
int fs_learn_get_weights(const fs_learn* n, float* out);   /* FP32 masters */
int fs_learn_set_weights(fs_learn* n, const float* in);
int fs_learn_samples(const fs_learn* n);
/* Solve the handle's current Poisson system with the multigrid-preconditioned
 * CG, one iteration at a time, appending the residual that reaches the learned
 * level in every cycle (with that level's operator) to the training set.
 * phi is left unchanged. */
int fs_learn_harvest(fs_learn* n, fs_handle* h, int32_t* added);
/* `steps` Adam steps on random batches of `batch` samples; loss[steps]. */
int fs_learn_train(fs_learn* n, int32_t steps, int32_t batch, float lr, uint64_t seed,
                   double* loss);
/* Loss and parameter gradient of a given batch (checks against a CPU model):
 * in5[B][H][W][5] = r/(s d) and the four coupling ratios, rhat[B][H][W],
 * op[B][3][H][W] = d, cx, cy.  grad[nparams], out[B][H][W] (e_hat) may be NULL. */
int fs_learn_loss_grad(fs_learn* n, int32_t B, int32_t H, int32_t W, const float* in5,
                       const float* rhat, const float* op, double* loss, float* grad, float* out);
/* Use the network in the handle's later solves when lapl_solver = 2 (NULL
 * detaches: zero correction on the learned level); the network must outlive
 * the attachment. */
int fs_learn_attach(fs_handle* h, fs_learn* n);

const char* fs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
