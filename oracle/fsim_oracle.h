/*
 * fsim_oracle.h -- CPU FP64 restatement of the fluidsim plasma-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * (libfsim_hip.so) and the timed CPU baseline.  Only tests/, the smoke() entry
 * and bench.py's cpu_baseline leg may load it.  The product never links it.
 *
 * It follows the reference Fortran line by line (citations in fsim_oracle.c):
 *   props.F90        constants and grid           (props.F90:24-47, 61-233)
 *   ptcl_props.F90   fits, SG flux, RK-Merson     (ptcl_props.F90:5-429)
 *   ptcl_lib.F90     RHS sweep, stage loop, dt    (ptcl_lib.F90:15-550)
 *   circ_lib.F90     electrode potential/circuit  (circ_lib.F90:19-223)
 *   sfc_lib.F90      dielectric wall charge       (sfc_lib.F90:20-90)
 *   lapl_lib.F90     semi-implicit Poisson        (lapl_lib.F90:57-636)
 *   main.F90         accept/reject, rotation      (main.F90:65-121)
 * All arrays use the reference's layout: (bx+2, by+2[, level]) FP64, column
 * major, i (x) fastest, one ghost layer, 1-based in the formulas.
 */
#ifndef FSIM_ORACLE_H
#define FSIM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Run configuration (all lengths in units of x0 = 1e-4 m, times in t0 = 1 us). */
typedef struct ocfg {
  int32_t nx, ny;            /* interior cells */
  int32_t grid_x, grid_y;    /* 0 uniform, 1 tanh (reference unif=F), 2 geometric */
  double ratio_x, ratio_y;   /* geometric ratio between neighbouring spacings */
  double l, w, ew, gap;      /* domain, electrode width, gap (nondimensional) */
  int32_t cyl;               /* axisymmetric r-z when 1 */
  int32_t rf;                /* 0 DC, 1 AC, 2 sin^3 pulse, 3 real pulse, 4 trapezoid */
  double p_torr, gam;        /* pressure [Torr], secondary emission */
  double vset;               /* source amplitude, nondimensional (volts / ph0) */
  double res_ohm;            /* ballast resistor [ohm] */
  double t_rise, t_width, t_fall;  /* trapezoid pulse (t0 units) */
  int32_t circuit;           /* 0 reference-literal, 1 active ballast */
  int32_t solver;            /* 0 banded LU, 1 BiCGStab+ILU(0), 2 Jacobi-PCG (sym.) */
  double dt0;
  double rtol;               /* Poisson relative residual */
  int32_t maxit;
  int32_t nthreads;          /* OpenMP threads (1 = serial) */
  int32_t rwall;             /* dielectric top wall with surface charge (props.F90:47) */
} ocfg;

typedef struct osim osim;

/* records: 10 doubles per ptcl_step attempt
 * [ts, t, dt_stage, err_cur, dt_next, accept_internal, lerr, iters, resid, t_m] */
#define OREC_LEN 10

osim* o_create(const ocfg* c);
void  o_destroy(osim* s);
int   o_dims(const osim* s, int32_t* out4);  /* bx, by, pitch, nlev */
int   o_get_grid(const osim* s, double* x, double* y, double* dx, double* dlx,
                 double* dy, double* dly, double* r, int32_t* type_x, int32_t* type_y);
/* field: 0 ni 1 ne 2 nt 3 nm 4 ph 5 ki 6 ke 7 kt 8 km ; level 1..3 (1..5 for k) */
int   o_get_field(const osim* s, int field, int level, double* out);
int   o_set_field(osim* s, int field, int level, const double* in);
void  o_init_uniform(osim* s);
void  o_init_gaussian(osim* s, uint64_t seed, double peak, double sigma, double noise);
/* scalars: [t, dt, err_prev, t_m, Vd_pl, Vd_mi, Vd_or, Ida, Idc, Vsrc, ts] */
void  o_get_scalars(const osim* s, double* out11);
void  o_set_scalars(osim* s, const double* in11);

void  o_rhs_sweep(osim* s, int stage);          /* ptclEqn over interior */
void  o_circ_step(osim* s, int stage);
void  o_sfc_step(osim* s, int stage);           /* wall charge (rwall) */
/* wall surface charge sig(bx+2), which: 0 sig_pl, 1 sig_mi, 2 sig_or */
int   o_get_sig(const osim* s, int which, double* out);
int   o_set_sig(osim* s, int which, const double* in);
void  o_rk_all(osim* s, int stage, double* nerr4);
int   o_ptcl_step(osim* s);                     /* <0 fatal */
int   o_lapl_assemble(osim* s, const double* ne, const double* ni, const double* nt,
                      double* b, double* A5);  /* A5: W,E,S,N,C per interior cell */
int   o_lapl_solve(osim* s, int* lerr, double* stats3);  /* iters, resid, seconds */
void  o_accept(osim* s);
void  o_reject(osim* s);
int   o_run(osim* s, int n_iters, double* times3);  /* ptcl, lapl, total seconds */

void  o_set_replay(osim* s, const double* recs, int n);
int   o_log_count(const osim* s);
int   o_get_log(const osim* s, double* out, int max);
void  o_clear_log(osim* s);

/* vectorised unit functions for known-answer tests */
void  o_consts(double p_torr, double gam, double* out);  /* see ORACLE_NCONST */
void  o_eval_fits(double p_torr, int n, const double* T, double* out12);
void  o_eval_flux(int n, const double* args7, double* out);
void  o_eval_dfdph(int n, const double* args7, double* out);
void  o_gaussian_field(const ocfg* c, uint64_t seed, double peak, double sigma,
                       double noise, int species, double* out);

#define ORACLE_NCONST 40

#ifdef __cplusplus
}
#endif
#endif
