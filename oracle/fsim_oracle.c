/*
 * fsim_oracle.c -- CPU FP64 restatement of the fluidsim plasma-transport path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libfsim_hip.so and the timed CPU
 * baseline.  Nothing in the product links or loads this file.
 *
 * Every routine follows one reference routine statement by statement, keeping
 * its operation order (left-to-right Fortran evaluation, no reassociation, no
 * FMA: build with -ffp-contract=off), its float32-widened literals and its
 * quirks.  Deliberate differences from the reference are marked "DIFF:".
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * of tests/test_oracle_pins.py.
 */
#include "fsim_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* constants: props.F90:24-44, ptcl_props.F90:5-25                           */
/* ------------------------------------------------------------------------ */
typedef struct oconsts {
  double pi, eps0, mu0, c0, e, kb, me;
  double x0, ph0, t0, wp02, w2;
  double Tg, p, ninf, n_zero, n_init;
  double mi, mui, Ti, vi, Di, mm, Dm, k_r, k_mp, k_2q, k_3q;
  double H_ir, H_ex, H_si, H_sc, gam;
  double log5119, log5667, log75; /* float32 breakpoints (ptcl_props.F90:213,337,366) */
  double ftol;                    /* get_Flux tol = 1e-12 (float literal, ptcl_props.F90:148) */
  double x0_3, x0_2, x0_6, Cap, kTe, nmin_t;
} oconsts;

#define F32(v) ((double)(v##f))

/* x0**6 correctly rounded, as gfortran folds the parameter expression
 * (ptcl_props.F90:18): x0 has a 24-bit significand, so x0**4 is exact in
 * binary128 and one extra rounding gives the double result. */
static double pow6_exact(double x) {
  __float128 q = (__float128)x;
  __float128 q2 = q * q, q4 = q2 * q2, q6 = q4 * q2;
  return (double)q6;
}

/* float32 log of a float32 literal, as gfortran folds log(5.119) */
static double flog32(float v) { return (double)(float)log((double)v); }

static void consts_init(oconsts* k, double p, double gam) {
  k->pi = 4.0 * atan(1.0);                       /* props.F90:24 */
  k->eps0 = F32(8.85418782e-12);                 /* props.F90:25 */
  k->mu0 = 4.0 * k->pi * F32(1e-7);              /* props.F90:26 */
  k->c0 = 1.0 / sqrt(k->eps0 * k->mu0);          /* props.F90:27 */
  k->e = F32(1.60217646e-19);                    /* props.F90:28 */
  k->kb = F32(1.3806503e-23);                    /* props.F90:29 */
  k->me = 9.10938188e-31;                        /* props.F90:30 */
  k->x0 = F32(1e-4);                             /* props.F90:33 */
  k->x0_2 = k->x0 * k->x0;                       /* exact (24-bit operand) */
  k->x0_3 = k->x0_2 * k->x0;                     /* single rounding = correctly rounded */
  k->x0_6 = pow6_exact(k->x0);
  k->ph0 = k->e / (k->eps0 * k->x0);             /* props.F90:34 */
  k->t0 = F32(1e-6);                             /* props.F90:35 */
  k->wp02 = (k->t0 * k->t0) * (k->e * k->e) / (k->eps0 * k->x0_3 * k->me); /* props.F90:36 */
  {
    double tw = k->t0 * F32(14e9);               /* props.F90:37 */
    k->w2 = tw * tw;
  }
  k->Tg = 300.0;                                 /* props.F90:40 */
  k->p = p;                                      /* props.F90:41 (runtime here) */
  k->ninf = p * 101325.0 / 760.0 / k->kb / k->Tg * k->x0_3;  /* props.F90:42 */
  k->n_zero = F32(1e8) * k->x0_3;                /* props.F90:43 */
  k->n_init = F32(1e11) * k->x0_3;               /* props.F90:44 */
  k->mi = 1.67262178e-27 * 39.948;               /* ptcl_props.F90:6 */
  k->mui = 1.45e3 / p * 1e-4 * k->ph0 * k->t0 / k->x0_2;       /* ptcl_props.F90:7 */
  k->Ti = k->Tg;                                 /* ptcl_props.F90:8 */
  k->vi = sqrt((8.0 * k->kb * k->Ti) / (k->pi * k->mi)) * k->t0 / k->x0;  /* :9 */
  k->Di = k->mui * (k->kb * k->Ti / k->e) / k->ph0;            /* ptcl_props.F90:10 */
  k->mm = k->mi;                                 /* ptcl_props.F90:13 */
  k->Dm = 2.42e18 * k->x0 / k->ninf / 1e2 * k->t0;             /* ptcl_props.F90:14 */
  k->k_r = 2e-7 / 1e6 / k->x0_3 * k->t0;         /* ptcl_props.F90:15 */
  k->k_mp = 6.2e-10 / 1.0e6 / k->x0_3 * k->t0;   /* ptcl_props.F90:16 */
  k->k_2q = 3e-15 / 1e6 / k->x0_3 * k->t0;       /* ptcl_props.F90:17 */
  k->k_3q = 1.1e-31 / 1.0e12 / k->x0_6 * k->t0;  /* ptcl_props.F90:18 */
  k->H_ir = 15.80 / k->ph0;                      /* ptcl_props.F90:21-24 */
  k->H_ex = 11.56 / k->ph0;
  k->H_si = 4.14 / k->ph0;
  k->H_sc = -11.56 / k->ph0;
  k->gam = gam;                                  /* ptcl_props.F90:25 (runtime here) */
  k->log5119 = flog32(5.119f);
  k->log5667 = flog32(5.667f);
  k->log75 = flog32(7.5f);
  k->ftol = F32(1e-12);
  k->Cap = 1e-11 * k->ph0 / k->e;                /* circ_lib.F90:221 */
  k->kTe = k->Tg * k->kb / k->e / k->ph0;        /* ptcl_lib.F90:28,508 */
  k->nmin_t = k->n_zero / k->ph0 / 1e2;          /* ptcl_lib.F90:56 */
}

/* order of the exported constant table; libfsim_hip exports the same table */
static void consts_table(const oconsts* k, double* o) {
  const double v[ORACLE_NCONST] = {
      k->pi,   k->eps0, k->mu0,  k->e,    k->kb,   k->me,    k->x0,   k->ph0,
      k->t0,   k->wp02, k->w2,   k->Tg,   k->p,    k->ninf,  k->n_zero, k->n_init,
      k->mi,   k->mui,  k->vi,   k->Di,   k->Dm,   k->k_r,   k->k_mp, k->k_2q,
      k->k_3q, k->H_ir, k->H_ex, k->H_si, k->H_sc, k->gam,   k->log5119, k->log5667,
      k->log75, k->ftol, k->Cap, k->x0_3, k->c0,   k->kTe,   k->nmin_t, k->x0_6};
  memcpy(o, v, sizeof(v));
}

/* ------------------------------------------------------------------------ */
/* integer powers: gfortran expands x**n (integer n) with the optimal power  */
/* tree x**n = x**(n-t) * x**t, t = table[n] (Knuth, as in gfortran).        */
/* Real exponents x**3. etc. are pow() calls (ptcl_props.F90:392,410,427);   */
/* x**2. folds to x*x.                                                       */
/* ------------------------------------------------------------------------ */
static const int powi_tab[12] = {0, 1, 1, 2, 2, 3, 3, 4, 4, 6, 5, 6};
static double ipow(double x, int n) {
  if (n == 0) return 1.0;
  if (n == 1) return x;
  int t = powi_tab[n];
  return ipow(x, n - t) * ipow(x, t);
}
static double rpow(double x, double y) { return y == 2.0 ? x * x : pow(x, y); }

/* ------------------------------------------------------------------------ */
/* fits: ptcl_props.F90:167-429                                              */
/* ------------------------------------------------------------------------ */
static int g_neg_density = 0; /* set when get_Te meets a negative density */

static double get_Te(double nte, double ne) { /* ptcl_props.F90:167-180 */
  if ((nte >= 0.0) && (ne >= 0.0)) return nte / ne;
  g_neg_density = 1; /* reference: "Error, negative density. Stop." */
  return 1e-8;
}

static double get_beta(const oconsts* k, double T) { /* ptcl_props.F90:182-192 */
  const double a = 11.604505e3;
  double v = T * k->ph0 / F32(1.5) * a;
  double x = log10(v < 5e4 ? v : 5e4);
  return pow(10.0, 0.944 - 0.665 * (x - F32(2.48))) * 1e-13 * k->t0 / k->x0_3;
}

static double clamp_log(double lo, double hi, double v) {
  double m = v < hi ? v : hi;  /* min(hi, v) */
  return log(lo > m ? lo : m); /* max(lo, .) */
}

static double get_mue(const oconsts* k, double T) { /* ptcl_props.F90:194-221 */
  const double a1 = F32(58.1133016145), b1 = F32(-0.984082217962), c1 = F32(-0.164770900119),
               d1 = F32(-0.142058042584), f1 = F32(-0.0637079234081), g1 = F32(0.0436642742558),
               a2 = F32(70.7846754548), b2 = F32(-22.7558138237), c2 = F32(12.8366493242),
               d2 = F32(-3.57242763244), f2 = F32(0.486276623664), g2 = F32(-0.0259045697422);
  double x = clamp_log(2.34e-1, 1.57e2, T * k->ph0);
  double s;
  if (x < k->log5119)
    s = a1 + b1 * x + c1 * ipow(x, 2) + d1 * ipow(x, 3) + f1 * ipow(x, 4) + g1 * ipow(x, 5);
  else
    s = a2 + b2 * x + c2 * ipow(x, 2) + d2 * ipow(x, 3) + f2 * ipow(x, 4) + g2 * ipow(x, 5);
  return exp(s) * k->x0 / k->ninf * k->t0 * k->ph0;
}

static double get_mut(const oconsts* k, double T) { /* ptcl_props.F90:223-250 */
  const double a1 = F32(58.1354622038), b1 = F32(-1.29918391109), c1 = F32(-0.106559658509),
               d1 = F32(-0.0202265418394), f1 = F32(-0.0286667333915), g1 = F32(0.022457210094),
               a2 = F32(68.2627663433), b2 = F32(-19.3524000831), c2 = F32(11.3251980583),
               d2 = F32(-3.24070386509), f2 = F32(0.449848156054), g2 = F32(-0.0242585762405);
  double x = clamp_log(2.34e-1, 1.57e2, T * k->ph0);
  double s;
  if (x < k->log5119)
    s = a1 + b1 * x + c1 * ipow(x, 2) + d1 * ipow(x, 3) + f1 * ipow(x, 4) + g1 * ipow(x, 5);
  else
    s = a2 + b2 * x + c2 * ipow(x, 2) + d2 * ipow(x, 3) + f2 * ipow(x, 4) + g2 * ipow(x, 5);
  return exp(s) * k->x0 / k->ninf * k->t0 * k->ph0;
}

static double get_mui_E(const oconsts* k, double E) { /* ptcl_props.F90:252-266 (never called) */
  const double a = F32(0.476914771378), b = F32(-0.227345106588), c = F32(0.0530705372631),
               d = F32(-0.00521915548354), f = F32(0.000176442914616);
  double v = E * k->ph0 / k->x0;
  double m = v > 1e2 ? v : 1e2;
  double x = log(m < 1e5 ? m : 1e5);
  return (a + b * x + c * ipow(x, 2) + d * ipow(x, 3) + f * ipow(x, 4)) * k->ph0 * k->t0 / k->x0_2;
}

static double get_De(const oconsts* k, double T) { /* ptcl_props.F90:268-291 */
  const double a = F32(57.7067018813), b = F32(-0.0699892875381), c = F32(-0.117645585949),
               d = F32(0.105390295278), f = F32(-0.102862612604), g = F32(-0.0469171521686),
               h = F32(0.0584908312121), i = F32(0.000578601715687), j = F32(-0.0122860884883),
               kk = F32(0.00426793748856), l = F32(-0.000590000082557), m = F32(3.00706533201e-05);
  double x = clamp_log(2.34e-1, 1.57e2, T * k->ph0);
  double s = a + b * x + c * ipow(x, 2) + d * ipow(x, 3) + f * ipow(x, 4) + g * ipow(x, 5) +
             h * ipow(x, 6) + i * ipow(x, 7) + j * ipow(x, 8) + kk * ipow(x, 9) +
             l * ipow(x, 10) + m * ipow(x, 11);
  return exp(s) * k->x0 / k->ninf * k->t0;
}

static double get_Dt(const oconsts* k, double T) { /* ptcl_props.F90:293-316 */
  const double a = F32(57.7214952841), b = F32(-0.353731464348), c = F32(-0.00505156731795),
               d = F32(0.161173720584), f = F32(-0.196610345869), g = F32(-0.0719115643218),
               h = F32(0.11874085868), i = F32(-0.00669712724784), j = F32(-0.0236445308504),
               kk = F32(0.00917326671764), l = F32(-0.00135453096483), m = F32(7.26379684461e-05);
  double x = clamp_log(2.34e-1, 1.57e2, T * k->ph0);
  double s = a + b * x + c * ipow(x, 2) + d * ipow(x, 3) + f * ipow(x, 4) + g * ipow(x, 5) +
             h * ipow(x, 6) + i * ipow(x, 7) + j * ipow(x, 8) + kk * ipow(x, 9) +
             l * ipow(x, 10) + m * ipow(x, 11);
  return exp(s) * k->x0 / k->ninf * k->t0;
}

static double get_k_ex(const oconsts* k, double T) { /* ptcl_props.F90:318-345 */
  const double a1 = F32(-54.4513453969), b1 = F32(17.1472529739), c1 = F32(-1.05188982824),
               d1 = F32(-10.5053010711), f1 = F32(7.51502551486), g1 = F32(-1.44763942525),
               a2 = F32(-150.266369415), b2 = F32(158.67012941), c2 = F32(-85.5163470769),
               d2 = F32(23.0655081891), f2 = F32(-3.0908688844), g2 = F32(0.16403226902);
  double x = clamp_log(9.611e-1, 1.57e2, T * k->ph0);
  double s;
  if (x < k->log5667)
    s = a1 + b1 * x + c1 * ipow(x, 2) + d1 * ipow(x, 3) + f1 * ipow(x, 4) + g1 * ipow(x, 5);
  else
    s = a2 + b2 * x + c2 * ipow(x, 2) + d2 * ipow(x, 3) + f2 * ipow(x, 4) + g2 * ipow(x, 5);
  return exp(s) * k->t0 / k->x0_3;
}

static double get_k_ir(const oconsts* k, double T) { /* ptcl_props.F90:347-374 */
  const double a1 = F32(-73.27080258), b1 = F32(-30.1312514721), c1 = F32(138.743842326),
               d1 = F32(-161.378167125), f1 = F32(81.9325486193), g1 = F32(-14.9080865672),
               a2 = F32(-168.925304837), b2 = F32(177.113452524), c2 = F32(-92.3047166553),
               d2 = F32(24.1814225482), f2 = F32(-3.15463258104), g2 = F32(0.16331188192);
  double x = clamp_log(1.37, 1.57e2, T * k->ph0);
  double s;
  if (x < k->log75)
    s = a1 + b1 * x + c1 * ipow(x, 2) + d1 * ipow(x, 3) + f1 * ipow(x, 4) + g1 * ipow(x, 5);
  else
    s = a2 + b2 * x + c2 * ipow(x, 2) + d2 * ipow(x, 3) + f2 * ipow(x, 4) + g2 * ipow(x, 5);
  return exp(s) * k->t0 / k->x0_3;
}

static double get_nu(const oconsts* k, double T) { /* ptcl_props.F90:376-395 */
  const double a = F32(-32.275912575), b = F32(1.45173283977), c = F32(0.00936933121094),
               d = F32(0.129397015353), f = F32(-0.0414865809044), g = F32(-0.0582934303409),
               h = F32(0.0309832277826), i = F32(-0.00542014733763), j = F32(0.000325615321708);
  double x = clamp_log(2.34e-1, 1.57e2, T * k->ph0);
  double s = a + b * x + c * rpow(x, 2.) + d * rpow(x, 3.) + f * rpow(x, 4.) + g * rpow(x, 5.) +
             h * rpow(x, 6.) + i * ipow(x, 7) + j * rpow(x, 8.);
  return exp(s) / k->x0_3 * k->ninf * k->t0;
}

static double get_k_sc(const oconsts* k, double T) { /* ptcl_props.F90:397-413 */
  const double a = F32(-21.4827864151), b = F32(0.457356923276), c = F32(-0.555439231606),
               d = F32(1.27257798891), f = F32(-0.67840685073), g = F32(0.10591014464);
  double v = T * k->ph0;
  double m = v > 5e-1 ? v : 5e-1;
  double x = log(m < 16.0 ? m : 16.0);
  double s = a + b * x + c * rpow(x, 2.) + d * rpow(x, 3.) + f * rpow(x, 4.) + g * rpow(x, 5.);
  return exp(s) / 1.0e6 / k->x0_3 * k->t0;
}

static double get_k_si(const oconsts* k, double T) { /* ptcl_props.F90:415-429 */
  const double a = F32(-43.1347385848), b = F32(43.9905424566), c = F32(-28.1169537586),
               d = F32(8.28853856817), f = F32(-0.931626144207);
  double v = T * k->ph0;
  double m = v > 5e-1 ? v : 5e-1;
  double x = log(m < 16.0 ? m : 16.0);
  double s = a + b * x + c * rpow(x, 2.) + d * rpow(x, 3.) + f * rpow(x, 4.);
  return exp(s) / 1.0e6 / k->x0_3 * k->t0;
}

/* Scharfetter-Gummel face flux: ptcl_props.F90:142-165 */
static double get_flux(const oconsts* k, double E, double dh, int q, double mu, double D,
                       double nl, double nr) {
  double tol = k->ftol;
  double v = (double)q * mu * E;
  double arg = v * dh / D;
  if (fabs((double)q * E) < tol) return D * (nl - nr) / dh;
  /* DIFF: where 1 - exp(-|arg|) rounds to 0 (|arg| below 2^-53, a field of
   * rounding size that passes the |qE| < tol test) the reference divides by
   * zero and stops on the NaN (ptcl_lib.F90:530-549); here the a -> 0 limit,
   * Fick's flux, as in the device kernels */
  if (arg > 0) {
    double ex = exp(fmax(-arg, -100.0));
    if (1.0 - ex == 0.0) return D * (nl - nr) / dh;
    return v * (nl - nr * ex) / (1.0 - ex);
  } else {
    double ex = exp(fmax(arg, -100.0));
    if (1.0 - ex == 0.0) return D * (nl - nr) / dh;
    return v * (nr - nl * ex) / (1.0 - ex);
  }
}

/* derivative of the SG flux w.r.t. the drift velocity: lapl_lib.F90:611-636 */
static int g_nan_dfdph = 0;
static double dfdph(double E_pl, double E_mi, int q, double mu, double D, double dx, double n1,
                    double n2) {
  (void)E_pl; /* only printed by the reference's NaN message */
  double tol = 1e-12, df;
  double v_mi = (double)q * mu * E_mi;
  double a_mi = v_mi * dx / D;
  if (fabs(v_mi) < tol) {
    df = 0.0;
  } else if (a_mi > 0) {
    double ex = exp(-a_mi), om = 1.0 - ex;
    df = (n1 - n2 * ex) / om + v_mi * ex * (n2 - n1) / (om * om) * dx / D;
    if (om == 0.0) df = 5e-1 * (n1 + n2); /* DIFF: the a -> 0 limit (see get_flux) */
  } else {
    double ex = exp(a_mi), om = 1.0 - ex;
    df = (n2 - n1 * ex) / om + v_mi * ex * (n2 - n1) / (om * om) * dx / D;
    if (om == 0.0) df = 5e-1 * (n1 + n2);
  }
  if (isnan(df)) g_nan_dfdph = 1; /* reference: "flx is NaN" + stop */
  return df;
}

/* ------------------------------------------------------------------------ */
/* simulation state                                                          */
/* ------------------------------------------------------------------------ */
struct osim {
  ocfg c;
  oconsts k;
  int bx, by, P, PL;       /* PL = P*(by+2): one plane */
  double *x, *y, *dx, *dlx, *dy, *dly, *r; /* 1-based */
  int *type_x, *type_y;    /* (bx,by) */
  double *ni, *ne, *nt, *nm; /* 3 levels */
  double *ki, *ke, *kt, *km; /* 5 stages */
  double *ph, *ph_mi, *Em;
  double t, dt, err_prev, t_m;
  double Vmax, Vsrc, Vd_pl, Vd_mi, Vd_or, Ida, Idc, Cap, Res, kv[6];
  double *sig_pl, *sig_mi, *sig_or; /* wall surface charge (bx+2), 1-based: sfc_lib.F90:7 */
  double *ks;                       /* its stage slopes (bx+2) x 6 */
  int ts, status;
  /* Poisson work */
  double *rhs, *A5, *sol, *wk, *part;
  /* replay / log */
  double* replay;
  int nreplay, ireplay;
  double* log;
  int nlog, caplog;
  double last_iters, last_resid, last_time;
};

#define A2(a, i, j) (a)[((size_t)(j)-1) * s->P + ((i)-1)]
#define A3(a, i, j, l) (a)[((size_t)(l)-1) * s->PL + ((size_t)(j)-1) * s->P + ((i)-1)]
#define TX(i, j) s->type_x[((j)-1) * s->bx + ((i)-1)]
#define TY(i, j) s->type_y[((j)-1) * s->bx + ((i)-1)]

static double wall(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* grid axis: props.F90:96-124 (x) and props.F90:134-163 (y) */
static void build_axis(int kind, double ratio, int n, double len, int is_y, double* x) {
  if (kind == 0) {
    double fl = (double)(float)(n + 1);
    for (int i = 1; i <= n + 2; i++) x[i] = len / fl * (double)(i - 1);
  } else if (kind == 1) {
    float xt = is_y ? 0.886f / (float)(n + 1) : (2.0f * 1.175f) / (float)(n + 1);
    double c = is_y ? F32(-0.886) : F32(-1.175);
    for (int i = 1; i <= n + 2; i++) x[i] = tanh(c + (double)xt * (double)(i - 1));
  } else {
    /* DIFF: geometric spacing (not in the reference): dx_k = ratio**min(k, n-k) */
    x[1] = 0.0;
    for (int kk = 0; kk <= n; kk++) {
      int m = kk < n - kk ? kk : n - kk;
      x[kk + 2] = x[kk + 1] + pow(ratio, (double)m);
    }
  }
  double t = x[1];
  for (int i = 1; i <= n + 2; i++) x[i] = x[i] - t;
  t = x[n + 2];
  for (int i = 1; i <= n + 2; i++) x[i] = x[i] / t;
  for (int i = 1; i <= n + 2; i++) x[i] = x[i] * len;
}

osim* o_create(const ocfg* c) {
  if (c->nx < 3 || c->ny < 1) return NULL;
  osim* s = (osim*)calloc(1, sizeof(osim));
  s->c = *c;
  if (c->ny == 1) { /* props.F90:64-65 */
    s->c.cyl = 0;
    s->c.rwall = 0;
  }
  consts_init(&s->k, c->p_torr, c->gam);
  s->bx = c->nx;
  s->by = c->ny;
  s->P = s->bx + 2;
  s->PL = s->P * (s->by + 2);
  int bx = s->bx, by = s->by;
  s->x = calloc(bx + 3, sizeof(double));
  s->y = calloc(by + 3, sizeof(double));
  s->dx = calloc(bx + 2, sizeof(double));
  s->dlx = calloc(bx + 1, sizeof(double));
  s->dy = calloc(by + 2, sizeof(double));
  s->dly = calloc(by + 1, sizeof(double));
  s->r = calloc(by + 3, sizeof(double));
  build_axis(c->grid_x, c->ratio_x, bx, c->l, 0, s->x);
  for (int i = 1; i <= bx + 1; i++) s->dx[i] = s->x[i + 1] - s->x[i];        /* props.F90:118-120 */
  for (int i = 1; i <= bx; i++) s->dlx[i] = 0.5 * (s->dx[i + 1] + s->dx[i]);  /* props.F90:122-124 */
  if (by > 1) {
    build_axis(c->grid_y, c->ratio_y, by, c->w, 1, s->y);
    for (int j = 1; j <= by + 1; j++) s->dy[j] = s->y[j + 1] - s->y[j];        /* props.F90:157-159 */
    for (int j = 1; j <= by; j++) s->dly[j] = 0.5 * (s->dy[j + 1] + s->dy[j]);  /* props.F90:161-163 */
    for (int j = 1; j <= by + 2; j++) s->r[j] = s->y[j];                       /* props.F90:165-166 */
  } else {
    /* one dimension (props.F90:175-178): y = 1 and dy = dx(1); the reference
     * has no dly and no r there (every y term is skipped), they are set to
     * dx(1) and 1 here so that the arrays hold finite values */
    for (int j = 1; j <= 3; j++) s->y[j] = 1.0;
    for (int j = 1; j <= 2; j++) s->dy[j] = s->dx[1];
    s->dly[1] = s->dx[1];
    for (int j = 1; j <= 3; j++) s->r[j] = 1.0;
  }
  s->type_x = calloc((size_t)bx * by, sizeof(int));
  s->type_y = calloc((size_t)bx * by, sizeof(int));
  for (int j = 1; j <= by; j++) /* props.F90:186-218 */
    for (int i = 1; i <= bx; i++) {
      if (j == 1) TY(i, j) = -1;
      if (j == by) TY(i, j) = s->c.rwall ? 3 : 1; /* props.F90:192-198 */
      int el = (s->y[j] <= c->ew) || (s->y[j] >= c->gap);
      if (i == 1) TX(i, j) = el ? -2 : -1;
      if (i == bx) TX(i, j) = el ? 2 : 1;
    }
  size_t PL = s->PL;
  s->ni = calloc(3 * PL, sizeof(double));
  s->ne = calloc(3 * PL, sizeof(double));
  s->nt = calloc(3 * PL, sizeof(double));
  s->nm = calloc(3 * PL, sizeof(double));
  s->ki = calloc(5 * PL, sizeof(double));
  s->ke = calloc(5 * PL, sizeof(double));
  s->kt = calloc(5 * PL, sizeof(double));
  s->km = calloc(5 * PL, sizeof(double));
  s->ph = calloc(PL, sizeof(double));
  s->ph_mi = calloc(PL, sizeof(double));
  s->Em = calloc(by + 3, sizeof(double)); /* em_lib.F90:104: Em = 0 */
  /* sfc_init (sfc_lib.F90:82-90): sig_pl = sig_mi = sig_or = 0 */
  s->sig_pl = calloc(bx + 3, sizeof(double));
  s->sig_mi = calloc(bx + 3, sizeof(double));
  s->sig_or = calloc(bx + 3, sizeof(double));
  s->ks = calloc((size_t)(bx + 3) * 6, sizeof(double));
  size_t N = (size_t)bx * by;
  s->rhs = calloc(N, sizeof(double));
  s->A5 = calloc(5 * N, sizeof(double));
  s->sol = calloc(N, sizeof(double));
  s->wk = NULL;
  o_init_uniform(s);
  /* circ_init (circ_lib.F90:213-223) */
  s->Vmax = s->Vsrc = s->Vd_pl = s->Vd_mi = s->Vd_or = c->vset;
  s->Cap = s->k.Cap;
  s->Res = c->res_ohm * s->k.e / (s->k.ph0 * s->k.t0); /* main.F90:286 */
  memset(s->kv, 0, sizeof(s->kv));
  s->t = 0.0;
  s->dt = c->dt0;
  s->err_prev = 1.0;
  s->t_m = F32(1e9);
  s->ts = 0;
  s->caplog = 0;
  return s;
}

void o_destroy(osim* s) {
  if (!s) return;
  free(s->x); free(s->y); free(s->dx); free(s->dlx); free(s->dy); free(s->dly); free(s->r);
  free(s->type_x); free(s->type_y);
  free(s->ni); free(s->ne); free(s->nt); free(s->nm);
  free(s->ki); free(s->ke); free(s->kt); free(s->km);
  free(s->ph); free(s->ph_mi); free(s->Em);
  free(s->sig_pl); free(s->sig_mi); free(s->sig_or); free(s->ks);
  free(s->rhs); free(s->A5); free(s->sol); free(s->wk); free(s->part);
  free(s->replay); free(s->log);
  free(s);
}

int o_dims(const osim* s, int32_t* o) {
  o[0] = s->bx; o[1] = s->by; o[2] = s->P; o[3] = 3;
  return 0;
}

int o_get_grid(const osim* s, double* x, double* y, double* dx, double* dlx, double* dy,
               double* dly, double* r, int32_t* tx, int32_t* ty) {
  if (x) memcpy(x, s->x + 1, (s->bx + 2) * sizeof(double));
  if (y) memcpy(y, s->y + 1, (s->by + 2) * sizeof(double));
  if (dx) memcpy(dx, s->dx + 1, (s->bx + 1) * sizeof(double));
  if (dlx) memcpy(dlx, s->dlx + 1, s->bx * sizeof(double));
  if (dy) memcpy(dy, s->dy + 1, (s->by + 1) * sizeof(double));
  if (dly) memcpy(dly, s->dly + 1, s->by * sizeof(double));
  if (r) memcpy(r, s->r + 1, (s->by + 2) * sizeof(double));
  for (int n = 0; n < s->bx * s->by; n++) {
    if (tx) tx[n] = s->type_x[n];
    if (ty) ty[n] = s->type_y[n];
  }
  return 0;
}

static double* field_ptr(const osim* s, int field, int level) {
  double* base[9] = {s->ni, s->ne, s->nt, s->nm, s->ph, s->ki, s->ke, s->kt, s->km};
  if (field < 0 || field > 8) return NULL;
  if (field == 4) return s->ph;
  int nl = field < 4 ? 3 : 5;
  if (level < 1 || level > nl) return NULL;
  return base[field] + (size_t)(level - 1) * s->PL;
}

int o_get_field(const osim* s, int field, int level, double* out) {
  double* p = field_ptr(s, field, level);
  if (!p) return -1;
  memcpy(out, p, s->PL * sizeof(double));
  return 0;
}

int o_set_field(osim* s, int field, int level, const double* in) {
  double* p = field_ptr(s, field, level);
  if (!p) return -1;
  memcpy(p, in, s->PL * sizeof(double));
  if (field == 4) memcpy(s->ph_mi, in, s->PL * sizeof(double));
  return 0;
}

/* ptcl_init: ptcl_lib.F90:15-30 */
void o_init_uniform(osim* s) {
  size_t PL = s->PL;
  memset(s->ki, 0, 5 * PL * sizeof(double));
  memset(s->ke, 0, 5 * PL * sizeof(double));
  memset(s->kt, 0, 5 * PL * sizeof(double));
  memset(s->km, 0, 5 * PL * sizeof(double));
  for (size_t n = 0; n < 3 * PL; n++) {
    s->ni[n] = s->k.n_init;
    s->ne[n] = s->k.n_init;
    s->nt[n] = s->ne[n] * s->k.Tg * s->k.kb / s->k.e / s->k.ph0;
    s->nm[n] = s->k.n_init;
  }
}

/* ---- Gaussian + seeded-noise initial condition (not in the reference) ---- */
static uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double unoise(uint64_t seed, int sp, int i, int j) {
  uint64_t h = splitmix64(splitmix64(seed) ^ ((uint64_t)sp << 56) ^ ((uint64_t)j << 28) ^
                          (uint64_t)i);
  return (double)(h >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

static void gaussian_fill(const osim* s, uint64_t seed, double peak, double sigma, double noise,
                          int sp, double* out) {
  const oconsts* k = &s->k;
  double xc = s->c.l * 0.5, yc = s->c.w * 0.5;
  for (int j = 1; j <= s->by + 2; j++)
    for (int i = 1; i <= s->bx + 2; i++) {
      /* in one dimension the blob is a profile in x alone */
      double ddx = s->x[i] - xc, ddy = s->by > 1 ? s->y[j] - yc : 0.0;
      double base = peak * exp(-(ddx * ddx + ddy * ddy) / (2.0 * sigma * sigma));
      double v;
      if (sp == 3) {
        v = base + k->n_zero;
      } else {
        int key = (sp == 2) ? 1 : sp; /* nt follows the noisy ne */
        double n = base * (1.0 + noise * unoise(seed, key, i, j)) + k->n_zero;
        v = (sp == 2) ? n * k->Tg * k->kb / k->e / k->ph0 : n;
      }
      out[(size_t)(j - 1) * s->P + (i - 1)] = v;
    }
}

void o_init_gaussian(osim* s, uint64_t seed, double peak, double sigma, double noise) {
  o_init_uniform(s);
  double* dst[4] = {s->ni, s->ne, s->nt, s->nm};
  for (int sp = 0; sp < 4; sp++) {
    gaussian_fill(s, seed, peak, sigma, noise, sp, dst[sp]);
    for (int l = 2; l <= 3; l++)
      memcpy(dst[sp] + (size_t)(l - 1) * s->PL, dst[sp], s->PL * sizeof(double));
  }
}

void o_gaussian_field(const ocfg* c, uint64_t seed, double peak, double sigma, double noise,
                      int species, double* out) {
  osim* s = o_create(c);
  gaussian_fill(s, seed, peak, sigma, noise, species, out);
  o_destroy(s);
}

void o_get_scalars(const osim* s, double* o) {
  o[0] = s->t; o[1] = s->dt; o[2] = s->err_prev; o[3] = s->t_m; o[4] = s->Vd_pl;
  o[5] = s->Vd_mi; o[6] = s->Vd_or; o[7] = s->Ida; o[8] = s->Idc; o[9] = s->Vsrc;
  o[10] = s->ts;
}
void o_set_scalars(osim* s, const double* o) {
  s->t = o[0]; s->dt = o[1]; s->err_prev = o[2]; s->t_m = o[3]; s->Vd_pl = o[4];
  s->Vd_mi = o[5]; s->Vd_or = o[6]; s->Ida = o[7]; s->Idc = o[8]; s->Vsrc = o[9];
  s->ts = (int)o[10];
}

/* ------------------------------------------------------------------------ */
/* ptclEqn: ptcl_lib.F90:99-550                                              */
/* ------------------------------------------------------------------------ */
static void ptcl_eqn(osim* s, int i, int j, int stage, double* tm_out) {
  const oconsts* K = &s->k;
  double a, Ex[3] = {0, 0, 0}, Ey[3] = {0, 0, 0}, Te[4] = {0, 0, 0, 0};
  double mue[3] = {0, 0, 0}, mut[3] = {0, 0, 0}, De[3] = {0, 0, 0}, Dt[3] = {0, 0, 0};
  double fluxi_x[3] = {0, 0, 0}, fluxi_y[3] = {0, 0, 0}, fluxe_x[3] = {0, 0, 0},
         fluxe_y[3] = {0, 0, 0}, fluxt_x[3] = {0, 0, 0}, fluxt_y[3] = {0, 0, 0},
         fluxm_x[3] = {0, 0, 0}, fluxm_y[3] = {0, 0, 0};
  double dfluxi_dx, dfluxi_dy = 0, dfluxe_dx, dfluxe_dy = 0, dfluxt_dx, dfluxt_dy = 0, dfluxm_dx,
         dfluxm_dy = 0;
  double ve, k_ir, k_ex, k_sc, k_si, nu, beta;
  const double mui = K->mui, Di = K->Di, Dm = K->Dm, vi = K->vi, gam = K->gam;
#define NI(a, b) A3(s->ni, a, b, 2)
#define NE(a, b) A3(s->ne, a, b, 2)
#define NT(a, b) A3(s->nt, a, b, 2)
#define NM(a, b) A3(s->nm, a, b, 2)
#define PH(a, b) A2(s->ph, a, b)
  int tx = TX(i - 1, j - 1), ty = TY(i - 1, j - 1);
  const double *dx = s->dx, *dy = s->dy, *r = s->r;

  /* X-dir fields: ptcl_lib.F90:112-121 */
  if (tx == -1) {
    Ex[1] = 0.0;
    Ex[2] = -(PH(i + 1, j) - PH(i, j)) / dx[i];
  } else if (tx == 1) {
    Ex[1] = -(PH(i, j) - PH(i - 1, j)) / dx[i - 1];
    Ex[2] = 0.0;
  } else {
    Ex[1] = -(PH(i, j) - PH(i - 1, j)) / dx[i - 1];
    Ex[2] = -(PH(i + 1, j) - PH(i, j)) / dx[i];
  }

  if (tx == 0) { /* centre: ptcl_lib.F90:125-161 */
    Te[1] = get_Te(NT(i - 1, j), NE(i - 1, j));
    Te[2] = get_Te(NT(i, j), NE(i, j));
    Te[3] = get_Te(NT(i + 1, j), NE(i + 1, j));
    mue[1] = 5e-1 * (get_mue(K, Te[1]) + get_mue(K, Te[2]));
    mue[2] = 5e-1 * (get_mue(K, Te[2]) + get_mue(K, Te[3]));
    De[1] = 5e-1 * (get_De(K, Te[1]) + get_De(K, Te[2]));
    De[2] = 5e-1 * (get_De(K, Te[2]) + get_De(K, Te[3]));
    mut[1] = 5.0 / 3.0 * mue[1]; mut[2] = 5.0 / 3.0 * mue[2];
    Dt[1] = 5.0 / 3.0 * De[1]; Dt[2] = 5.0 / 3.0 * De[2];
    fluxi_x[1] = get_flux(K, Ex[1], dx[i - 1], 1, mui, Di, NI(i - 1, j), NI(i, j));
    fluxe_x[1] = get_flux(K, Ex[1], dx[i - 1], -1, mue[1], De[1], NE(i - 1, j), NE(i, j));
    fluxt_x[1] = get_flux(K, Ex[1], dx[i - 1], -1, mut[1], Dt[1], NT(i - 1, j), NT(i, j));
    fluxm_x[1] = -Dm * (NM(i, j) - NM(i - 1, j)) / dx[i - 1];
    fluxi_x[2] = get_flux(K, Ex[2], dx[i], 1, mui, Di, NI(i, j), NI(i + 1, j));
    fluxe_x[2] = get_flux(K, Ex[2], dx[i], -1, mue[2], De[2], NE(i, j), NE(i + 1, j));
    fluxt_x[2] = get_flux(K, Ex[2], dx[i], -1, mut[2], Dt[2], NT(i, j), NT(i + 1, j));
    fluxm_x[2] = -Dm * (NM(i + 1, j) - NM(i, j)) / dx[i];
  } else if (tx < 0) { /* left: ptcl_lib.F90:164-221 */
    Te[2] = get_Te(NT(i, j), NE(i, j));
    Te[3] = get_Te(NT(i + 1, j), NE(i + 1, j));
    mue[2] = 5e-1 * (get_mue(K, Te[2]) + get_mue(K, Te[3]));
    De[2] = 5e-1 * (get_De(K, Te[2]) + get_De(K, Te[3]));
    mut[1] = 5.0 / 3.0 * mue[1]; mut[2] = 5.0 / 3.0 * mue[2];
    Dt[1] = 5.0 / 3.0 * De[1]; Dt[2] = 5.0 / 3.0 * De[2];
    fluxi_x[2] = get_flux(K, Ex[2], dx[i], 1, mui, Di, NI(i, j), NI(i + 1, j));
    fluxe_x[2] = get_flux(K, Ex[2], dx[i], -1, mue[2], De[2], NE(i, j), NE(i + 1, j));
    fluxt_x[2] = get_flux(K, Ex[2], dx[i], -1, mut[2], Dt[2], NT(i, j), NT(i + 1, j));
    fluxm_x[2] = -Dm * (NM(i + 1, j) - NM(i, j)) / dx[i];
    if (tx == -2) {
      a = (PH(i - 1, j) > PH(i, j)) ? 1.0 : 0.0;
      mue[1] = get_mue(K, Te[2]);
      mut[1] = 5.0 / 3.0 * mue[1]; mut[2] = 5.0 / 3.0 * mue[2];
      ve = sqrt((16.0 * K->e * K->ph0 * Te[2]) / (3.0 * K->pi * K->me)) * K->t0 / K->x0;
      fluxi_x[1] = (1.0 - a) * mui * Ex[1] * NI(i, j) - 2.5e-1 * vi * NI(i, j);
      fluxe_x[1] = -(a * mue[1] * Ex[1] * NE(i, j)) - 2.5e-1 * ve * NE(i, j) - gam * fluxi_x[1];
      fluxt_x[1] = -(a * mut[1] * Ex[1] * NT(i, j)) - 1.0 / 3.0 * ve * NT(i, j) -
                   gam * Te[2] * fluxi_x[1];
      fluxm_x[1] = -(5e-1 * vi * NM(i, j));
    } else if (tx == -1) {
      fluxi_x[1] = 0.0; fluxe_x[1] = 0.0; fluxt_x[1] = 0.0; fluxm_x[1] = 0.0;
    }
  } else if (tx > 0) { /* right: ptcl_lib.F90:224-282 */
    Te[1] = get_Te(NT(i - 1, j), NE(i - 1, j));
    Te[2] = get_Te(NT(i, j), NE(i, j));
    mue[1] = 5e-1 * (get_mue(K, Te[1]) + get_mue(K, Te[2]));
    De[1] = 5e-1 * (get_De(K, Te[1]) + get_De(K, Te[2]));
    mut[1] = 5.0 / 3.0 * mue[1]; mut[2] = 5.0 / 3.0 * mue[2];
    Dt[1] = 5.0 / 3.0 * De[1]; Dt[2] = 5.0 / 3.0 * De[2];
    fluxi_x[1] = get_flux(K, Ex[1], dx[i - 1], 1, mui, Di, NI(i - 1, j), NI(i, j));
    fluxe_x[1] = get_flux(K, Ex[1], dx[i - 1], -1, mue[1], De[1], NE(i - 1, j), NE(i, j));
    fluxt_x[1] = get_flux(K, Ex[1], dx[i - 1], -1, mut[1], Dt[1], NT(i - 1, j), NT(i, j));
    fluxm_x[1] = -Dm * (NM(i, j) - NM(i - 1, j)) / dx[i - 1];
    if (tx == 2) {
      a = (PH(i + 1, j) > PH(i, j)) ? 1.0 : 0.0;
      mue[2] = get_mue(K, Te[2]);
      mut[1] = 5.0 / 3.0 * mue[1]; mut[2] = 5.0 / 3.0 * mue[2];
      ve = sqrt((16.0 * K->e * K->ph0 * Te[2]) / (3.0 * K->pi * K->me)) * K->t0 / K->x0;
      fluxi_x[2] = (1.0 - a) * mui * Ex[2] * NI(i, j) + 2.5e-1 * vi * NI(i, j);
      fluxe_x[2] = -(a * mue[2] * Ex[2] * NE(i, j)) + 2.5e-1 * ve * NE(i, j) - gam * fluxi_x[2];
      fluxt_x[2] = -(a * mut[2] * Ex[2] * NT(i, j)) + 1.0 / 3.0 * ve * NT(i, j) -
                   gam * Te[2] * fluxi_x[2];
      fluxm_x[2] = 5e-1 * vi * NM(i, j);
    } else if (tx == 1) {
      fluxi_x[2] = 0.0; fluxe_x[2] = 0.0; fluxt_x[2] = 0.0; fluxm_x[2] = 0.0;
    }
  }

  /* Y-dir: ptcl_lib.F90:285-443, only for ny > 1; in one dimension Ey,
   * the y fluxes and their divergence keep their initial zeros (the
   * reference's saved locals, ptcl_lib.F90:103-107, are never assigned) */
  const int twod = s->by > 1;
  if (!twod) {
  } else if (ty == -1) {
    Ey[1] = 0.0;
    Ey[2] = -(PH(i, j + 1) - PH(i, j)) / dy[j];
  } else if (ty == 1) {
    Ey[1] = -(PH(i, j) - PH(i, j - 1)) / dy[j - 1];
    Ey[2] = 0.0;
  } else if (ty == 3) { /* dielectric wall: ptcl_lib.F90:292-294 */
    Ey[1] = -(PH(i, j) - PH(i, j - 1)) / dy[j - 1];
    Ey[2] = -s->sig_mi[i];
  } else {
    Ey[1] = -(PH(i, j) - PH(i, j - 1)) / dy[j - 1];
    Ey[2] = -(PH(i, j + 1) - PH(i, j)) / dy[j];
  }
  if (!twod) {
  } else if (ty == 0) {
    Te[1] = get_Te(NT(i, j - 1), NE(i, j - 1));
    Te[2] = get_Te(NT(i, j), NE(i, j));
    Te[3] = get_Te(NT(i, j + 1), NE(i, j + 1));
    mue[1] = 5e-1 * (get_mue(K, Te[1]) + get_mue(K, Te[2]));
    mue[2] = 5e-1 * (get_mue(K, Te[2]) + get_mue(K, Te[3]));
    De[1] = 5e-1 * (get_De(K, Te[1]) + get_De(K, Te[2]));
    De[2] = 5e-1 * (get_De(K, Te[2]) + get_De(K, Te[3]));
    mut[1] = 5.0 / 3.0 * mue[1]; mut[2] = 5.0 / 3.0 * mue[2];
    Dt[1] = 5.0 / 3.0 * De[1]; Dt[2] = 5.0 / 3.0 * De[2];
    fluxi_y[1] = get_flux(K, Ey[1], dy[j - 1], 1, mui, Di, NI(i, j - 1), NI(i, j));
    fluxe_y[1] = get_flux(K, Ey[1], dy[j - 1], -1, mue[1], De[1], NE(i, j - 1), NE(i, j));
    fluxt_y[1] = get_flux(K, Ey[1], dy[j - 1], -1, mut[1], Dt[1], NT(i, j - 1), NT(i, j));
    fluxm_y[1] = -Dm * (NM(i, j) - NM(i, j - 1)) / dy[j - 1];
    fluxi_y[2] = get_flux(K, Ey[2], dy[j], 1, mui, Di, NI(i, j), NI(i, j + 1));
    fluxe_y[2] = get_flux(K, Ey[2], dy[j], -1, mue[2], De[2], NE(i, j), NE(i, j + 1));
    fluxt_y[2] = get_flux(K, Ey[2], dy[j], -1, mut[2], Dt[2], NT(i, j), NT(i, j + 1));
    fluxm_y[2] = -Dm * (NM(i, j + 1) - NM(i, j)) / dy[j];
  } else if (ty < 0) {
    Te[2] = get_Te(NT(i, j), NE(i, j));
    Te[3] = get_Te(NT(i, j + 1), NE(i, j + 1));
    mue[2] = 5e-1 * (get_mue(K, Te[2]) + get_mue(K, Te[3]));
    De[2] = 5e-1 * (get_De(K, Te[2]) + get_De(K, Te[3]));
    mut[1] = 5.0 / 3.0 * mue[1]; mut[2] = 5.0 / 3.0 * mue[2];
    Dt[1] = 5.0 / 3.0 * De[1]; Dt[2] = 5.0 / 3.0 * De[2];
    fluxi_y[2] = get_flux(K, Ey[2], dy[j], 1, mui, Di, NI(i, j), NI(i, j + 1));
    fluxe_y[2] = get_flux(K, Ey[2], dy[j], -1, mue[2], De[2], NE(i, j), NE(i, j + 1));
    fluxt_y[2] = get_flux(K, Ey[2], dy[j], -1, mut[2], Dt[2], NT(i, j), NT(i, j + 1));
    fluxm_y[2] = -Dm * (NM(i, j + 1) - NM(i, j)) / dy[j];
    fluxi_y[1] = 0.0; fluxe_y[1] = 0.0; fluxt_y[1] = 0.0; fluxm_y[1] = 0.0;
  } else if (ty > 0) {
    Te[1] = get_Te(NT(i, j - 1), NE(i, j - 1));
    Te[2] = get_Te(NT(i, j), NE(i, j));
    mue[1] = 5e-1 * (get_mue(K, Te[1]) + get_mue(K, Te[2]));
    De[1] = 5e-1 * (get_De(K, Te[1]) + get_De(K, Te[2]));
    mut[1] = 5.0 / 3.0 * mue[1]; mut[2] = 5.0 / 3.0 * mue[2];
    Dt[1] = 5.0 / 3.0 * De[1]; Dt[2] = 5.0 / 3.0 * De[2];
    fluxi_y[1] = get_flux(K, Ey[1], dy[j - 1], 1, mui, Di, NI(i, j - 1), NI(i, j));
    fluxe_y[1] = get_flux(K, Ey[1], dy[j - 1], -1, mue[1], De[1], NE(i, j - 1), NE(i, j));
    fluxt_y[1] = get_flux(K, Ey[1], dy[j - 1], -1, mut[1], Dt[1], NT(i, j - 1), NT(i, j));
    fluxm_y[1] = -Dm * (NM(i, j) - NM(i, j - 1)) / dy[j - 1];
    if (ty == 3) { /* flux into the dielectric: ptcl_lib.F90:393-409 (type 2 is never assigned) */
      Te[2] = get_Te(NT(i, j), NE(i, j));
      mue[2] = get_mue(K, Te[2]);
      De[2] = get_De(K, Te[2]);
      mut[2] = get_mut(K, Te[2]);
      Dt[2] = get_Dt(K, Te[2]);
      fluxi_y[2] = get_flux(K, Ey[2], dy[j], 1, mui, Di, NI(i, j), 0.0);
      fluxe_y[2] = get_flux(K, Ey[2], dy[j], -1, mue[2], De[2], NE(i, j), 0.0);
      fluxt_y[2] = get_flux(K, Ey[2], dy[j], -1, mut[2], Dt[2], NT(i, j), 0.0);
      fluxm_y[2] = Dm * NM(i, j) / dy[j];
    } else {
      fluxi_y[2] = 0.0; fluxe_y[2] = 0.0; fluxt_y[2] = 0.0; fluxm_y[2] = 0.0;
    }
  }

  /* divergence: ptcl_lib.F90:446-480 */
  dfluxi_dx = (fluxi_x[2] - fluxi_x[1]) / s->dlx[i - 1];
  dfluxe_dx = (fluxe_x[2] - fluxe_x[1]) / s->dlx[i - 1];
  dfluxt_dx = (fluxt_x[2] - fluxt_x[1]) / s->dlx[i - 1];
  dfluxm_dx = (fluxm_x[2] - fluxm_x[1]) / s->dlx[i - 1];
  fluxe_x[1] = (fluxe_x[2] + fluxe_x[1]) / 2.0;
  Ex[1] = (Ex[1] + Ex[2]) / 2.0;
  if (!twod) {
  } else if (s->c.cyl) {
    dfluxi_dy = (fluxi_y[2] * (r[j + 1] + r[j]) / 2.0 - fluxi_y[1] * (r[j] + r[j - 1]) / 2.0) /
                s->dly[j - 1] / r[j];
    dfluxe_dy = (fluxe_y[2] * (r[j + 1] + r[j]) / 2.0 - fluxe_y[1] * (r[j] + r[j - 1]) / 2.0) /
                s->dly[j - 1] / r[j];
    dfluxt_dy = (fluxt_y[2] * (r[j + 1] + r[j]) / 2.0 - fluxt_y[1] * (r[j] + r[j - 1]) / 2.0) /
                s->dly[j - 1] / r[j];
    dfluxm_dy = (fluxm_y[2] * (r[j + 1] + r[j]) / 2.0 - fluxm_y[1] * (r[j] + r[j - 1]) / 2.0) /
                s->dly[j - 1] / r[j];
  } else {
    dfluxi_dy = (fluxi_y[2] - fluxi_y[1]) / s->dly[j - 1];
    dfluxe_dy = (fluxe_y[2] - fluxe_y[1]) / s->dly[j - 1];
    dfluxt_dy = (fluxt_y[2] - fluxt_y[1]) / s->dly[j - 1];
    dfluxm_dy = (fluxm_y[2] - fluxm_y[1]) / s->dly[j - 1];
  }
  if (twod) {
    fluxe_y[1] = (fluxe_y[2] + fluxe_y[1]) / 2.0;
    Ey[1] = (Ey[1] + Ey[2]) / 2.0;
  }

  /* rates: ptcl_lib.F90:483-488 */
  k_ir = get_k_ir(K, Te[2]);
  k_sc = get_k_sc(K, Te[2]);
  k_si = get_k_si(K, Te[2]);
  k_ex = get_k_ex(K, Te[2]);
  nu = get_nu(K, Te[2]);
  beta = get_beta(K, Te[2]);
  (void)k_sc;

  /* sources: ptcl_lib.F90:491-514 */
  double ne_c = NE(i, j), ni_c = NI(i, j), nm_c = NM(i, j), ninf = K->ninf;
  double term_sie = k_ir * ninf * ne_c - beta * ni_c * ne_c + k_si * nm_c * ne_c +
                    K->k_mp * ipow(nm_c, 2);
  double term_sm = k_ex * ninf * ne_c - k_si * nm_c * ne_c - k_sc * nm_c * ne_c -
                   K->k_r * nm_c * ne_c - 2.0 * K->k_mp * ipow(nm_c, 2) -
                   K->k_2q * ninf * nm_c - K->k_3q * ipow(ninf, 2) * nm_c;
  double term_st1 = -(fluxe_x[1] * Ex[1] + fluxe_y[1] * Ey[1]);
  double term_st2 = -(ne_c * nu * K->me / K->mi * (Te[2] / 1.5 - K->kTe));
  double term_st3 = -(K->H_ir * k_ir * ninf * ne_c) - K->H_si * k_si * nm_c * ne_c -
                    K->H_ex * k_ex * ninf * ne_c +
                    ne_c * K->wp02 * s->Em[j] / sqrt(K->w2 + ipow(nu, 2));

  double kiv = -dfluxi_dx - dfluxi_dy + term_sie;
  double kev = -dfluxe_dx - dfluxe_dy + term_sie;
  double ktv = -dfluxt_dx - dfluxt_dy + term_st1 + term_st2 + term_st3;
  double kmv = -dfluxm_dx - dfluxm_dy + term_sm;
  A3(s->ki, i, j, stage) = kiv;
  A3(s->ke, i, j, stage) = kev;
  A3(s->kt, i, j, stage) = ktv;
  A3(s->km, i, j, stage) = kmv;

  if (stage == 5) { /* ptcl_lib.F90:526-528 */
    double v = 1.0 / (mue[2] * ne_c + mui * ni_c);
    if (v < *tm_out) *tm_out = v;
  }
  if (isnan(kiv) || isnan(kev) || isnan(ktv) || isnan(kmv)) s->status = -3;
#undef NI
#undef NE
#undef NT
#undef NM
#undef PH
}

void o_rhs_sweep(osim* s, int stage) { /* ptcl_lib.F90:44-48 */
  double tm = s->t_m;
  int nth = s->c.nthreads > 0 ? s->c.nthreads : 1;
#pragma omp parallel for num_threads(nth) schedule(static) reduction(min : tm)
  for (int j = 2; j <= s->by + 1; j++)
    for (int i = 2; i <= s->bx + 1; i++) ptcl_eqn(s, i, j, stage, &tm);
  s->t_m = tm;
  if (g_neg_density) s->status = -2;
}

/* ------------------------------------------------------------------------ */
/* circ_step: circ_lib.F90:19-210                                            */
/* ------------------------------------------------------------------------ */
static double vsource(const osim* s) {
  const double t = s->t, Vmax = s->Vmax, pi = s->k.pi;
  switch (s->c.rf) {
    case 0: return Vmax;
    case 1: return Vmax * sin(pi * t);
    case 2: return t < 2.5e-1 ? Vmax * ipow(sin(4.0 * pi * t), 3) : 0.0;
    case 3:
      if (t < 2.3e-1) return Vmax * ipow(sin(4.0 * pi * t), 2);
      if (t < 3.35e-1) return Vmax * (sin(3.81 * pi * (t + 5.85e-2)) * 7.19e-1 + 2.81e-1);
      if (t < 8.2e-1) return Vmax * (ipow(sin(3.81 * pi * (t + 1.2e-1) / 3.5), 3) * (-4.38e-1));
      return 0.0;
    case 4: { /* DIFF: trapezoid pulse (not in the reference) */
      double tr = s->c.t_rise, tw = s->c.t_width, tf = s->c.t_fall;
      if (t < tr) return Vmax * (t / tr);
      if (t < tr + tw) return Vmax;
      double te = tr + tw + tf;
      if (t < te) return Vmax * ((te - t) / tf);
      return 0.0;
    }
  }
  return Vmax;
}

void o_circ_step(osim* s, int stage) {
  const oconsts* K = &s->k;
  int bx = s->bx, by = s->by, active = (s->c.circuit == 1);
  s->Vsrc = vsource(s);
#define NI(a, b) A3(s->ni, a, b, 2)
#define NE(a, b) A3(s->ne, a, b, 2)
#define NT(a, b) A3(s->nt, a, b, 2)
#define PH(a, b) A2(s->ph, a, b)
  int i = 3;
  s->Ida = 0;
  for (int j = 2; j <= by + 1; j++) {
    if (TX(i - 2, j - 1) == -2) {
      double dr;
      if (by > 1) {
        dr = s->c.cyl ? s->dy[j - 1] * s->r[j - 1] * 2.0 * K->pi : s->dy[j - 1] * s->c.w;
        if (!active && s->r[j] < s->c.ew) dr = 0.0;
      } else {
        dr = s->c.w * s->c.w * K->pi; /* circ_lib.F90:98-99, 131-132 */
      }
      double Ex = -(PH(i, j) - PH(i - 1, j)) / s->dx[i - 1];
      double T1 = get_Te(NT(i - 1, j), NE(i - 1, j)), T2 = get_Te(NT(i, j), NE(i, j));
      double mue = 5e-1 * (get_mue(K, T1) + get_mue(K, T2));
      double De = 5e-1 * (get_De(K, T1) + get_De(K, T2));
      double flxi = get_flux(K, Ex, s->dx[i - 1], 1, K->mui, K->Di, NI(i - 1, j), NI(i, j));
      double flxe = get_flux(K, Ex, s->dx[i - 1], -1, mue, De, NE(i - 1, j), NE(i, j));
      s->Ida = s->Ida + dr * (flxi - flxe);
    }
  }
  i = bx;
  s->Idc = 0;
  for (int j = 2; j <= by + 1; j++) {
    if (TX(i, j - 1) == 2) {
      double dr;
      if (by > 1) {
        dr = s->c.cyl ? s->dy[j - 1] * s->r[j - 1] * 2.0 * K->pi : s->dy[j - 1] * s->c.w;
        if (!active && s->r[j] < s->c.ew) dr = 0.0;
      } else {
        dr = s->c.w * s->c.w * K->pi; /* circ_lib.F90:98-99, 131-132 */
      }
      double Ex = -(PH(i + 1, j) - PH(i, j)) / s->dx[i];
      double T1 = get_Te(NT(i, j), NE(i, j)), T2 = get_Te(NT(i + 1, j), NE(i + 1, j));
      double mue = 5e-1 * (get_mue(K, T1) + get_mue(K, T2));
      double De = 5e-1 * (get_De(K, T1) + get_De(K, T2));
      double flxi = get_flux(K, Ex, s->dx[i], 1, K->mui, K->Di, NI(i, j), NI(i + 1, j));
      double flxe = get_flux(K, Ex, s->dx[i], -1, mue, De, NE(i, j), NE(i + 1, j));
      s->Idc = s->Idc + dr * (flxi - flxe);
    }
  }
  s->kv[stage] = -(s->Idc - (s->Vsrc - s->Vd_mi) / s->Res) / s->Cap;
  double dt = s->dt, *kv = s->kv;
  if (stage == 1) s->Vd_mi = s->Vd_or + dt * kv[1] / 3.0;
  else if (stage == 2) s->Vd_mi = s->Vd_or + dt * (kv[1] + kv[2]) / 6.0;
  else if (stage == 3) s->Vd_mi = s->Vd_or + dt * (kv[1] + kv[3] * 3.0) / 8.0;
  else if (stage == 4) s->Vd_mi = s->Vd_or + dt * (kv[1] - kv[3] * 3.0 + kv[4] * 4.0) / 2.0;
  else s->Vd_pl = s->Vd_or + dt * (kv[1] + kv[4] * 4.0 + kv[5]) / 6.0;
  if (s->c.rf != 1 && !active) { /* circ_lib.F90:176-177 */
    s->Vd_pl = s->Vsrc;
    s->Vd_mi = s->Vsrc;
  }
  i = 2;
  if (by > 1) {
    for (int j = 2; j <= by + 1; j++) { /* circ_lib.F90:180-198 */
      if (s->r[j] <= s->c.ew) {
        if (TX(i - 1, j - 1) == -2) PH(i - 1, j) = (stage == 5 ? s->Vd_pl : s->Vd_mi) / 2.0;
        if (TX(bx, j - 1) == 2) PH(bx + 2, j) = (stage == 5 ? s->Vd_pl : s->Vd_mi) / (-2.0);
      }
    }
  } else { /* one dimension: the source voltage itself, circ_lib.F90:199-202 */
    if (TX(1, 1) == -2) PH(1, 2) = s->Vsrc / 2.0;
    if (TX(bx, 1) == 2) PH(bx + 2, 2) = s->Vsrc / (-2.0);
  }
#undef NI
#undef NE
#undef NT
#undef PH
}

/* ------------------------------------------------------------------------ */
/* rk_step order 4 (Merson): ptcl_props.F90:71-138                          */
/* ------------------------------------------------------------------------ */
static void rk_step(osim* s, double* n, const double* k, int stage, double dt, double* nerr,
                    double n_min) {
  int bx = s->bx, by = s->by, nth = s->c.nthreads > 0 ? s->c.nthreads : 1;
  size_t PL = s->PL;
#define N(l) A3(n, i, j, l)
#define K(l) A3(k, i, j, l)
  int out = (stage == 5) ? 1 : 2;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int j = 2; j <= by + 1; j++)
    for (int i = 2; i <= bx + 1; i++) {
      if (stage == 1) N(2) = N(3) + K(1) * dt / 3.0;
      else if (stage == 2) N(2) = N(3) + dt * (K(1) + K(2)) / 6.0;
      else if (stage == 3) N(2) = N(3) + dt * (K(1) + K(3) * 3.0) / 8.0;
      else if (stage == 4) N(2) = N(3) + dt * (K(1) - K(3) * 3.0 + K(4) * 4.0) / 2.0;
      else N(1) = N(3) + dt * (K(1) + K(4) * 4.0 + K(5)) / 6.0;
    }
  double* o = n + (size_t)(out - 1) * PL; /* clamp covers the ghosts too */
  for (size_t m = 0; m < PL; m++) o[m] = o[m] > n_min ? o[m] : n_min;
  if (stage == 5) { /* whole array, column-major, sequential (ptcl_props.F90:125-134) */
    const double abs_tol = 1e-3, rel_tol = 1e-3;
    double sum = 0.0;
    for (int j = 1; j <= by + 2; j++)
      for (int i = 1; i <= bx + 2; i++) {
        double e = fabs(dt * (K(1) * 2.0 / 30.0 - K(3) * 3.0 / 10.0 + K(4) * 4.0 / 15.0 -
                              K(5) / 30.0));
        double q = e / (abs_tol + rel_tol * fabs(N(3)));
        sum = sum + q * q;
      }
    *nerr = sum;
  }
#undef N
#undef K
}

void o_rk_all(osim* s, int stage, double* nerr4) {
  double e[4] = {0, 0, 0, 0};
  rk_step(s, s->ni, s->ki, stage, s->dt, &e[0], s->k.n_zero); /* ptcl_lib.F90:54-57 */
  rk_step(s, s->ne, s->ke, stage, s->dt, &e[1], s->k.n_zero);
  rk_step(s, s->nt, s->kt, stage, s->dt, &e[2], s->k.nmin_t);
  rk_step(s, s->nm, s->km, stage, s->dt, &e[3], s->k.n_zero);
  if (nerr4) memcpy(nerr4, e, sizeof(e));
}

static void copy_level(osim* s, int dst, int src) {
  double* a[4] = {s->ni, s->ne, s->nt, s->nm};
  for (int q = 0; q < 4; q++)
    memcpy(a[q] + (size_t)(dst - 1) * s->PL, a[q] + (size_t)(src - 1) * s->PL,
           s->PL * sizeof(double));
}

static void log_push(osim* s, const double* rec) {
  if (s->nlog == s->caplog) {
    s->caplog = s->caplog ? 2 * s->caplog : 256;
    s->log = realloc(s->log, (size_t)s->caplog * OREC_LEN * sizeof(double));
  }
  memcpy(s->log + (size_t)s->nlog * OREC_LEN, rec, OREC_LEN * sizeof(double));
  s->nlog++;
}

/* ptcl_step: ptcl_lib.F90:32-97 */
int o_ptcl_step(osim* s) {
  int stage = 0;
  s->status = 0;
  g_neg_density = 0;
  for (;;) {
    stage = stage + 1;
    if (stage > 5) break;
    o_rhs_sweep(s, stage);
    if (s->status < 0) return s->status;
    o_circ_step(s, stage);
    if (s->c.rwall) o_sfc_step(s, stage); /* ptcl_lib.F90:52 */
    double e[4];
    o_rk_all(s, stage, e);
    if (stage == 5) {
      double dt_stage = s->dt;
      double err_cur = pow(e[0] + e[1] + e[3] + e[2], 0.5); /* ni+ne+nm+nt, :60 */
      double scfac = 8e-1 * pow(err_cur, -7e-1 / 4.0) * pow(s->err_prev, 4e-1 / 4.0);
      scfac = fmin(2.5, fmax(3e-1, scfac));
      s->dt = scfac * s->dt;
      int accept = err_cur <= 1.0;
      if (s->replay && s->ireplay < s->nreplay) { /* replayed controller output */
        const double* rec = s->replay + (size_t)s->ireplay * OREC_LEN;
        s->dt = rec[4];
        accept = rec[5] != 0.0;
        s->ireplay++;
      }
      double rec[OREC_LEN] = {s->ts, s->t, dt_stage, err_cur, s->dt, accept, -1, 0, 0, s->t_m};
      log_push(s, rec);
      if (s->dt <= 1.1e-12) return -4; /* ptcl_lib.F90:68-72 */
      if (accept) {
        s->err_prev = err_cur;
      } else { /* restart: ptcl_lib.F90:76-93 */
        stage = 0;
        copy_level(s, 1, 3);
        copy_level(s, 2, 3);
        s->Vd_pl = s->Vd_or;
        s->Vd_mi = s->Vd_or;
        for (int i = 1; i <= s->bx + 2; i++) s->sig_pl[i] = s->sig_mi[i] = s->sig_or[i];
      }
    }
  }
  return 0;
}

/* sfc_step: sfc_lib.F90:20-79, the dielectric wall's surface charge, Merson
 * stages with the circuit's tableau (ks = ion minus electron flux into the
 * wall).  Two departures from the text, both in the reference's own call:
 *  - ptcl_step calls sfc_step(g, stage, ni(:,:,2), ne(:,:,2), nt(:,:,2))
 *    (ptcl_lib.F90:52) against the dummy order (g, stage, ne, ni, nt), so
 *    inside it `ne` is the ion density and `ni` the electron density: kept
 *    (Te = nt / n_ion, the "ion" flux from the electron density and so on);
 *  - the row test reads type_y(i-1, by+1), one column past the (bx, by)
 *    array: undefined in the reference; here the top row type_y(i-1, by),
 *    the row the wall belongs to (props.F90:192-198). */
void o_sfc_step(osim* s, int stage) {
  const oconsts* K = &s->k;
  const int bx = s->bx, by = s->by, j = by + 1;
#define NE_D(a, b) A3(s->ni, a, b, 2) /* dummy ne <- actual ni */
#define NI_D(a, b) A3(s->ne, a, b, 2) /* dummy ni <- actual ne */
#define NT_D(a, b) A3(s->nt, a, b, 2)
  for (int i = 2; i <= bx + 1; i++) {
    if (TY(i - 1, by) != 3) continue;
    const double Te = get_Te(NT_D(i, j), NE_D(i, j));
    const double mue = get_mue(K, Te), De = get_De(K, Te);
    const double fluxi = get_flux(K, -s->sig_mi[i], s->dy[j], 1, K->mui, K->Di, NI_D(i, j), 0.0);
    const double fluxe = get_flux(K, -s->sig_mi[i], s->dy[j], -1, mue, De, NE_D(i, j), 0.0);
    s->ks[(size_t)i * 6 + stage] = fluxi - fluxe;
  }
#undef NE_D
#undef NI_D
#undef NT_D
  const double dt = s->dt;
  for (int i = 2; i <= bx + 1; i++) {
    if (TY(i - 1, by) != 3) continue;
    const double* k = s->ks + (size_t)i * 6;
    if (stage == 1) s->sig_mi[i] = s->sig_or[i] + dt * k[1] / 3.0;
    else if (stage == 2) s->sig_mi[i] = s->sig_or[i] + dt * (k[1] + k[2]) / 6.0;
    else if (stage == 3) s->sig_mi[i] = s->sig_or[i] + dt * (k[1] + k[3] * 3.0) / 8.0;
    else if (stage == 4)
      s->sig_mi[i] = s->sig_or[i] + dt * (k[1] - k[3] * 3.0 + k[4] * 4.0) / 2.0;
    else s->sig_pl[i] = s->sig_or[i] + dt * (k[1] + k[4] * 4.0 + k[5]) / 6.0;
  }
  /* the halo exchange of sig (sfc_lib.F90:60-72) is MPI_PROC_NULL at one rank */
}

int o_get_sig(const osim* s, int which, double* out) {
  const double* a = which == 0 ? s->sig_pl : which == 1 ? s->sig_mi : which == 2 ? s->sig_or : 0;
  if (!a) return -1;
  memcpy(out, a + 1, (size_t)(s->bx + 2) * sizeof(double));
  return 0;
}
int o_set_sig(osim* s, int which, const double* in) {
  double* a = which == 0 ? s->sig_pl : which == 1 ? s->sig_mi : which == 2 ? s->sig_or : 0;
  if (!a) return -1;
  memcpy(a + 1, in, (size_t)(s->bx + 2) * sizeof(double));
  return 0;
}

/* ------------------------------------------------------------------------ */
/* laplEqn: lapl_lib.F90:123-379                                             */
/* ------------------------------------------------------------------------ */
static double lapl_eqn(osim* s, int i, int j, const double* ne, const double* ni,
                       const double* nt, double sig) {
  const oconsts* K = &s->k;
  double dfdx = 0, dfdy = 0, dflxe_x = 0, dflxe_y = 0, dflxi_x = 0, dflxi_y = 0;
  double Ex[3] = {0, 0, 0}, Ey[3] = {0, 0, 0}, mue[3] = {0, 0, 0}, De[3] = {0, 0, 0};
  double Te[4] = {0, 0, 0, 0}, ve, a, flxi[3] = {0, 0, 0}, flxe[3] = {0, 0, 0};
  const double mui = K->mui, Di = K->Di, vi = K->vi, gam = K->gam;
  const double *dx = s->dx, *dy = s->dy, *r = s->r;
#define NE(a, b) A2(ne, a, b)
#define NI(a, b) A2(ni, a, b)
#define NT(a, b) A2(nt, a, b)
#define PH(a, b) A2(s->ph, a, b)
  int tx = TX(i - 1, j - 1), ty = TY(i - 1, j - 1);
  if (tx == -1) {
    Ex[1] = 0.0;
    Ex[2] = -(PH(i + 1, j) - PH(i, j)) / dx[i];
  } else if (tx == 1) {
    Ex[1] = -(PH(i, j) - PH(i - 1, j)) / dx[i - 1];
    Ex[2] = 0.0;
  } else {
    Ex[1] = -(PH(i, j) - PH(i - 1, j)) / dx[i - 1];
    Ex[2] = -(PH(i + 1, j) - PH(i, j)) / dx[i];
  }
  if (tx == 0) {
    Te[1] = get_Te(NT(i - 1, j), NE(i - 1, j));
    Te[2] = get_Te(NT(i, j), NE(i, j));
    Te[3] = get_Te(NT(i + 1, j), NE(i + 1, j));
    mue[1] = 5e-1 * (get_mue(K, Te[1]) + get_mue(K, Te[2]));
    mue[2] = 5e-1 * (get_mue(K, Te[2]) + get_mue(K, Te[3]));
    De[1] = 5e-1 * (get_De(K, Te[1]) + get_De(K, Te[2]));
    De[2] = 5e-1 * (get_De(K, Te[2]) + get_De(K, Te[3]));
    flxi[1] = get_flux(K, Ex[1], dx[i - 1], 1, mui, Di, NI(i - 1, j), NI(i, j));
    flxe[1] = get_flux(K, Ex[1], dx[i - 1], -1, mue[1], De[1], NE(i - 1, j), NE(i, j));
    flxi[2] = get_flux(K, Ex[2], dx[i], 1, mui, Di, NI(i, j), NI(i + 1, j));
    flxe[2] = get_flux(K, Ex[2], dx[i], -1, mue[2], De[2], NE(i, j), NE(i + 1, j));
  } else if (tx < 0) {
    Te[2] = get_Te(NT(i, j), NE(i, j));
    Te[3] = get_Te(NT(i + 1, j), NE(i + 1, j));
    mue[2] = 5e-1 * (get_mue(K, Te[2]) + get_mue(K, Te[3]));
    De[2] = 5e-1 * (get_De(K, Te[2]) + get_De(K, Te[3]));
    flxi[2] = get_flux(K, Ex[2], dx[i], 1, mui, Di, NI(i, j), NI(i + 1, j));
    flxe[2] = get_flux(K, Ex[2], dx[i], -1, mue[2], De[2], NE(i, j), NE(i + 1, j));
    if (tx == -2) {
      a = (PH(i - 1, j) > PH(i, j)) ? 1.0 : 0.0;
      mue[1] = get_mue(K, Te[2]);
      ve = sqrt((16.0 * K->e * K->ph0 * Te[2]) / (3.0 * K->pi * K->me)) * K->t0 / K->x0;
      flxi[1] = (1.0 - a) * mui * Ex[1] * NI(i, j) - 2.5e-1 * vi * NI(i, j);
      flxe[1] = -(a * mue[1] * Ex[1] * NE(i, j)) - 2.5e-1 * ve * NE(i, j) - gam * flxi[1];
    } else if (tx == -1) {
      flxi[1] = 0.0;
      flxe[1] = 0.0;
    }
  } else if (tx > 0) {
    Te[1] = get_Te(NT(i - 1, j), NE(i - 1, j));
    Te[2] = get_Te(NT(i, j), NE(i, j));
    mue[1] = 5e-1 * (get_mue(K, Te[1]) + get_mue(K, Te[2]));
    De[1] = 5e-1 * (get_De(K, Te[1]) + get_De(K, Te[2]));
    flxi[1] = get_flux(K, Ex[1], dx[i - 1], 1, mui, Di, NI(i - 1, j), NI(i, j));
    flxe[1] = get_flux(K, Ex[1], dx[i - 1], -1, mue[1], De[1], NE(i - 1, j), NE(i, j));
    if (tx == 2) {
      a = (PH(i + 1, j) > PH(i, j)) ? 1.0 : 0.0;
      mue[2] = get_mue(K, Te[2]);
      ve = sqrt((16.0 * K->e * K->ph0 * Te[2]) / (3.0 * K->pi * K->me)) * K->t0 / K->x0;
      flxi[2] = (1.0 - a) * mui * Ex[2] * NI(i, j) + 2.5e-1 * vi * NI(i, j);
      flxe[2] = -(a * mue[2] * Ex[2] * NE(i, j)) + 2.5e-1 * ve * NE(i, j) - gam * flxi[2];
    } else if (tx == 1) {
      flxi[2] = 0.0;
      flxe[2] = 0.0;
    }
  }
  dfdx = -(Ex[2] - Ex[1]) / s->dlx[i - 1];
  dflxe_x = (flxe[2] - flxe[1]) / s->dlx[i - 1];
  dflxi_x = (flxi[2] - flxi[1]) / s->dlx[i - 1];

  /* y part: lapl_lib.F90:255-372, only for ny > 1 (in one dimension dfdy,
   * dflxi_y and dflxe_y keep their initial zeros, lapl_lib.F90:128-130) */
  if (s->by > 1) {
    if (ty == -1) {
      Ey[1] = 0.0;
      Ey[2] = -(PH(i, j + 1) - PH(i, j)) / dy[j];
    } else if (ty == 1) {
      Ey[1] = -(PH(i, j) - PH(i, j - 1)) / dy[j - 1];
      Ey[2] = 0.0;
    } else if (ty == 3) { /* dielectric wall: lapl_lib.F90:263-265 */
      Ey[1] = -(PH(i, j) - PH(i, j - 1)) / dy[j - 1];
      Ey[2] = -sig;
    } else {
      Ey[1] = -(PH(i, j) - PH(i, j - 1)) / dy[j - 1];
      Ey[2] = -(PH(i, j + 1) - PH(i, j)) / dy[j];
    }
    if (ty == 0) {
      Te[1] = get_Te(NT(i, j - 1), NE(i, j - 1));
      Te[2] = get_Te(NT(i, j), NE(i, j));
      Te[3] = get_Te(NT(i, j + 1), NE(i, j + 1));
      mue[1] = 5e-1 * (get_mue(K, Te[1]) + get_mue(K, Te[2]));
      mue[2] = 5e-1 * (get_mue(K, Te[2]) + get_mue(K, Te[3]));
      De[1] = 5e-1 * (get_De(K, Te[1]) + get_De(K, Te[2]));
      De[2] = 5e-1 * (get_De(K, Te[2]) + get_De(K, Te[3]));
      flxi[1] = get_flux(K, Ey[1], dy[j - 1], 1, mui, Di, NI(i, j - 1), NI(i, j));
      flxe[1] = get_flux(K, Ey[1], dy[j - 1], -1, mue[1], De[1], NE(i, j - 1), NE(i, j));
      flxi[2] = get_flux(K, Ey[2], dy[j], 1, mui, Di, NI(i, j), NI(i, j + 1));
      flxe[2] = get_flux(K, Ey[2], dy[j], -1, mue[2], De[2], NE(i, j), NE(i, j + 1));
    } else if (ty < 0) {
      Te[2] = get_Te(NT(i, j), NE(i, j));
      Te[3] = get_Te(NT(i, j + 1), NE(i, j + 1));
      mue[2] = 5e-1 * (get_mue(K, Te[2]) + get_mue(K, Te[3]));
      De[2] = 5e-1 * (get_De(K, Te[2]) + get_De(K, Te[3]));
      flxi[2] = get_flux(K, Ey[2], dy[j], 1, mui, Di, NI(i, j), NI(i, j + 1));
      flxe[2] = get_flux(K, Ey[2], dy[j], -1, mue[2], De[2], NE(i, j), NE(i, j + 1));
      flxi[1] = 0.0;
      flxe[1] = 0.0;
    } else if (ty > 0) {
      Te[1] = get_Te(NT(i, j - 1), NE(i, j - 1));
      Te[2] = get_Te(NT(i, j), NE(i, j));
      mue[1] = 5e-1 * (get_mue(K, Te[1]) + get_mue(K, Te[2]));
      De[1] = 5e-1 * (get_De(K, Te[1]) + get_De(K, Te[2]));
      flxi[1] = get_flux(K, Ey[1], dy[j - 1], 1, mui, Di, NI(i, j - 1), NI(i, j));
      flxe[1] = get_flux(K, Ey[1], dy[j - 1], -1, mue[1], De[1], NE(i, j - 1), NE(i, j));
      if (ty == 3) { /* lapl_lib.F90:324-335 */
        Te[2] = get_Te(NT(i, j), NE(i, j));
        mue[2] = get_mue(K, Te[2]);
        De[2] = get_De(K, Te[2]);
        flxi[2] = get_flux(K, Ey[2], dy[j], 1, mui, Di, NI(i, j), 0.0);
        flxe[2] = get_flux(K, Ey[2], dy[j], -1, mue[2], De[2], NE(i, j), 0.0);
      } else {
        flxi[2] = 0.0;
        flxe[2] = 0.0;
      }
    }
    if (s->c.cyl) {
      dfdy = -(Ey[2] * (r[j + 1] + r[j]) / 2.0 - Ey[1] * (r[j] + r[j - 1]) / 2.0) / s->dly[j - 1] /
             r[j];
      dflxi_y = (flxi[2] * (r[j + 1] + r[j]) / 2.0 - flxi[1] * (r[j] + r[j - 1]) / 2.0) /
                s->dly[j - 1] / r[j];
      dflxe_y = (flxe[2] * (r[j + 1] + r[j]) / 2.0 - flxe[1] * (r[j] + r[j - 1]) / 2.0) /
                s->dly[j - 1] / r[j];
    } else {
      dfdy = -(Ey[2] - Ey[1]) / s->dly[j - 1];
      dflxi_y = (flxi[2] - flxi[1]) / s->dly[j - 1];
      dflxe_y = (flxe[2] - flxe[1]) / s->dly[j - 1];
    }
  }
  double b = dfdx + dfdy + NI(i, j) - NE(i, j) - s->dt * (dflxi_x + dflxi_y - dflxe_x - dflxe_y);
#undef NE
#undef NI
#undef NT
#undef PH
  return b;
}

/* laplJ: lapl_lib.F90:381-609.  A5 = {W, E, S, N, C} (0 where the reference
 * stores no column). */
static void lapl_j(osim* s, int i, int j, const double* ne, const double* ni, const double* nt,
                   double* A) {
  const oconsts* K = &s->k;
  double Te1, Te2, mue, De, E_pl, E_mi, dfxi_pl, dfxi_mi, dfxe_pl, dfxe_mi, dfyi_pl = 0,
                                                                          dfyi_mi = 0,
                                                                          dfye_pl = 0,
                                                                          dfye_mi = 0;
  const double mui = K->mui, Di = K->Di, gam = K->gam, dt = s->dt;
  const double *dx = s->dx, *dy = s->dy, *r = s->r, *dlx = s->dlx, *dly = s->dly;
  int ac;
#define NE(a, b) A2(ne, a, b)
#define NI(a, b) A2(ni, a, b)
#define NT(a, b) A2(nt, a, b)
#define PH(a, b) A2(s->ph, a, b)
#define PHM(a, b) A2(s->ph_mi, a, b)
  int tx = TX(i - 1, j - 1), ty = TY(i - 1, j - 1);
  A[0] = A[1] = A[2] = A[3] = A[4] = 0.0;
  if (tx >= 0) {
    Te1 = get_Te(NT(i - 1, j), NE(i - 1, j));
    Te2 = get_Te(NT(i, j), NE(i, j));
    mue = 5e-1 * (get_mue(K, Te1) + get_mue(K, Te2));
    De = 5e-1 * (get_De(K, Te1) + get_De(K, Te2));
    E_pl = -(PH(i, j) - PH(i - 1, j)) / dx[i - 1];
    E_mi = -(PHM(i, j) - PHM(i - 1, j)) / dx[i - 1];
    dfxe_mi = dfdph(E_pl, E_mi, -1, mue, De, dx[i - 1], NE(i - 1, j), NE(i, j));
    dfxe_mi = -(mue / dx[i - 1] * dfxe_mi);
    dfxi_mi = dfdph(E_pl, E_mi, 1, mui, Di, dx[i - 1], NI(i - 1, j), NI(i, j));
    dfxi_mi = mui / dx[i - 1] * dfxi_mi;
    A[0] = (1.0 / dx[i - 1] + dt * (dfxi_mi - dfxe_mi)) / dlx[i - 1];
  } else if (tx == -2) {
    ac = (PH(i - 1, j) > PH(i, j)) ? 1 : 0;
    Te2 = get_Te(NT(i, j), NE(i, j));
    mue = get_mue(K, Te2);
    dfxi_mi = (double)(1 - ac) * mui * NI(i, j) / dx[i - 1];
    dfxe_mi = -((double)ac * mue * NE(i, j) / dx[i - 1]) - gam * dfxi_mi;
  } else {
    dfxi_mi = 0.0;
    dfxe_mi = 0.0;
  }
  if (tx <= 0) {
    Te1 = get_Te(NT(i, j), NE(i, j));
    Te2 = get_Te(NT(i + 1, j), NE(i + 1, j));
    mue = 5e-1 * (get_mue(K, Te1) + get_mue(K, Te2));
    De = 5e-1 * (get_De(K, Te1) + get_De(K, Te2));
    E_pl = -(PH(i + 1, j) - PH(i, j)) / dx[i];
    E_mi = -(PHM(i + 1, j) - PHM(i, j)) / dx[i];
    dfxe_pl = dfdph(E_pl, E_mi, -1, mue, De, dx[i], NE(i, j), NE(i + 1, j));
    dfxe_pl = mue / dx[i] * dfxe_pl;
    dfxi_pl = dfdph(E_pl, E_mi, 1, mui, Di, dx[i], NI(i, j), NI(i + 1, j));
    dfxi_pl = -(mui / dx[i] * dfxi_pl);
    A[1] = (1.0 / dx[i] - dt * (dfxi_pl - dfxe_pl)) / dlx[i - 1];
  } else if (tx == 2) {
    ac = (PH(i + 1, j) > PH(i, j)) ? 1 : 0;
    Te2 = get_Te(NT(i, j), NE(i, j));
    mue = get_mue(K, Te2);
    dfxi_pl = (double)(ac - 1) * mui * NI(i, j) / dx[i];
    dfxe_pl = (double)ac * mue * NE(i, j) / dx[i] - gam * dfxi_pl;
  } else {
    dfxi_pl = 0.0;
    dfxe_pl = 0.0;
  }
  /* y: lapl_lib.F90:462-564, only for ny > 1 (one dimension: no y columns) */
  const int twod = s->by > 1;
  if (!twod) {
    dfyi_mi = dfye_mi = dfyi_pl = dfye_pl = 0.0;
  } else if (ty >= 0) {
    Te1 = get_Te(NT(i, j - 1), NE(i, j - 1));
    Te2 = get_Te(NT(i, j), NE(i, j));
    mue = 5e-1 * (get_mue(K, Te1) + get_mue(K, Te2));
    De = 5e-1 * (get_De(K, Te1) + get_De(K, Te2));
    E_pl = -(PH(i, j) - PH(i, j - 1)) / dy[j - 1];
    E_mi = -(PHM(i, j) - PHM(i, j - 1)) / dy[j - 1];
    dfye_mi = dfdph(E_pl, E_mi, -1, mue, De, dy[j - 1], NE(i, j - 1), NE(i, j));
    dfye_mi = -(mue / dy[j - 1] * dfye_mi);
    dfyi_mi = dfdph(E_pl, E_mi, 1, mui, Di, dy[j - 1], NI(i, j - 1), NI(i, j));
    dfyi_mi = mui / dy[j - 1] * dfyi_mi;
    if (s->c.cyl)
      A[2] = 5e-1 * (r[j] + r[j - 1]) / dy[j - 1] / dly[j - 1] / r[j] +
             dt * (dfyi_mi - dfye_mi) * (r[j] + r[j - 1]) / 2.0 / dly[j - 1] / r[j];
    else
      A[2] = 1.0 / dy[j - 1] / dly[j - 1] + dt * (dfyi_mi - dfye_mi) / dly[j - 1];
  } else {
    dfyi_mi = 0.0;
    dfye_mi = 0.0;
  }
  if (!twod) {
  } else if (ty <= 0) {
    Te1 = get_Te(NT(i, j), NE(i, j));
    Te2 = get_Te(NT(i, j + 1), NE(i, j + 1));
    mue = 5e-1 * (get_mue(K, Te1) + get_mue(K, Te2));
    De = 5e-1 * (get_De(K, Te1) + get_De(K, Te2));
    E_pl = -(PH(i, j + 1) - PH(i, j)) / dy[j];
    E_mi = -(PHM(i, j + 1) - PHM(i, j)) / dy[j];
    dfye_pl = dfdph(E_pl, E_mi, -1, mue, De, dy[j], NE(i, j), NE(i, j + 1));
    dfye_pl = mue / dy[j] * dfye_pl;
    dfyi_pl = dfdph(E_pl, E_mi, 1, mui, Di, dy[j], NI(i, j), NI(i, j + 1));
    dfyi_pl = -(mui / dy[j] * dfyi_pl);
    if (s->c.cyl)
      A[3] = 5e-1 * (r[j + 1] + r[j]) / dy[j] / dly[j - 1] / r[j] -
             dt * (dfyi_pl - dfye_pl) * (r[j + 1] + r[j]) / 2.0 / dly[j - 1] / r[j];
    else
      A[3] = 1.0 / dy[j] / dly[j - 1] - dt * (dfyi_pl - dfye_pl) / dly[j - 1];
  } else {
    dfyi_pl = 0.0;
    dfye_pl = 0.0;
  }
  /* diagonal: lapl_lib.F90:566-607 */
  double C = dt * (dfxi_pl - dfxi_mi - dfxe_pl + dfxe_mi) / dlx[i - 1];
  if (tx == -1 || tx == 1)
    C = C - 1.0 / dx[i - 1] / dlx[i - 1];
  else
    C = C - 1.0 / dlx[i - 1] * (1.0 / dx[i] + 1.0 / dx[i - 1]);
  if (twod) { /* lapl_lib.F90:577-607 */
    if (s->c.cyl)
      C = C + dt * ((dfyi_pl - dfye_pl) * (r[j + 1] + r[j]) / 2.0 -
                    (dfyi_mi - dfye_mi) * (r[j] + r[j - 1]) / 2.0) /
                  dly[j - 1] / r[j];
    else
      C = C + dt * (dfyi_pl - dfyi_mi - dfye_pl + dfye_mi) / dly[j - 1];
    if (ty == -1) {
      if (s->c.cyl)
        C = C - 1.0 / dy[j - 1] / dly[j - 1] * (r[j + 1] + r[j]) / 2.0 / r[j];
      else
        C = C - 1.0 / dy[j - 1] / dly[j - 1];
    } else if (ty == 1) {
      if (s->c.cyl)
        C = C - 1.0 / dy[j - 1] / dly[j - 1] * (r[j] + r[j - 1]) / 2.0 / r[j];
      else
        C = C - 1.0 / dy[j - 1] / dly[j - 1];
    } else {
      if (s->c.cyl)
        C = C - 5e-1 * ((r[j + 1] + r[j]) / dy[j] + (r[j] + r[j - 1]) / dy[j - 1]) / dly[j - 1] /
                    r[j];
      else
        C = C - 1.0 / dly[j - 1] * (1.0 / dy[j] + 1.0 / dy[j - 1]);
    }
  }
  A[4] = C;
#undef NE
#undef NI
#undef NT
#undef PH
#undef PHM
}

int o_lapl_assemble(osim* s, const double* ne, const double* ni, const double* nt, double* b,
                    double* A5) {
  int bx = s->bx, by = s->by, nth = s->c.nthreads > 0 ? s->c.nthreads : 1;
  g_nan_dfdph = 0;
  g_neg_density = 0;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int j = 2; j <= by + 1; j++)
    for (int i = 2; i <= bx + 1; i++) {
      size_t n = (size_t)(i - 2) + (size_t)(j - 2) * bx; /* props.F90:229-230 */
      /* sig0 = sig(i) on the dielectric row, else 0 (lapl_lib.F90:72-76); main
       * passes sig_pl (main.F90:80) */
      const double sig0 = TY(i - 1, j - 1) == 3 ? s->sig_pl[i] : 0.0;
      b[n] = lapl_eqn(s, i, j, ne, ni, nt, sig0);
      lapl_j(s, i, j, ne, ni, nt, A5 + 5 * n);
    }
  if (g_nan_dfdph) return -5;
  if (g_neg_density) return -2;
  return 0;
}

/* ---- linear algebra for A x = rhs (A in reference row form) ---- */
static void spmv5(const osim* s, const double* A5, const double* x, double* y) {
  int bx = s->bx, by = s->by, nth = s->c.nthreads > 0 ? s->c.nthreads : 1;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int j = 0; j < by; j++)
    for (int i = 0; i < bx; i++) {
      size_t n = (size_t)j * bx + i;
      const double* a = A5 + 5 * n;
      double v = a[4] * x[n];
      if (i > 0) v += a[0] * x[n - 1];
      if (i < bx - 1) v += a[1] * x[n + 1];
      if (j > 0) v += a[2] * x[n - bx];
      if (j < by - 1) v += a[3] * x[n + bx];
      y[n] = v;
    }
}

/* deterministic dot: fixed row partials summed in order (same for any thread count) */
static double dotp(const osim* s, const double* a, const double* b) {
  int bx = s->bx, by = s->by, nth = s->c.nthreads > 0 ? s->c.nthreads : 1;
  double* part = s->part;
#pragma omp parallel for num_threads(nth) schedule(static)
  for (int j = 0; j < by; j++) {
    double acc = 0.0;
    for (int i = 0; i < bx; i++) acc += a[(size_t)j * bx + i] * b[(size_t)j * bx + i];
    part[j] = acc;
  }
  double t = 0.0;
  for (int j = 0; j < by; j++) t += part[j];
  return t;
}

static double true_resid(const osim* s, const double* A5, const double* x, const double* rhs,
                         double* tmp) {
  size_t N = (size_t)s->bx * s->by;
  spmv5(s, A5, x, tmp);
  double rr = 0, bb = 0;
  for (size_t n = 0; n < N; n++) {
    double d = rhs[n] - tmp[n];
    rr += d * d;
    bb += rhs[n] * rhs[n];
  }
  return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/* banded LU without pivoting (A is an M-matrix by rows) */
static int solve_banded(const osim* s, const double* A5, const double* rhs, double* x) {
  int bx = s->bx, by = s->by;
  size_t N = (size_t)bx * by, W = 2 * (size_t)bx + 1;
  double* band = calloc(N * W, sizeof(double));
  if (!band) return -1;
#define BD(row, col) band[(size_t)(row)*W + ((col) - (row) + bx)]
  for (int j = 0; j < by; j++)
    for (int i = 0; i < bx; i++) {
      size_t n = (size_t)j * bx + i;
      const double* a = A5 + 5 * n;
      BD(n, n) = a[4];
      if (i > 0) BD(n, n - 1) = a[0];
      if (i < bx - 1) BD(n, n + 1) = a[1];
      if (j > 0) BD(n, n - bx) = a[2];
      if (j < by - 1) BD(n, n + bx) = a[3];
    }
  memcpy(x, rhs, N * sizeof(double));
  for (size_t k = 0; k < N; k++) {
    double piv = BD(k, k);
    size_t rmax = k + bx < N - 1 ? k + bx : N - 1;
    for (size_t r = k + 1; r <= rmax; r++) {
      double f = BD(r, k);
      if (f == 0.0) continue;
      f = f / piv;
      for (size_t c = k; c <= rmax; c++) BD(r, c) -= f * BD(k, c);
      x[r] -= f * x[k];
    }
  }
  for (size_t kk = N; kk-- > 0;) {
    double v = x[kk];
    size_t cmax = kk + bx < N - 1 ? kk + bx : N - 1;
    for (size_t c = kk + 1; c <= cmax; c++) v -= BD(kk, c) * x[c];
    x[kk] = v / BD(kk, kk);
  }
#undef BD
  free(band);
  return 0;
}

/* ILU(0) for the 5-point pattern (natural order) */
static void ilu0(const osim* s, const double* A5, double* lW, double* lS, double* uC) {
  int bx = s->bx, by = s->by;
  for (int j = 0; j < by; j++)
    for (int i = 0; i < bx; i++) {
      size_t n = (size_t)j * bx + i;
      const double* a = A5 + 5 * n;
      double c = a[4];
      lW[n] = 0;
      lS[n] = 0;
      if (i > 0) {
        lW[n] = a[0] / uC[n - 1];
        c -= lW[n] * A5[5 * (n - 1) + 1];
      }
      if (j > 0) {
        lS[n] = a[2] / uC[n - bx];
        c -= lS[n] * A5[5 * (n - bx) + 3];
      }
      uC[n] = c;
    }
}
static void ilu_apply(const osim* s, const double* A5, const double* lW, const double* lS,
                      const double* uC, const double* v, double* z) {
  int bx = s->bx, by = s->by;
  size_t N = (size_t)bx * by;
  for (size_t n = 0; n < N; n++) {
    double y = v[n];
    if (n % bx) y -= lW[n] * z[n - 1];
    if (n >= (size_t)bx) y -= lS[n] * z[n - bx];
    z[n] = y;
  }
  for (size_t n = N; n-- > 0;) {
    double y = z[n];
    if ((n + 1) % bx) y -= A5[5 * n + 1] * z[n + 1];
    if (n + bx < N) y -= A5[5 * n + 3] * z[n + bx];
    z[n] = y / uC[n];
  }
}

static int solve_bicgstab_ilu(osim* s, const double* A5, const double* b, double* x,
                              double* iters) {
  size_t N = (size_t)s->bx * s->by;
  double* w = s->wk;
  double *r = w, *r0 = w + N, *p = w + 2 * N, *v = w + 3 * N, *ph = w + 4 * N, *sv = w + 5 * N,
         *tv = w + 6 * N, *sh = w + 7 * N;
  double* lW = malloc(3 * N * sizeof(double));
  double *lS = lW + N, *uC = lW + 2 * N;
  ilu0(s, A5, lW, lS, uC);
  for (size_t n = 0; n < N; n++) {
    x[n] = 0;
    r[n] = b[n];
    r0[n] = b[n];
    p[n] = 0;
    v[n] = 0;
  }
  double bn = sqrt(dotp(s, b, b)), rho = 1, alpha = 1, omega = 1;
  int it, ok = 0;
  if (bn == 0) { free(lW); *iters = 0; return 0; }
  for (it = 1; it <= s->c.maxit; it++) {
    double rho1 = dotp(s, r0, r);
    if (rho1 == 0) break;
    double beta = (rho1 / rho) * (alpha / omega);
    for (size_t n = 0; n < N; n++) p[n] = r[n] + beta * (p[n] - omega * v[n]);
    ilu_apply(s, A5, lW, lS, uC, p, ph);
    spmv5(s, A5, ph, v);
    alpha = rho1 / dotp(s, r0, v);
    for (size_t n = 0; n < N; n++) sv[n] = r[n] - alpha * v[n];
    if (sqrt(dotp(s, sv, sv)) <= s->c.rtol * bn) {
      for (size_t n = 0; n < N; n++) x[n] += alpha * ph[n];
      ok = 1;
      break;
    }
    ilu_apply(s, A5, lW, lS, uC, sv, sh);
    spmv5(s, A5, sh, tv);
    double tt = dotp(s, tv, tv);
    omega = tt > 0 ? dotp(s, tv, sv) / tt : 0.0;
    for (size_t n = 0; n < N; n++) {
      x[n] += alpha * ph[n] + omega * sh[n];
      r[n] = sv[n] - omega * tv[n];
    }
    rho = rho1;
    if (sqrt(dotp(s, r, r)) <= s->c.rtol * bn) { ok = 1; break; }
    if (omega == 0) break;
  }
  *iters = it;
  free(lW);
  return ok ? 0 : 1;
}

/* Jacobi-PCG on the volume-symmetrised system  -V A x = -V rhs */
static int solve_pcg(osim* s, const double* A5, const double* b, double* x, double* iters) {
  int bx = s->bx, by = s->by;
  size_t N = (size_t)bx * by;
  double* w = s->wk;
  double *r = w, *z = w + N, *p = w + 2 * N, *q = w + 3 * N, *M = w + 4 * N, *bs = w + 9 * N;
  for (int j = 0; j < by; j++)
    for (int i = 0; i < bx; i++) {
      size_t n = (size_t)j * bx + i;
      double V = s->dlx[i + 1] * s->dly[j + 1] * (s->c.cyl ? s->r[j + 2] : 1.0);
      for (int c = 0; c < 5; c++) M[5 * n + c] = -V * A5[5 * n + c];
      bs[n] = -V * b[n];
    }
  double bn = sqrt(dotp(s, bs, bs));
  for (size_t n = 0; n < N; n++) {
    x[n] = 0;
    r[n] = bs[n];
    z[n] = r[n] / M[5 * n + 4];
    p[n] = z[n];
  }
  if (bn == 0) { *iters = 0; return 0; }
  double rz = dotp(s, r, z);
  int it, ok = 0;
  for (it = 1; it <= s->c.maxit; it++) {
    spmv5(s, M, p, q);
    double alpha = rz / dotp(s, p, q);
    for (size_t n = 0; n < N; n++) {
      x[n] += alpha * p[n];
      r[n] -= alpha * q[n];
      z[n] = r[n] / M[5 * n + 4];
    }
    if (sqrt(dotp(s, r, r)) <= s->c.rtol * bn) { ok = 1; break; }
    double rz1 = dotp(s, r, z), beta = rz1 / rz;
    rz = rz1;
    for (size_t n = 0; n < N; n++) p[n] = z[n] + beta * p[n];
  }
  *iters = it;
  return ok ? 0 : 1;
}

/* lapl_solve: lapl_lib.F90:57-121 */
int o_lapl_solve(osim* s, int* lerr, double* stats) {
  double t0 = wall();
  size_t N = (size_t)s->bx * s->by;
  const double *ne = s->ne, *ni = s->ni, *nt = s->nt; /* level 1 (main.F90:80) */
  if (!s->wk) s->wk = calloc(11 * N, sizeof(double));
  if (!s->part) s->part = calloc(s->by + 8, sizeof(double));
  int st = o_lapl_assemble(s, ne, ni, nt, s->sol, s->A5);
  if (st < 0) return st;
  for (size_t n = 0; n < N; n++) s->rhs[n] = -s->sol[n]; /* VecSetValues(-b) */
  double iters = 1;
  int fail;
  if (s->c.solver == 0) fail = solve_banded(s, s->A5, s->rhs, s->sol) != 0;
  else if (s->c.solver == 1) fail = solve_bicgstab_ilu(s, s->A5, s->rhs, s->sol, &iters);
  else fail = solve_pcg(s, s->A5, s->rhs, s->sol, &iters);
  double res = true_resid(s, s->A5, s->sol, s->rhs, s->wk + 10 * N);
  if (fail) {
    *lerr = 1; /* lapl_lib.F90:106-108 */
  } else {
    *lerr = 0;
    for (int j = 2; j <= s->by + 1; j++)
      for (int i = 2; i <= s->bx + 1; i++)
        A2(s->ph, i, j) = A2(s->ph, i, j) + s->sol[(size_t)(i - 2) + (size_t)(j - 2) * s->bx];
    memcpy(s->ph_mi, s->ph, s->PL * sizeof(double)); /* lapl_lib.F90:119 */
  }
  s->last_iters = iters;
  s->last_resid = res;
  s->last_time = wall() - t0;
  if (stats) { stats[0] = iters; stats[1] = res; stats[2] = s->last_time; }
  return 0;
}

/* main.F90:83-101 */
void o_accept(osim* s) {
  s->ts = s->ts + 1;
  s->t = s->t + s->dt;
  copy_level(s, 3, 2);
  copy_level(s, 2, 1);
  s->Vd_or = s->Vd_mi;
  s->Vd_mi = s->Vd_pl;
  for (int i = 1; i <= s->bx + 2; i++) { /* main.F90:100-101 */
    s->sig_or[i] = s->sig_mi[i];
    s->sig_mi[i] = s->sig_pl[i];
  }
}

/* main.F90:103-121 */
void o_reject(osim* s) {
  s->dt = s->dt / 2.0;
  copy_level(s, 1, 3);
  copy_level(s, 2, 3);
  s->Vd_pl = s->Vd_or;
  s->Vd_mi = s->Vd_or;
  for (int i = 1; i <= s->bx + 2; i++) s->sig_pl[i] = s->sig_mi[i] = s->sig_or[i];
}

/* n_iters passes of the main loop body: main.F90:68-121 */
int o_run(osim* s, int n_iters, double* times) {
  double tp = 0, tl = 0, t00 = wall();
  for (int it = 0; it < n_iters; it++) {
    s->t_m = F32(1e9);
    if (s->ts > 1) {
      double a = wall();
      int st = o_ptcl_step(s);
      tp += wall() - a;
      if (st < 0) return st;
    }
    int lerr = 0;
    double a = wall();
    int st = o_lapl_solve(s, &lerr, NULL);
    tl += wall() - a;
    if (st < 0) return st;
    if (s->nlog > 0 && s->log[(size_t)(s->nlog - 1) * OREC_LEN + 6] < 0) {
      double* rec = s->log + (size_t)(s->nlog - 1) * OREC_LEN;
      rec[6] = lerr;
      rec[7] = s->last_iters;
      rec[8] = s->last_resid;
    }
    if (lerr == 0) o_accept(s);
    else o_reject(s);
  }
  if (times) { times[0] = tp; times[1] = tl; times[2] = wall() - t00; }
  return 0;
}

void o_set_replay(osim* s, const double* recs, int n) {
  free(s->replay);
  s->replay = NULL;
  s->nreplay = 0;
  s->ireplay = 0;
  if (n > 0) {
    s->replay = malloc((size_t)n * OREC_LEN * sizeof(double));
    memcpy(s->replay, recs, (size_t)n * OREC_LEN * sizeof(double));
    s->nreplay = n;
  }
}
int o_log_count(const osim* s) { return s->nlog; }
int o_get_log(const osim* s, double* out, int max) {
  int n = s->nlog < max ? s->nlog : max;
  memcpy(out, s->log, (size_t)n * OREC_LEN * sizeof(double));
  return n;
}
void o_clear_log(osim* s) { s->nlog = 0; }

/* ---- vectorised unit functions ---- */
void o_consts(double p, double gam, double* out) {
  oconsts k;
  consts_init(&k, p, gam);
  consts_table(&k, out);
}

/* out[12*n + m]: Te-independent of species; columns
 * mue, mut, De, Dt, k_ex, k_ir, nu, k_sc, k_si, beta, mui(E=T), unused */
void o_eval_fits(double p, int n, const double* T, double* out) {
  oconsts k;
  consts_init(&k, p, 0.1);
  for (int m = 0; m < n; m++) {
    double t = T[m], *o = out + 12 * (size_t)m;
    o[0] = get_mue(&k, t);
    o[1] = get_mut(&k, t);
    o[2] = get_De(&k, t);
    o[3] = get_Dt(&k, t);
    o[4] = get_k_ex(&k, t);
    o[5] = get_k_ir(&k, t);
    o[6] = get_nu(&k, t);
    o[7] = get_k_sc(&k, t);
    o[8] = get_k_si(&k, t);
    o[9] = get_beta(&k, t);
    o[10] = get_mui_E(&k, t);
    o[11] = 0;
  }
}

/* args7 = {E, dh, q, mu, D, nl, nr} */
void o_eval_flux(int n, const double* a, double* out) {
  oconsts k;
  consts_init(&k, 10.0, 0.1);
  for (int m = 0; m < n; m++) {
    const double* v = a + 7 * (size_t)m;
    out[m] = get_flux(&k, v[0], v[1], (int)v[2], v[3], v[4], v[5], v[6]);
  }
}

/* args7 = {E, dx, q, mu, D, n1, n2} */
void o_eval_dfdph(int n, const double* a, double* out) {
  for (int m = 0; m < n; m++) {
    const double* v = a + 7 * (size_t)m;
    out[m] = dfdph(v[0], v[0], (int)v[2], v[3], v[4], v[1], v[5], v[6]);
  }
}
