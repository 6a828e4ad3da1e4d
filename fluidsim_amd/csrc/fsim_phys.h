// fsim_phys.h -- physical constants and per-cell / per-face physics of the
// fluidsim plasma-transport path, as HIP device functions for gfx950.
//
// The functions keep the reference's operation order (ptcl_props.F90,
// lapl_lib.F90), compiled with -ffp-contract=off (no FMA contraction), except
// where the stage kernel trades a division or a transcendental for a
// multiplication, each a difference of a few ulp from a left-to-right Fortran
// evaluation: the rate fits multiply by precomputed products of their unit
// conversions, beta uses exp10 for 10**y, and the stage kernel multiplies by
// reciprocal grid spacings and shares one SG exponential between electrons and
// electron energy (their exponent arguments are equal up to rounding:
// mut/Dt = mue/De).  The stage kernel's logarithm is fs_log below (within one
// ulp, like the library's).  The Poisson assembly and the mobility/diffusion fits keep
// the reference's order exactly: the assembled residual is a small difference
// of large terms, where ulp changes in E or in the mobility show at 1e-9.
//
// Hoisting (bit-identical to the reference's repeated calls, because the same
// expression on the same operands gives the same value):
//  * mue/De are evaluated once per cell and averaged per face, instead of the
//    reference's 8+8 calls per cell (ptcl_lib.F90:131-135,307-311);
//  * the rate fits share the cell's log(clamp(T*ph0, 0.234, 157)) whenever
//    their own clamp does not bind; when it binds the argument is a constant
//    whose log is taken on the host with the same libm as the CPU oracle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

struct FsPhys {
  double pi, eps0, mu0, c0, e, kb, me;
  double x0, ph0, t0, wp02, w2;
  double Tg, p, ninf, n_zero, n_init;
  double mi, mui, vi, Di, Dm, k_r, k_mp, k_2q, k_3q;
  double H_ir, H_ex, H_si, H_sc, gam;
  double log5119, log5667, log75, ftol;
  double Cap, x0_2, x0_3, x0_6, kTe, nmin_t;
  // products the reference folds at compile time (parameter operands only)
  double c53;         // 5d0/3d0                     (ptcl_lib.F90:142)
  double ve_num;      // 16d0*e*ph0                  (ptcl_lib.F90:198)
  double ve_den;      // 3d0*pi*me
  double qvi;         // 2.5d-1*vi                   (ptcl_lib.F90:202)
  double hvi;         // 5d-1*vi                     (ptcl_lib.F90:212)
  double k2q_ninf;    // k_2q*ninf                   (ptcl_lib.F90:501)
  double k3q_ninf2;   // k_3q*ninf**2                (ptcl_lib.F90:502)
  double kmp2;        // 2d0*k_mp                    (ptcl_lib.F90:500)
  // host-libm logs of the clamp bounds that bind in the rate fits
  double log_kex_lo, log_kir_lo, log_s_lo, log_s_hi;
  double exp_m100;    // exp(-100), the SG exponent clamp (ptcl_props.F90:159)
  // products of the rate fits' unit conversions, formed once (the fits
  // multiply exp(poly) by them instead of dividing by x0^3, 1e6 each time)
  double c_rate;      // t0 / x0^3              (k_ex, k_ir)
  double c_rate6;     // t0 / (1e6 x0^3)        (k_sc, k_si)
  double c_nu;        // ninf * t0 / x0^3       (nu)
  double c_beta;      // 1e-13 * t0 / x0^3      (beta)
  double me_mi;       // me / mi
  // the stage kernel's folded forms (each a few ulp from the reference's
  // left-to-right evaluation)
  double c_mue;       // x0 / ninf * t0 * ph0   (get_mue's unit conversion)
  double c_de;        // x0 / ninf * t0         (get_De's)
  double r_Di;        // 1 / Di                 (ion SG exponent)
  double beta_a;      // ln10 (0.944 + 0.665 * 2.48) - 0.665 ln(11604.505 / 1.5)
  double beta_hi;     // beta at the w = 5e4 clamp
};

// Host: fill FsPhys for runtime pressure and gamma (props.F90:24-44,
// ptcl_props.F90:5-25; the reference has p=10 and gam=0.1 as parameters).
void fs_phys_init(FsPhys* K, double p_torr, double gam);
void fs_phys_table(const FsPhys* K, double* out40);

// float32-widened literal, as gfortran widens a default-real literal
#define F32(v) ((double)(v##f))

// ---------------------------------------------------------------------------
// log for the stage kernel's arguments in fewer operations than the general
// library function (no special cases: normal positive arguments); within one
// ulp of the exact value and correctly rounded for nearly all arguments (a CPU
// test compares it with long-double libm).
//
// log: v = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1 (exact); with
// s = f / (2 + f), log m = 2 artanh(s) = f - (f^2/2 - s (f^2/2 + R)),
// R = 2 s^2/3 + 2 s^4/5 + ... (the Taylor series of artanh to s^21; |s| <= 0.1716,
// truncation below 2^-56 relative), so the leading term f is exact and the
// rounding errors sit in a correction at most a fifth of the result;
// log v = e ln2 + log m (the order of the sums as in Sun's fdlibm).  For normal
// v > 0.
__host__ __device__ inline double fs_log(double v) {
  constexpr double kLn2Hi = 0.6931471803691238;
  constexpr double kLn2Lo = 1.9082149292705877e-10;
  int e;
  double m = frexp(v, &e);  // [0.5, 1)
  if (m < 0.7071067811865476) {
    m = m * 2.0;
    e = e - 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double q = 2.0 / 21.0;
  q = fma(q, z, 2.0 / 19.0);
  q = fma(q, z, 2.0 / 17.0);
  q = fma(q, z, 2.0 / 15.0);
  q = fma(q, z, 2.0 / 13.0);
  q = fma(q, z, 2.0 / 11.0);
  q = fma(q, z, 2.0 / 9.0);
  q = fma(q, z, 2.0 / 7.0);
  q = fma(q, z, 2.0 / 5.0);
  q = fma(q, z, 2.0 / 3.0);
  const double R = z * q;
  const double hfsq = 0.5 * f * f;
  const double de = (double)e;
  // e ln2_lo joins the correction before f, and e ln2_hi (exact) comes last:
  // log v = e ln2_hi - ((f^2/2 - (s (f^2/2 + R) + e ln2_lo)) - f)
  return de * kLn2Hi - ((hfsq - (s * (hfsq + R) + de * kLn2Lo)) - f);
}

// ---------------------------------------------------------------------------
// integer powers by gfortran's optimal power tree: x^n = x^(n-t) * x^t
struct Pow11 {
  double x, x2, x3, x4, x5, x6, x7, x8, x9, x10, x11;
};
__device__ __forceinline__ Pow11 pow_chain(double x) {
  Pow11 q;
  q.x = x;
  q.x2 = x * x;
  q.x3 = q.x2 * x;
  q.x4 = q.x2 * q.x2;
  q.x5 = q.x2 * q.x3;
  q.x6 = q.x3 * q.x3;
  q.x7 = q.x3 * q.x4;
  q.x8 = q.x4 * q.x4;
  q.x9 = q.x3 * q.x6;
  q.x10 = q.x5 * q.x5;
  q.x11 = q.x5 * q.x6;
  return q;
}

// log(max(0.234, min(157, T*ph0)))  (ptcl_props.F90:211)
__device__ __forceinline__ double fit_x(double v) {
  double m = v < 1.57e2 ? v : 1.57e2;
  return log(2.34e-1 > m ? 2.34e-1 : m);
}
// the stage kernel's form (fs_log)
__device__ __forceinline__ double fit_x_stage(double v) {
  double m = v < 1.57e2 ? v : 1.57e2;
  return fs_log(2.34e-1 > m ? 2.34e-1 : m);
}

// get_mue from x (ptcl_props.F90:194-221): exponent, then the value
__device__ __forceinline__ double mue_poly(const FsPhys& K, double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3;
  if (x < K.log5119)
    return F32(58.1133016145) + F32(-0.984082217962) * x + F32(-0.164770900119) * x2 +
           F32(-0.142058042584) * x3 + F32(-0.0637079234081) * x4 + F32(0.0436642742558) * x5;
  return F32(70.7846754548) + F32(-22.7558138237) * x + F32(12.8366493242) * x2 +
         F32(-3.57242763244) * x3 + F32(0.486276623664) * x4 + F32(-0.0259045697422) * x5;
}
__device__ __forceinline__ double mue_of_x(const FsPhys& K, double x) {
  return exp(mue_poly(K, x)) * K.x0 / K.ninf * K.t0 * K.ph0;
}
// the stage kernel's form: one product with the folded conversion
__device__ __forceinline__ double mue_of_x_folded(const FsPhys& K, double x) {
  return exp(mue_poly(K, x)) * K.c_mue;
}

// get_mut from x (ptcl_props.F90:223-250)
__device__ __forceinline__ double mut_of_x(const FsPhys& K, double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3;
  double s;
  if (x < K.log5119)
    s = F32(58.1354622038) + F32(-1.29918391109) * x + F32(-0.106559658509) * x2 +
        F32(-0.0202265418394) * x3 + F32(-0.0286667333915) * x4 + F32(0.022457210094) * x5;
  else
    s = F32(68.2627663433) + F32(-19.3524000831) * x + F32(11.3251980583) * x2 +
        F32(-3.24070386509) * x3 + F32(0.449848156054) * x4 + F32(-0.0242585762405) * x5;
  return exp(s) * K.x0 / K.ninf * K.t0 * K.ph0;
}

// get_De from x (ptcl_props.F90:268-291)
__device__ __forceinline__ double de_poly(double x) {
  Pow11 q = pow_chain(x);
  return F32(57.7067018813) + F32(-0.0699892875381) * x + F32(-0.117645585949) * q.x2 +
         F32(0.105390295278) * q.x3 + F32(-0.102862612604) * q.x4 +
         F32(-0.0469171521686) * q.x5 + F32(0.0584908312121) * q.x6 +
         F32(0.000578601715687) * q.x7 + F32(-0.0122860884883) * q.x8 +
         F32(0.00426793748856) * q.x9 + F32(-0.000590000082557) * q.x10 +
         F32(3.00706533201e-05) * q.x11;
}
__device__ __forceinline__ double de_of_x(const FsPhys& K, double x) {
  return exp(de_poly(x)) * K.x0 / K.ninf * K.t0;
}
__device__ __forceinline__ double de_of_x_folded(const FsPhys& K, double x) {
  return exp(de_poly(x)) * K.c_de;
}

// get_Dt from x (ptcl_props.F90:293-316)
__device__ __forceinline__ double dt_of_x(const FsPhys& K, double x) {
  Pow11 q = pow_chain(x);
  double s = F32(57.7214952841) + F32(-0.353731464348) * x + F32(-0.00505156731795) * q.x2 +
             F32(0.161173720584) * q.x3 + F32(-0.196610345869) * q.x4 +
             F32(-0.0719115643218) * q.x5 + F32(0.11874085868) * q.x6 +
             F32(-0.00669712724784) * q.x7 + F32(-0.0236445308504) * q.x8 +
             F32(0.00917326671764) * q.x9 + F32(-0.00135453096483) * q.x10 +
             F32(7.26379684461e-05) * q.x11;
  return exp(s) * K.x0 / K.ninf * K.t0;
}

// get_k_ex (ptcl_props.F90:318-345); x = log(max(0.9611, min(157, T*ph0)))
__device__ __forceinline__ double kex_of_x(const FsPhys& K, double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3;
  double s;
  if (x < K.log5667)
    s = F32(-54.4513453969) + F32(17.1472529739) * x + F32(-1.05188982824) * x2 +
        F32(-10.5053010711) * x3 + F32(7.51502551486) * x4 + F32(-1.44763942525) * x5;
  else
    s = F32(-150.266369415) + F32(158.67012941) * x + F32(-85.5163470769) * x2 +
        F32(23.0655081891) * x3 + F32(-3.0908688844) * x4 + F32(0.16403226902) * x5;
  return exp(s) * K.c_rate;
}

// get_k_ir (ptcl_props.F90:347-374); x = log(max(1.37, min(157, T*ph0)))
__device__ __forceinline__ double kir_of_x(const FsPhys& K, double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3;
  double s;
  if (x < K.log75)
    s = F32(-73.27080258) + F32(-30.1312514721) * x + F32(138.743842326) * x2 +
        F32(-161.378167125) * x3 + F32(81.9325486193) * x4 + F32(-14.9080865672) * x5;
  else
    s = F32(-168.925304837) + F32(177.113452524) * x + F32(-92.3047166553) * x2 +
        F32(24.1814225482) * x3 + F32(-3.15463258104) * x4 + F32(0.16331188192) * x5;
  return exp(s) * K.c_rate;
}

// get_nu (ptcl_props.F90:376-395).  The reference's real exponents x**3. ..
// x**8. are libm pow() calls; here they are power-tree products (<= 2 ulp apart).
__device__ __forceinline__ double nu_of_x(const FsPhys& K, double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3, x6 = x3 * x3, x7 = x3 * x4,
         x8 = x4 * x4;
  double s = F32(-32.275912575) + F32(1.45173283977) * x + F32(0.00936933121094) * x2 +
             F32(0.129397015353) * x3 + F32(-0.0414865809044) * x4 +
             F32(-0.0582934303409) * x5 + F32(0.0309832277826) * x6 +
             F32(-0.00542014733763) * x7 + F32(0.000325615321708) * x8;
  return exp(s) * K.c_nu;
}

// get_k_sc / get_k_si (ptcl_props.F90:397-429); x = log(min(16, max(0.5, T*ph0)))
__device__ __forceinline__ double ksc_of_x(const FsPhys& K, double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3;
  double s = F32(-21.4827864151) + F32(0.457356923276) * x + F32(-0.555439231606) * x2 +
             F32(1.27257798891) * x3 + F32(-0.67840685073) * x4 + F32(0.10591014464) * x5;
  return exp(s) * K.c_rate6;
}
__device__ __forceinline__ double ksi_of_x(const FsPhys& K, double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
  double s = F32(-43.1347385848) + F32(43.9905424566) * x + F32(-28.1169537586) * x2 +
             F32(8.28853856817) * x3 + F32(-0.931626144207) * x4;
  return exp(s) * K.c_rate6;
}

// get_beta (ptcl_props.F90:182-192), v = T*ph0, xa = fit_x(v):
// 10**(0.944 - 0.665 (log10(w) - 2.48)) with w = v 11604.505/1.5 clamped at
// 5e4 is exp(beta_a - 0.665 ln v) below the clamp, and ln v is xa there unless
// v < 0.234 (fit_x's lower clamp).  Above the clamp the value is a constant.
__device__ __forceinline__ double beta_of_v(const FsPhys& K, double v, double xa) {
  double w = v * (11.604505e3 / F32(1.5));
  if (!(w < 5e4)) return K.beta_hi;
  double lv = v >= 2.34e-1 ? xa : log(v);
  return exp(K.beta_a - 0.665 * lv) * K.c_beta;
}

// The six reaction-rate fits at one cell from v = T*ph0 and xa = fit_x(v).
struct Rates {
  double k_ir, k_sc, k_si, k_ex, nu, beta;
};
__device__ __forceinline__ Rates cell_rates(const FsPhys& K, double v, double xa) {
  Rates R;
  double x_ex = v >= 9.611e-1 ? xa : K.log_kex_lo;
  double x_ir = v >= 1.37 ? xa : K.log_kir_lo;
  double x_s = v < 5e-1 ? K.log_s_lo : (v > 16.0 ? K.log_s_hi : xa);
  R.k_ir = kir_of_x(K, x_ir);
  R.k_sc = ksc_of_x(K, x_s);
  R.k_si = ksi_of_x(K, x_s);
  R.k_ex = kex_of_x(K, x_ex);
  R.nu = nu_of_x(K, xa);
  R.beta = beta_of_v(K, v, xa);
  return R;
}

// Scharfetter-Gummel face flux (ptcl_props.F90:142-165); qs = +1 ions, -1
// electrons/energy.  One exponential: exp(max(-|arg|, -100)).
// rh = 1/h.
__device__ __forceinline__ double sg_flux(const FsPhys& K, double E, double h, double rh,
                                          double qs, double mu, double D, double nl, double nr) {
  double v = qs * mu * E;
  double arg = v * h / D;
  if (fabs(qs * E) < K.ftol) return D * (nl - nr) * rh;
  double na = arg > 0 ? -arg : arg;
  double ex = na < -100.0 ? K.exp_m100 : exp(na);
  double num = arg > 0 ? (nl - nr * ex) : (nr - nl * ex);
  return v * num / (1.0 - ex);
}

// Ion flux (qs = +1, mu = mui, D = Di) with the exponent argument formed by
// multiplying by 1/Di.
__device__ __forceinline__ double sg_flux_ion(const FsPhys& K, double E, double h, double rh,
                                              double nl, double nr) {
  double v = K.mui * E;
  double arg = v * h * K.r_Di;
  if (fabs(E) < K.ftol) return K.Di * (nl - nr) * rh;
  double na = arg > 0 ? -arg : arg;
  double ex = na < -100.0 ? K.exp_m100 : exp(na);
  double num = arg > 0 ? (nl - nr * ex) : (nr - nl * ex);
  return v * num / (1.0 - ex);
}

// Electron and electron-energy fluxes of one face (qs = -1; mut = c53 mue,
// Dt = c53 De): one exponential of the electron argument for both.
__device__ __forceinline__ void sg_flux_et(const FsPhys& K, double E, double h, double rh,
                                           double mue, double de, double mut, double dt,
                                           double neL, double neR, double ntL, double ntR,
                                           double& fe, double& ft) {
  if (fabs(E) < K.ftol) {
    fe = de * (neL - neR) * rh;
    ft = dt * (ntL - ntR) * rh;
    return;
  }
  const double ve = -(mue * E), vt = -(mut * E);
  const double arg = ve * h / de;
  const double na = arg > 0 ? -arg : arg;
  const double ex = na < -100.0 ? K.exp_m100 : exp(na);
  const double om = 1.0 / (1.0 - ex);
  fe = ve * (arg > 0 ? (neL - neR * ex) : (neR - neL * ex)) * om;
  ft = vt * (arg > 0 ? (ntL - ntR * ex) : (ntR - ntL * ex)) * om;
}

// SG flux and its derivative w.r.t. the drift velocity (dFdPh,
// lapl_lib.F90:611-636) sharing one exponential; E_pl == E_mi at every face
// the reference evaluates (ph_mi == ph after each solve, lapl_lib.F90:119).
__device__ __forceinline__ void sg_flux_dfdph(const FsPhys& K, double E, double h, double qs,
                                              double mu, double D, double nl, double nr,
                                              double& flux, double& df) {
  double v = qs * mu * E;
  double arg = v * h / D;
  double na = arg > 0 ? -arg : arg;
  double ef = exp(na);                         // dFdPh: exp(-|a|), unclamped
  double ex = na < -100.0 ? K.exp_m100 : ef;   // get_Flux: exp(max(-|a|, -100))
  if (fabs(qs * E) < K.ftol) {
    flux = D * (nl - nr) / h;
  } else {
    double num = arg > 0 ? (nl - nr * ex) : (nr - nl * ex);
    flux = v * num / (1.0 - ex);
  }
  if (fabs(v) < 1e-12) {
    df = 0.0;
  } else {
    double om = 1.0 - ef;
    double t1 = arg > 0 ? (nl - nr * ef) / om : (nr - nl * ef) / om;
    df = t1 + v * ef * (nr - nl) / (om * om) * h / D;
  }
}
