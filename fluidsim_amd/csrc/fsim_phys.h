// fsim_phys.h -- physical constants and per-cell / per-face physics of the
// fluidsim plasma-transport path, as HIP device functions for gfx950.
//
// The Poisson assembly's functions keep the reference's operation order
// (ptcl_props.F90, lapl_lib.F90) exactly, compiled with -ffp-contract=off (no
// FMA contraction): the assembled residual is a small difference of large
// terms, where ulp changes in E or in the mobility show at 1e-9.  The stage
// kernel, bound by FP64 instruction issue, uses shorter forms, each within a
// few ulp of a left-to-right Fortran evaluation (host tests bound each one):
// the fits in Horner form with FMA, its own exp (fs_exp), log (fs_log) and
// Newton-refined division (fs_div, fs_rcp); the rate fits multiply by
// precomputed products of their unit conversions, beta is exp of the cell's
// log; reciprocal grid spacings; one SG exponential shared by electrons and
// electron energy (their exponent arguments are equal up to rounding:
// mut/Dt = mue/De).
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
// Short arithmetic for the stage kernel.  Its cost is FP64 VALU issue: every
// wave64 instruction holds a SIMD for 4 cycles, a coefficient moved into a
// VGPR costs two more, so the forms below keep constants in SGPRs and drop the
// special-value handling that the kernel's argument ranges never need.

// a*b + c with c uniform across the wave (a constant or a kernel-wide value):
// one VOP3 v_fma_f64 reading c from an SGPR pair.  (Left to itself the
// compiler turns a Horner step into v_fmac_f64, whose accumulator must first
// be copied into VGPRs: three instructions instead of one.)  c must not vary
// across lanes.
__host__ __device__ inline double fs_fma_c(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}

// 1/b refined from the hardware estimate (v_rcp_f64) by two Newton steps
// r <- r + r (1 - b r), each squaring the relative error; a/b is then
// q = a r corrected once by q + r (a - b q) (Markstein's final step).  Within
// one ulp for normal operands with a normal quotient (a host test checks the
// iteration from a float-precision seed); no special values.
__host__ __device__ inline double fs_rcp(double b) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(b);
#else
  double r = (double)(1.0f / (float)b);
#endif
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  return r;
}
__host__ __device__ inline double fs_div(double a, double b) {
  const double r = fs_rcp(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}

// exp(y) for |y| < 700: y = k ln2 + r with ln2 split in two (Cody and Waite;
// the high part has 32 significant bits, so k ln2_hi is exact), |r| <= ln2/2,
// exp(r) by its Taylor series through r^13/13! (truncation below 5e-18
// relative) in Horner form, the result 2^k exp(r).  Within about one ulp; no
// overflow, underflow or NaN handling (the stage kernel's arguments lie in
// [-200, 100]).
__host__ __device__ inline double fs_exp(double y) {
  constexpr double kLog2e = 1.4426950408889634;
  constexpr double kLn2Hi = 6.93147180369123816490e-01;
  constexpr double kLn2Lo = 1.90821492927058770002e-10;
  const double k = rint(y * kLog2e);
  double r = fma(-k, kLn2Hi, y);
  r = fma(-k, kLn2Lo, r);
  double p = fs_fma_c(1.0 / 6227020800.0, r, 1.0 / 479001600.0);  // 1/13!, 1/12!
  p = fs_fma_c(p, r, 1.0 / 39916800.0);
  p = fs_fma_c(p, r, 1.0 / 3628800.0);
  p = fs_fma_c(p, r, 1.0 / 362880.0);
  p = fs_fma_c(p, r, 1.0 / 40320.0);
  p = fs_fma_c(p, r, 1.0 / 5040.0);
  p = fs_fma_c(p, r, 1.0 / 720.0);
  p = fs_fma_c(p, r, 1.0 / 120.0);
  p = fs_fma_c(p, r, 1.0 / 24.0);
  p = fs_fma_c(p, r, 1.0 / 6.0);
  p = fs_fma_c(p, r, 0.5);
  p = fs_fma_c(p, r, 1.0);
  p = fs_fma_c(p, r, 1.0);
  return ldexp(p, (int)k);
}

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
  const double s = fs_div(f, 2.0 + f);
  const double z = s * s;
  double q = fs_fma_c(2.0 / 21.0, z, 2.0 / 19.0);
  q = fs_fma_c(q, z, 2.0 / 17.0);
  q = fs_fma_c(q, z, 2.0 / 15.0);
  q = fs_fma_c(q, z, 2.0 / 13.0);
  q = fs_fma_c(q, z, 2.0 / 11.0);
  q = fs_fma_c(q, z, 2.0 / 9.0);
  q = fs_fma_c(q, z, 2.0 / 7.0);
  q = fs_fma_c(q, z, 2.0 / 5.0);
  q = fs_fma_c(q, z, 2.0 / 3.0);
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
__host__ __device__ inline Pow11 pow_chain(double x) {
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
__host__ __device__ inline double mue_poly(const FsPhys& K, double x) {
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
// The stage kernel's forms of the fits: the same polynomials in Horner form
// with FMA (the exponent moves by a few ulp of the largest power-sum term from
// the reference's evaluation order; a host test bounds the difference), fs_exp,
// and one product with the folded unit conversion.
__host__ __device__ inline double mue_poly_h(double x, double log5119) {
  double lo = fs_fma_c(F32(0.0436642742558), x, F32(-0.0637079234081));
  lo = fs_fma_c(lo, x, F32(-0.142058042584));
  lo = fs_fma_c(lo, x, F32(-0.164770900119));
  lo = fs_fma_c(lo, x, F32(-0.984082217962));
  lo = fs_fma_c(lo, x, F32(58.1133016145));
  double hi = fs_fma_c(F32(-0.0259045697422), x, F32(0.486276623664));
  hi = fs_fma_c(hi, x, F32(-3.57242763244));
  hi = fs_fma_c(hi, x, F32(12.8366493242));
  hi = fs_fma_c(hi, x, F32(-22.7558138237));
  hi = fs_fma_c(hi, x, F32(70.7846754548));
  return x < log5119 ? lo : hi;
}
__device__ __forceinline__ double mue_of_x_folded(const FsPhys& K, double x) {
  return fs_exp(mue_poly_h(x, K.log5119)) * K.c_mue;
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
__host__ __device__ inline double de_poly(double x) {
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
__host__ __device__ inline double de_poly_h(double x) {
  double q = fs_fma_c(F32(3.00706533201e-05), x, F32(-0.000590000082557));
  q = fs_fma_c(q, x, F32(0.00426793748856));
  q = fs_fma_c(q, x, F32(-0.0122860884883));
  q = fs_fma_c(q, x, F32(0.000578601715687));
  q = fs_fma_c(q, x, F32(0.0584908312121));
  q = fs_fma_c(q, x, F32(-0.0469171521686));
  q = fs_fma_c(q, x, F32(-0.102862612604));
  q = fs_fma_c(q, x, F32(0.105390295278));
  q = fs_fma_c(q, x, F32(-0.117645585949));
  q = fs_fma_c(q, x, F32(-0.0699892875381));
  return fs_fma_c(q, x, F32(57.7067018813));
}
__device__ __forceinline__ double de_of_x_folded(const FsPhys& K, double x) {
  return fs_exp(de_poly_h(x)) * K.c_de;
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

// The rate fits as the stage kernel evaluates them (Horner with FMA, fs_exp,
// one product with the folded conversion); the reference's power-sum forms
// are kept below for the host test that bounds the difference.
// get_k_ex (ptcl_props.F90:318-345); x = log(max(0.9611, min(157, T*ph0)))
__host__ __device__ inline double kex_poly_h(double x, double log5667) {
  double lo = fs_fma_c(F32(-1.44763942525), x, F32(7.51502551486));
  lo = fs_fma_c(lo, x, F32(-10.5053010711));
  lo = fs_fma_c(lo, x, F32(-1.05188982824));
  lo = fs_fma_c(lo, x, F32(17.1472529739));
  lo = fs_fma_c(lo, x, F32(-54.4513453969));
  double hi = fs_fma_c(F32(0.16403226902), x, F32(-3.0908688844));
  hi = fs_fma_c(hi, x, F32(23.0655081891));
  hi = fs_fma_c(hi, x, F32(-85.5163470769));
  hi = fs_fma_c(hi, x, F32(158.67012941));
  hi = fs_fma_c(hi, x, F32(-150.266369415));
  return x < log5667 ? lo : hi;
}
__device__ __forceinline__ double kex_of_x(const FsPhys& K, double x) {
  return fs_exp(kex_poly_h(x, K.log5667)) * K.c_rate;
}

// get_k_ir (ptcl_props.F90:347-374); x = log(max(1.37, min(157, T*ph0)))
__host__ __device__ inline double kir_poly_h(double x, double log75) {
  double lo = fs_fma_c(F32(-14.9080865672), x, F32(81.9325486193));
  lo = fs_fma_c(lo, x, F32(-161.378167125));
  lo = fs_fma_c(lo, x, F32(138.743842326));
  lo = fs_fma_c(lo, x, F32(-30.1312514721));
  lo = fs_fma_c(lo, x, F32(-73.27080258));
  double hi = fs_fma_c(F32(0.16331188192), x, F32(-3.15463258104));
  hi = fs_fma_c(hi, x, F32(24.1814225482));
  hi = fs_fma_c(hi, x, F32(-92.3047166553));
  hi = fs_fma_c(hi, x, F32(177.113452524));
  hi = fs_fma_c(hi, x, F32(-168.925304837));
  return x < log75 ? lo : hi;
}
__device__ __forceinline__ double kir_of_x(const FsPhys& K, double x) {
  return fs_exp(kir_poly_h(x, K.log75)) * K.c_rate;
}

// get_nu (ptcl_props.F90:376-395)
__host__ __device__ inline double nu_poly_h(double x) {
  double q = fs_fma_c(F32(0.000325615321708), x, F32(-0.00542014733763));
  q = fs_fma_c(q, x, F32(0.0309832277826));
  q = fs_fma_c(q, x, F32(-0.0582934303409));
  q = fs_fma_c(q, x, F32(-0.0414865809044));
  q = fs_fma_c(q, x, F32(0.129397015353));
  q = fs_fma_c(q, x, F32(0.00936933121094));
  q = fs_fma_c(q, x, F32(1.45173283977));
  return fs_fma_c(q, x, F32(-32.275912575));
}
__device__ __forceinline__ double nu_of_x(const FsPhys& K, double x) {
  return fs_exp(nu_poly_h(x)) * K.c_nu;
}

// get_k_sc / get_k_si (ptcl_props.F90:397-429); x = log(min(16, max(0.5, T*ph0)))
__host__ __device__ inline double ksc_poly_h(double x) {
  double q = fs_fma_c(F32(0.10591014464), x, F32(-0.67840685073));
  q = fs_fma_c(q, x, F32(1.27257798891));
  q = fs_fma_c(q, x, F32(-0.555439231606));
  q = fs_fma_c(q, x, F32(0.457356923276));
  return fs_fma_c(q, x, F32(-21.4827864151));
}
__host__ __device__ inline double ksi_poly_h(double x) {
  double q = fs_fma_c(F32(-0.931626144207), x, F32(8.28853856817));
  q = fs_fma_c(q, x, F32(-28.1169537586));
  q = fs_fma_c(q, x, F32(43.9905424566));
  return fs_fma_c(q, x, F32(-43.1347385848));
}
__device__ __forceinline__ double ksc_of_x(const FsPhys& K, double x) {
  return fs_exp(ksc_poly_h(x)) * K.c_rate6;
}
__device__ __forceinline__ double ksi_of_x(const FsPhys& K, double x) {
  return fs_exp(ksi_poly_h(x)) * K.c_rate6;
}

// The reference's power-sum exponents of the rate fits (left to right, powers
// by the power tree), for the host test of the Horner forms above.
__host__ __device__ inline double kex_poly_ref(double x, double log5667) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3;
  if (x < log5667)
    return F32(-54.4513453969) + F32(17.1472529739) * x + F32(-1.05188982824) * x2 +
           F32(-10.5053010711) * x3 + F32(7.51502551486) * x4 + F32(-1.44763942525) * x5;
  return F32(-150.266369415) + F32(158.67012941) * x + F32(-85.5163470769) * x2 +
         F32(23.0655081891) * x3 + F32(-3.0908688844) * x4 + F32(0.16403226902) * x5;
}
__host__ __device__ inline double kir_poly_ref(double x, double log75) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3;
  if (x < log75)
    return F32(-73.27080258) + F32(-30.1312514721) * x + F32(138.743842326) * x2 +
           F32(-161.378167125) * x3 + F32(81.9325486193) * x4 + F32(-14.9080865672) * x5;
  return F32(-168.925304837) + F32(177.113452524) * x + F32(-92.3047166553) * x2 +
         F32(24.1814225482) * x3 + F32(-3.15463258104) * x4 + F32(0.16331188192) * x5;
}
__host__ __device__ inline double nu_poly_ref(double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3, x6 = x3 * x3, x7 = x3 * x4,
         x8 = x4 * x4;
  return F32(-32.275912575) + F32(1.45173283977) * x + F32(0.00936933121094) * x2 +
         F32(0.129397015353) * x3 + F32(-0.0414865809044) * x4 + F32(-0.0582934303409) * x5 +
         F32(0.0309832277826) * x6 + F32(-0.00542014733763) * x7 + F32(0.000325615321708) * x8;
}
__host__ __device__ inline double ksc_poly_ref(double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x2 * x3;
  return F32(-21.4827864151) + F32(0.457356923276) * x + F32(-0.555439231606) * x2 +
         F32(1.27257798891) * x3 + F32(-0.67840685073) * x4 + F32(0.10591014464) * x5;
}
__host__ __device__ inline double ksi_poly_ref(double x) {
  double x2 = x * x, x3 = x2 * x, x4 = x2 * x2;
  return F32(-43.1347385848) + F32(43.9905424566) * x + F32(-28.1169537586) * x2 +
         F32(8.28853856817) * x3 + F32(-0.931626144207) * x4;
}

// get_beta (ptcl_props.F90:182-192), v = T*ph0, xa = fit_x(v):
// 10**(0.944 - 0.665 (log10(w) - 2.48)) with w = v 11604.505/1.5 clamped at
// 5e4 is exp(beta_a - 0.665 ln v) below the clamp, and ln v is xa there unless
// v < 0.234 (fit_x's lower clamp).  Above the clamp the value is a constant.
__device__ __forceinline__ double beta_of_v(const FsPhys& K, double v, double xa) {
  double w = v * (11.604505e3 / F32(1.5));
  if (!(w < 5e4)) return K.beta_hi;
  double lv = v >= 2.34e-1 ? xa : log(v);
  return fs_exp(K.beta_a - 0.665 * lv) * K.c_beta;
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
  double om = 1.0 - ex;
  // |arg| below 2^-53: the reference divides by zero here and stops on the
  // NaN; the a -> 0 limit of the flux is Fick's
  return om == 0.0 ? D * (nl - nr) * rh : v * num / om;
}

// The stage kernel's SG fluxes, with fs_exp and the Newton-refined division
// (the face's 1 - exp(-|arg|) carries their ulp as eps/|arg| relative, as it
// carries the library exp's).
// Ion flux (qs = +1, mu = mui, D = Di) with the exponent argument formed by
// multiplying by 1/Di.
__host__ __device__ inline double sg_flux_ion(const FsPhys& K, double E, double h, double rh,
                                              double nl, double nr) {
  double v = K.mui * E;
  double arg = v * h * K.r_Di;
  if (fabs(E) < K.ftol) return K.Di * (nl - nr) * rh;
  double na = arg > 0 ? -arg : arg;
  double ex = na < -100.0 ? K.exp_m100 : fs_exp(na);
  double num = arg > 0 ? (nl - nr * ex) : (nr - nl * ex);
  const double om = 1.0 - ex;
  return om == 0.0 ? K.Di * (nl - nr) * rh : fs_div(v * num, om);  // a -> 0: Fick
}

// Electron and electron-energy fluxes of one face (qs = -1; mut = c53 mue,
// Dt = c53 De): one exponential of the electron argument for both.
__host__ __device__ inline void sg_flux_et(const FsPhys& K, double E, double h, double rh,
                                           double mue, double de, double mut, double dt,
                                           double neL, double neR, double ntL, double ntR,
                                           double& fe, double& ft) {
  if (fabs(E) < K.ftol) {
    fe = de * (neL - neR) * rh;
    ft = dt * (ntL - ntR) * rh;
    return;
  }
  const double ve = -(mue * E), vt = -(mut * E);
  const double arg = fs_div(ve * h, de);
  const double na = arg > 0 ? -arg : arg;
  const double ex = na < -100.0 ? K.exp_m100 : fs_exp(na);
  const double om1 = 1.0 - ex;
  const double om = fs_rcp(om1);
  fe = ve * (arg > 0 ? (neL - neR * ex) : (neR - neL * ex)) * om;
  ft = vt * (arg > 0 ? (ntL - ntR * ex) : (ntR - ntL * ex)) * om;
  if (om1 == 0.0) {  // |arg| below 2^-53: the a -> 0 (Fick) limit
    fe = de * (neL - neR) * rh;
    ft = dt * (ntL - ntR) * rh;
  }
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
  if (fabs(qs * E) < K.ftol || 1.0 - ex == 0.0) {
    // (|arg| below 2^-53: the reference divides by zero and stops; the
    // a -> 0 limit is Fick's flux)
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
    if (om == 0.0) df = 5e-1 * (nl + nr);  // the a -> 0 limit of dF/dv
  }
}
