/*
 * kinc_math.h — the one FP32 exp/log/sqrt implementation shared by the HIP
 * kernels and the CPU code, so that both produce identical bits.
 *
 * The reference calls unqualified exp/log/sqrt on float arguments
 * (src/core/pairwise_gmm.cpp:46,177,180,216,219,262,271,277,356,391;
 *  src/core/pairwise_linalg.cpp:103; src/core/pairwise_pearson.cpp:48).
 * Those resolve to the FP32 overloads here.  exp and log follow published
 * methods:
 *   - expf: Cody-Waite reduction x = g + n*ln2 (W. J. Cody and W. Waite,
 *           "Software Manual for the Elementary Functions", 1980) with ln2
 *           split as in Cephes, C1 = 0.693359375, C2 = -2.12194440e-4; then
 *           e^g = 1 + g (1 + g P(g)) on |g| <= 0.3467 with P of degree 4,
 *           its coefficients from a Remez exchange for the relative error of
 *           e^g (3.1e-9 before rounding the coefficients to float), evaluated
 *           in Horner form with six FMAs.  Against the correctly rounded
 *           result, over every 7th float in [-104, 89]: at most 0.88 ulp,
 *           0.47% of results not correctly rounded.
 *   - logf: Cephes (S. L. Moshier, "Methods and Programs for Mathematical
 *           Functions", 1989; logf.c): x = m * 2^e with m in
 *           [sqrt(1/2), sqrt(2)), the degree-9 minimax polynomial of logf.c
 *           for log(1+f), and the same ln2 split.
 * Only correctly rounded IEEE-754 single-precision operations are used:
 * +, -, *, explicit fused multiply-add (fmaf), ldexp, frexp, rint and
 * comparisons; no hardware transcendental.  Each of these has exactly one
 * correct result, so the host CPU (gcc, FMA3) and gfx950 (v_fma_f32,
 * v_ldexp_f32, v_frexp_*, v_rndne_f32) produce identical bits as long as
 * implicit FP contraction stays off and denormals are preserved.  sqrt is
 * the correctly rounded IEEE sqrtf.
 *
 * Accuracy against glibc (checked by the tests): expf and logf within 2 ulp.
 */
#ifndef KINC_MATH_H
#define KINC_MATH_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP_DEVICE_COMPILE__)
#include <hip/hip_runtime.h>
#define KINC_HD __host__ __device__ __forceinline__
#else
#define KINC_HD static inline
#endif

KINC_HD float kinc_as_float(uint32_t u)
{
   float f;
   memcpy(&f, &u, sizeof(f));
   return f;
}

KINC_HD uint32_t kinc_as_uint(float f)
{
   uint32_t u;
   memcpy(&u, &f, sizeof(u));
   return u;
}

/* 2^n for -126 <= n <= 127, built from the exponent bits (exact). */
KINC_HD float kinc_pow2i(int n)
{
   return kinc_as_float((uint32_t)(n + 127) << 23);
}

/* p * 2^n for -150 <= n <= 128 with one IEEE rounding (ldexp).  The host
 * form keeps the first product normal and exact, so the second applies the
 * only rounding, which is the correctly rounded ldexp result. */
KINC_HD float kinc_ldexpf(float p, int n)
{
#if defined(__HIP_DEVICE_COMPILE__)
   return __builtin_amdgcn_ldexpf(p, n);
#else
   int n1 = (n > 127) ? 127 : ((n < -126) ? n + 126 : n);
   int n2 = (n > 127) ? n - 127 : ((n < -126) ? -126 : 0);
   return (p * kinc_pow2i(n1)) * kinc_pow2i(n2);
#endif
}

KINC_HD float kinc_fmaf(float a, float b, float c)
{
#if defined(__HIP_DEVICE_COMPILE__)
   return __builtin_fmaf(a, b, c);
#else
   return fmaf(a, b, c);
#endif
}

KINC_HD float kinc_rintf(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
   return __builtin_rintf(x);
#else
   return rintf(x);
#endif
}

/* x = m * 2^e with m in [0.5, 1) for finite nonzero x (exact). */
KINC_HD float kinc_frexpf(float x, int* e)
{
#if defined(__HIP_DEVICE_COMPILE__)
   *e = __builtin_amdgcn_frexp_expf(x);
   return __builtin_amdgcn_frexp_mantf(x);
#else
   return frexpf(x, e);
#endif
}

/* min(max(x, lo), hi) for x that is not NaN (a NaN x gives an unspecified
 * value: callers replace the result for NaN input). */
KINC_HD float kinc_clampf(float x, float lo, float hi)
{
#if defined(__HIP_DEVICE_COMPILE__)
   return __builtin_amdgcn_fmed3f(x, lo, hi);
#else
   return fminf(fmaxf(x, lo), hi);
#endif
}

/* The integer value of v, an integral float in [-150, 128], or 0 when v is
 * NaN.  On the device this is the hardware conversion v_cvt_i32_f32, whose
 * NaN result (0) the ISA defines; it is written as an instruction because a
 * C++ float-to-int conversion of NaN is undefined (poison to the compiler),
 * and kinc_expf_le0 lets NaN reach here.  Not a memory access. */
KINC_HD int kinc_f2i(float v)
{
#if defined(__HIP_DEVICE_COMPILE__)
   int r;
   asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(v));
   return r;
#else
   return (v != v) ? 0 : (int)v;
#endif
}

/* The arithmetic of kinc_expf after the argument is clamped. */
KINC_HD float kinc_expf_reduced(float xc)
{
   const float LOG2EF = 1.44269504088896341f;
   const float C1 = 0.693359375f;
   const float C2 = -2.12194440e-4f;
   float nf = kinc_rintf(xc * LOG2EF);
   float g = kinc_fmaf(-nf, C1, xc);
   g = kinc_fmaf(-nf, C2, g);
   /* e^g = 1 + g (1 + g P(g)), P of degree 4: six dependent FMAs */
   float q = 1.3814560370519757e-3f;
   q = kinc_fmaf(q, g, 8.368735201656818e-3f);
   q = kinc_fmaf(q, g, 4.166838899254799e-2f);
   q = kinc_fmaf(q, g, 1.666652113199234e-1f);
   q = kinc_fmaf(q, g, 4.999999403953552e-1f);
   q = kinc_fmaf(q, g, 1.0f);
   float p = kinc_fmaf(q, g, 1.0f);
   return kinc_ldexpf(p, kinc_f2i(nf));    /* -150 <= nf <= 128, or NaN */
}

KINC_HD float kinc_expf(float x)
{
   /* Outside [-104, 89] the result is 0 or +inf; clamping there lets the
    * final ldexp produce them with no extra case (exp(-104) < 2^-150 rounds
    * to 0, exp(89) overflows). */
#if defined(__HIP_DEVICE_COMPILE__)
   /* NaN-propagating clamp (v_minimum3_f32 / v_maximum3_f32): NaN stays NaN
    * through the reduction */
   return kinc_expf_reduced(__builtin_elementwise_maximum(__builtin_elementwise_minimum(x, 89.0f), -104.0f));
#else
   const float r = kinc_expf_reduced(kinc_clampf(x, -104.0f, 89.0f));
   return (x != x) ? x : r;
#endif
}

/* kinc_expf for x <= 0 or NaN, the arguments v - max(v) of a log-sum-exp.
 * On the device: the lower clamp alone, as a NaN-propagating maximum
 * (v_maximum3_f32); the NaN then stays NaN through the reduction (kinc_f2i
 * maps it to 0, and ldexp keeps the NaN).  The same value as kinc_expf for
 * such x, which the host evaluates directly. */
KINC_HD float kinc_expf_le0(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
   return kinc_expf_reduced(__builtin_elementwise_maximum(x, -104.0f));
#else
   return kinc_expf(x);
#endif
}

/* The arithmetic of kinc_logf for positive finite x; other x give
 * unspecified values (NaN stays NaN). */
KINC_HD float kinc_logf_core(float x)
{
   const float SQRTHF = 0.707106781186547524f;
   const float C1 = 0.693359375f;
   const float C2 = -2.12194440e-4f;

   /* m in [0.5, 1) for positive finite x; kinc_logf replaces the values
    * computed for zero, negative, infinite and NaN x */
   int e = 0;
   float m = kinc_frexpf(x, &e);

   int lo = m < SQRTHF;
   e = lo ? e - 1 : e;
   float f = lo ? (m + m) - 1.0f : m - 1.0f;

   float z = f * f;
   float y = 7.0376836292E-2f;
   y = kinc_fmaf(y, f, -1.1514610310E-1f);
   y = kinc_fmaf(y, f, 1.1676998740E-1f);
   y = kinc_fmaf(y, f, -1.2420140846E-1f);
   y = kinc_fmaf(y, f, 1.4249322787E-1f);
   y = kinc_fmaf(y, f, -1.6668057665E-1f);
   y = kinc_fmaf(y, f, 2.0000714765E-1f);
   y = kinc_fmaf(y, f, -2.4999993993E-1f);
   y = kinc_fmaf(y, f, 3.3333331174E-1f);
   y = y * f;
   y = y * z;

   float fe = (float)e;
   y = kinc_fmaf(fe, C2, y);
   y = kinc_fmaf(-0.5f, z, y);
   float r = f + y;
   return kinc_fmaf(fe, C1, r);
}

KINC_HD float kinc_logf(float x)
{
   const float r = kinc_logf_core(x);
   /* log(+inf) = +inf, log(+-0) = -inf, log(x < 0) = log(NaN) = NaN */
   const int regular = (x > 0.0f) && (x < INFINITY);
   const float special = (x == 0.0f) ? -INFINITY : ((x > 0.0f) ? x : NAN);
   return regular ? r : special;
}

/* kinc_logf for finite x >= 1 or NaN, the sums of a log-sum-exp whose
 * largest term is exp(0) = 1: the special-value cases never apply, so they
 * are left out.  The same bits as kinc_logf for finite x >= 1, NaN for NaN. */
KINC_HD float kinc_logf_ge1(float x)
{
   return kinc_logf_core(x);
}

#endif
