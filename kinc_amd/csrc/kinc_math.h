/*
 * kinc_math.h — the one FP32 exp/log/sqrt implementation shared by the HIP
 * kernels and the CPU code, so that both produce identical bits.
 *
 * The reference calls unqualified exp/log/sqrt on float arguments
 * (src/core/pairwise_gmm.cpp:46,177,180,216,219,262,271,277,356,391;
 *  src/core/pairwise_linalg.cpp:103; src/core/pairwise_pearson.cpp:48).
 * Those resolve to the FP32 overloads here.  exp and log are written from a
 * published method, Cephes (S. L. Moshier, "Methods and Programs for
 * Mathematical Functions", 1989; Cephes Math Library expf.c / logf.c):
 *   - expf: Cody-Waite reduction x = g + n*ln2 with ln2 split as
 *           C1 = 0.693359375, C2 = -2.12194440e-4, then the degree-5
 *           minimax polynomial of expf.c for e^g on |g| <= ln2/2.
 *   - logf: x = m * 2^e with m in [sqrt(1/2), sqrt(2)), the degree-9
 *           minimax polynomial of logf.c for log(1+f), and the same ln2 split.
 * Only IEEE-754 single-precision +, -, * and comparisons are used (no FMA,
 * no division, no hardware transcendental), so any compiler that keeps
 * FP contraction off and preserves denormals produces the same result on
 * the host CPU and on gfx950.  sqrt is the correctly rounded IEEE sqrtf.
 *
 * Accuracy against glibc (checked by the tests): expf within 2 ulp, logf
 * within 2 ulp over the ranges the pair pipeline uses.
 */
#ifndef KINC_MATH_H
#define KINC_MATH_H

#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__) || defined(__HIP_DEVICE_COMPILE__)
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

/* Round to nearest integer, ties to even, for |x| < 2^22 (exact on both sides:
 * adding and subtracting 1.5*2^23 forces the rounding in the FP adder). */
KINC_HD float kinc_rint_small(float x)
{
   const float magic = 12582912.0f; /* 1.5 * 2^23 */
   float t = x + magic;
   return t - magic;
}

KINC_HD float kinc_expf(float x)
{
   const float MAXLOGF = 88.72283905206835f;
   const float MINLOGF = -103.97207708f; /* below this the result rounds to 0 */
   const float LOG2EF = 1.44269504088896341f;
   const float C1 = 0.693359375f;
   const float C2 = -2.12194440e-4f;

   /* straight-line form: clamp, evaluate, then select the special cases */
   float xc = (x < MINLOGF) ? MINLOGF : x;
   xc = (xc > MAXLOGF) ? MAXLOGF : xc;
   xc = (x != x) ? 0.0f : xc;

   float nf = kinc_rint_small(LOG2EF * xc);
   float g = xc - nf * C1;
   g = g - nf * C2;
   float z = g * g;
   float p = 1.9875691500E-4f;
   p = p * g + 1.3981999507E-3f;
   p = p * g + 8.3334519073E-3f;
   p = p * g + 4.1665795894E-2f;
   p = p * g + 1.6666665459E-1f;
   p = p * g + 5.0000001201E-1f;
   p = p * z + g;
   p = p + 1.0f;

   /* p * 2^n with one rounding: the first factor keeps the product normal
    * and exact, the second applies the remaining power of two. */
   int n = (int)nf;                      /* -150 <= n <= 128 */
   int n1 = (n > 127) ? 127 : ((n < -126) ? n + 126 : n);
   int n2 = (n > 127) ? n - 127 : ((n < -126) ? -126 : 0);
   float r = (p * kinc_pow2i(n1)) * kinc_pow2i(n2);

   r = (x > MAXLOGF) ? INFINITY : r;
   r = (x < MINLOGF) ? 0.0f : r;
   r = (x != x) ? x : r;
   return r;
}

KINC_HD float kinc_logf(float x)
{
   const float SQRTHF = 0.707106781186547524f;
   const float C1 = 0.693359375f;
   const float C2 = -2.12194440e-4f;

   /* denormals are scaled into the normal range first */
   uint32_t u0 = kinc_as_uint(x);
   int den = (u0 & 0x7f800000u) == 0u;
   float xs = den ? x * 33554432.0f : x;   /* 2^25, exact */
   uint32_t u = kinc_as_uint(xs);
   int e = (int)((u >> 23) & 0xffu) - 126 - (den ? 25 : 0);
   float m = kinc_as_float((u & 0x007fffffu) | 0x3f000000u);   /* [0.5, 1) */

   int lo = m < SQRTHF;
   e = lo ? e - 1 : e;
   float f = lo ? (m + m) - 1.0f : m - 1.0f;

   float z = f * f;
   float y = 7.0376836292E-2f;
   y = y * f - 1.1514610310E-1f;
   y = y * f + 1.1676998740E-1f;
   y = y * f - 1.2420140846E-1f;
   y = y * f + 1.4249322787E-1f;
   y = y * f - 1.6668057665E-1f;
   y = y * f + 2.0000714765E-1f;
   y = y * f - 2.4999993993E-1f;
   y = y * f + 3.3333331174E-1f;
   y = y * f;
   y = y * z;

   float fe = (float)e;
   y = y + fe * C2;
   y = y - 0.5f * z;
   float r = f + y;
   r = r + fe * C1;

   r = (x == INFINITY) ? INFINITY : r;
   r = (x == 0.0f) ? -INFINITY : r;
   r = (x < 0.0f) ? NAN : r;
   r = (x != x) ? x : r;
   return r;
}

#endif
