/*
 * kinc_oracle_rmt.c — CPU oracle for the RMT threshold analytic.
 *
 * TEST INFRASTRUCTURE ONLY, like kinc_oracle.c: the tests use it as the
 * checker of the GPU RMT path and bench.py times it as the CPU baseline;
 * nothing in kinc_amd/ links, loads or calls it.
 *
 * Restates RMT::process (src/core/rmt.cpp:33-109) and its helpers on a
 * dense correlation matrix [N][N][K] as CorrelationMatrix::dumpRawData
 * (src/core/correlationmatrix.cpp:144-177) builds it.  Two steps of the
 * reference live in a third-party library that is absent here, the GNU
 * Scientific Library (KINC links the system GSL 2.x):
 *   - gsl_eigen_symmv + gsl_eigen_symmv_sort(GSL_EIGEN_SORT_ABS_ASC)
 *     (rmt.cpp:243-275): the eigenvalues of a real symmetric matrix in
 *     double, ordered by increasing absolute value.  Any backward-stable
 *     solver gives them to double rounding; this oracle uses the cyclic
 *     Jacobi method (Golub & Van Loan, Matrix Computations, 8.5), which is
 *     independent of the LAPACK-style solver the device uses;
 *   - gsl_interp_akima (rmt.cpp:392-439): Akima's 1970 spline with GSL's
 *     non-periodic end conditions m[-2] = 3m[0] - 2m[1], m[-1] = 2m[0] - m[1]
 *     (and the mirror at the right end), node slopes weighted by
 *     |m[i+1]-m[i]| and |m[i-1]-m[i-2]|, a straight segment where both
 *     weights vanish, and evaluation y_i + dx (b_i + dx (c_i + d_i dx)).
 * Float and double types follow the reference statement by statement.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* RMT::computeMaximums (rmt.cpp:160-189): per (gene, cluster) the largest
 * |r| over the other genes, starting from 0. */
void oracle_rmt_maximums(const float* matrix, int N, int K, float* maximums)
{
   for ( int i = 0; i < N * K; ++i ) maximums[i] = 0;
   for ( int i = 0; i < N; ++i )
   {
      for ( int j = 0; j < i; ++j )
      {
         for ( int k = 0; k < K; ++k )
         {
            float c = fabsf(matrix[(size_t)i * N * K + (size_t)j * K + k]);
            if ( maximums[i * K + k] < c ) maximums[i * K + k] = c;
            if ( maximums[j * K + k] < c ) maximums[j * K + k] = c;
         }
      }
   }
}

/* RMT::computePruneMatrix (rmt.cpp:196-236): the rows/columns whose maximum
 * reaches the threshold, in (gene, cluster) order; strictly lower entries
 * of the same cluster with |r| >= threshold, ones on the diagonal, zeros
 * elsewhere.  Writes the indices and the size x size matrix (row-major);
 * returns the size. */
int oracle_rmt_prune(const float* matrix, const float* maximums, int N, int K, float threshold,
                     int32_t* indices, float* prune)
{
   int n = 0;
   for ( int i = 0; i < N * K; ++i )
   {
      if ( maximums[i] >= threshold ) indices[n++] = i;
   }
   memset(prune, 0, sizeof(float) * (size_t)n * (size_t)n);
   for ( int i = 0; i < n; ++i )
   {
      for ( int j = 0; j < i; ++j )
      {
         if ( indices[i] % K != indices[j] % K ) continue;
         float c = matrix[(size_t)(indices[i] / K) * N * K + (size_t)(indices[j] / K) * K + indices[i] % K];
         if ( fabsf(c) >= threshold ) prune[(size_t)i * n + j] = c;
      }
      prune[(size_t)i * n + i] = 1;
   }
   return n;
}

/* Eigenvalues of the symmetric matrix whose lower triangle is a[n][n]
 * (double, overwritten), by cyclic Jacobi sweeps until the off-diagonal
 * mass is negligible. */
static void jacobi_eigenvalues(double* a, int n, double* w)
{
   for ( int i = 0; i < n; ++i )
      for ( int j = i + 1; j < n; ++j ) a[(size_t)i * n + j] = a[(size_t)j * n + i];
   for ( int sweep = 0; sweep < 100; ++sweep )
   {
      double off = 0, diag = 0;
      for ( int i = 0; i < n; ++i )
      {
         diag += a[(size_t)i * n + i] * a[(size_t)i * n + i];
         for ( int j = i + 1; j < n; ++j ) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
      }
      if ( off <= 1e-30 * (diag + off) || off == 0 ) break;
      for ( int p = 0; p < n; ++p )
      {
         for ( int q = p + 1; q < n; ++q )
         {
            const double apq = a[(size_t)p * n + q];
            if ( apq == 0 ) continue;
            const double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
            /* rotation that annihilates a[p][q] (Golub & Van Loan 8.5.2) */
            const double tau = (aqq - app) / (2 * apq);
            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1 + tau * tau));
            const double c = 1 / sqrt(1 + t * t), s = t * c;
            for ( int k = 0; k < n; ++k )
            {
               const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
               a[(size_t)k * n + p] = c * akp - s * akq;
               a[(size_t)k * n + q] = s * akp + c * akq;
            }
            for ( int k = 0; k < n; ++k )
            {
               const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
               a[(size_t)p * n + k] = c * apk - s * aqk;
               a[(size_t)q * n + k] = s * apk + c * aqk;
            }
            a[(size_t)p * n + q] = 0;
            a[(size_t)q * n + p] = 0;
         }
      }
   }
   for ( int i = 0; i < n; ++i ) w[i] = a[(size_t)i * n + i];
}

/* gsl_eigen_symmv_sort(GSL_EIGEN_SORT_ABS_ASC): selection sort by |value|,
 * the first of equal keys kept in place. */
static void sort_abs_asc(double* w, int n)
{
   for ( int i = 0; i + 1 < n; ++i )
   {
      int k = i;
      for ( int j = i + 1; j < n; ++j )
      {
         if ( fabs(w[j]) < fabs(w[k]) ) k = j;
      }
      if ( k != i )
      {
         double t = w[i];
         w[i] = w[k];
         w[k] = t;
      }
   }
}

/* RMT::computeEigenvalues (rmt.cpp:243-275): float matrix to double,
 * eigenvalues, |.|-ascending order, back to float. */
void oracle_rmt_eigenvalues(const float* prune, int n, float* eigens)
{
   double* a = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
   double* w = (double*)malloc(sizeof(double) * (size_t)n);
   for ( size_t i = 0; i < (size_t)n * (size_t)n; ++i ) a[i] = prune[i];
   jacobi_eigenvalues(a, n, w);
   sort_abs_asc(w, n);
   for ( int i = 0; i < n; ++i ) eigens[i] = (float)w[i];
   free(a);
   free(w);
}

/* ------------------------------------------------------------- Akima */

typedef struct
{
   int n;
   const double* x;
   const double* y;
   double* b;
   double* c;
   double* d;
} akima_t;

static int akima_init(akima_t* s, const double* x, const double* y, int n)
{
   if ( n < 5 ) return -1;
   for ( int i = 0; i + 1 < n; ++i )
   {
      if ( !(x[i] < x[i + 1]) ) return -2;   /* GSL: x values must be strictly increasing */
   }
   double* mm = (double*)malloc(sizeof(double) * (size_t)(n + 3));
   double* m = mm + 2;
   for ( int i = 0; i <= n - 2; ++i ) m[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
   m[-2] = 3.0 * m[0] - 2.0 * m[1];
   m[-1] = 2.0 * m[0] - m[1];
   m[n - 1] = 2.0 * m[n - 2] - m[n - 3];
   m[n] = 3.0 * m[n - 2] - 2.0 * m[n - 3];
   s->n = n;
   s->x = x;
   s->y = y;
   s->b = (double*)malloc(sizeof(double) * (size_t)n);
   s->c = (double*)malloc(sizeof(double) * (size_t)n);
   s->d = (double*)malloc(sizeof(double) * (size_t)n);
   for ( int i = 0; i < n - 1; ++i )
   {
      const double ne = fabs(m[i + 1] - m[i]) + fabs(m[i - 1] - m[i - 2]);
      if ( ne == 0.0 )
      {
         s->b[i] = m[i];
         s->c[i] = 0.0;
         s->d[i] = 0.0;
      }
      else
      {
         const double h = x[i + 1] - x[i];
         const double ne_next = fabs(m[i + 2] - m[i + 1]) + fabs(m[i] - m[i - 1]);
         const double alpha = fabs(m[i - 1] - m[i - 2]) / ne;
         double t_next;
         if ( ne_next == 0.0 )
         {
            t_next = m[i];
         }
         else
         {
            const double alpha_next = fabs(m[i] - m[i - 1]) / ne_next;
            t_next = (1.0 - alpha_next) * m[i] + alpha_next * m[i + 1];
         }
         s->b[i] = (1.0 - alpha) * m[i - 1] + alpha * m[i];
         s->c[i] = (3.0 * m[i] - 2.0 * s->b[i] - t_next) / h;
         s->d[i] = (s->b[i] + t_next - 2.0 * m[i]) / (h * h);
      }
   }
   free(mm);
   return 0;
}

static void akima_free(akima_t* s)
{
   free(s->b);
   free(s->c);
   free(s->d);
}

/* interval: the largest i < n-1 with x[i] <= v (gsl_interp_bsearch) */
static double akima_eval(const akima_t* s, double v)
{
   int lo = 0, hi = s->n - 1;
   while ( hi > lo + 1 )
   {
      int mid = (lo + hi) / 2;
      if ( s->x[mid] > v ) hi = mid;
      else lo = mid;
   }
   const double dx = v - s->x[lo];
   return s->y[lo] + dx * (s->b[lo] + dx * (s->c[lo] + s->d[lo] * dx));
}

/* Akima spline through (x, y) evaluated at xe; 0 or < 0 as akima_init. */
int oracle_akima(const double* x, const double* y, int n, const double* xe, int ne, double* ye)
{
   akima_t s;
   int rc = akima_init(&s, x, y, n);
   if ( rc ) return rc;
   for ( int i = 0; i < ne; ++i ) ye[i] = akima_eval(&s, xe[i]);
   akima_free(&s);
   return 0;
}

/* ------------------------------------------------------------ chi-square */

/* RMT::degenerate (rmt.cpp:366-383): eigenvalues 1..n-1 without the ones
 * within 1e-6 of the last kept value. */
static int degenerate(const float* e, int n, float* u)
{
   const float EPSILON = 1e-6f;
   int m = 0;
   for ( int i = 1; i < n; ++i )
   {
      if ( m == 0 || fabsf(e[i] - u[m - 1]) > EPSILON ) u[m++] = e[i];
   }
   return m;
}

/* RMT::unfold (rmt.cpp:392-439); returns < 0 where GSL would fail. */
static int unfold(const float* e, int n, int pace, float* spacings)
{
   const int ns = n / pace;
   double* x = (double*)malloc(sizeof(double) * (size_t)ns);
   double* y = (double*)malloc(sizeof(double) * (size_t)ns);
   for ( int i = 0; i < ns; ++i )
   {
      x[i] = (double)e[i * pace];
      y[i] = (double)(i * pace + 1) / n;
   }
   x[ns - 1] = e[n - 1];
   y[ns - 1] = 1.0;
   akima_t s;
   int rc = akima_init(&s, x, y, ns);
   if ( rc == 0 )
   {
      float* se = (float*)malloc(sizeof(float) * (size_t)n);
      se[0] = 0.0f;
      se[n - 1] = 1.0f;
      for ( int i = 1; i < n - 1; ++i )
      {
         const double v = e[i];
         if ( v < x[0] || v > x[ns - 1] ) rc = -3;   /* GSL_EDOM */
         se[i] = (float)akima_eval(&s, v);
      }
      for ( int i = 0; i < n - 1; ++i ) spacings[i] = (se[i + 1] - se[i]) * n;
      free(se);
      akima_free(&s);
   }
   free(x);
   free(y);
   return rc;
}

/* RMT::computePaceChiSquare (rmt.cpp:320-359) */
static int pace_chi(const float* e, int n, int pace, int bins, float* chi_out)
{
   float* sp = (float*)malloc(sizeof(float) * (size_t)n);
   int rc = unfold(e, n, pace, sp);
   if ( rc )
   {
      free(sp);
      return rc;
   }
   const float hmin = 0, hmax = 3;
   const float width = (hmax - hmin) / bins;
   float* hist = (float*)calloc((size_t)bins, sizeof(float));
   for ( int i = 0; i < n - 1; ++i )
   {
      if ( hmin <= sp[i] && sp[i] < hmax ) ++hist[(int)((sp[i] - hmin) / width)];
   }
   float chi = 0.0f;
   for ( int i = 0; i < bins; ++i )
   {
      const float O = hist[i];
      const float E = (expf(-i * width) - expf(-(i + 1) * width)) * n;
      chi += (O - E) * (O - E) / E;
   }
   free(sp);
   free(hist);
   *chi_out = chi;
   return 0;
}

/* RMT::computeChiSquare (rmt.cpp:282-313): -1 with fewer than min_eigen
 * unique eigenvalues; -2 where the reference's GSL calls would fail (knots
 * not strictly increasing). */
float oracle_rmt_chi_square(const float* eigens, int n, int min_eigen, int min_pace, int max_pace,
                            int bins)
{
   float* u = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
   const int m = degenerate(eigens, n, u);
   if ( m < min_eigen )
   {
      free(u);
      return -1;
   }
   float chi = 0.0f;
   int count = 0;
   for ( int pace = min_pace; pace <= max_pace; ++pace )
   {
      if ( m / pace < 5 ) break;
      float c;
      if ( pace_chi(u, m, pace, bins, &c) )
      {
         free(u);
         return -2;
      }
      chi += c;
      ++count;
   }
   free(u);
   chi /= count;
   return chi;
}

/*
 * RMT::process (rmt.cpp:33-109): thresholds from start down by step (in
 * float, as the reference accumulates it) until the largest chi-square
 * after the last sub-critical one reaches chi2.  Per evaluated threshold t
 * the log line is (t, size, chi).  Returns the number of thresholds
 * evaluated and the final threshold, or -1 if the stop threshold was
 * passed first (the reference's "RMT Threshold Error"), -2 on a GSL failure
 * (that step is not logged: the reference aborts inside it); *steps_out =
 * the thresholds logged.
 */
int oracle_rmt_run(const float* matrix, int N, int K, float start, float step, float stop,
                   float chi1, float chi2, int min_eigen, int min_pace, int max_pace, int bins,
                   float* log_t, int32_t* log_size, float* log_chi, int cap, float* final_threshold,
                   int* steps_out)
{
   float* maxima = (float*)malloc(sizeof(float) * (size_t)N * K);
   int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (size_t)N * K);
   oracle_rmt_maximums(matrix, N, K, maxima);
   float finalThreshold = 0, finalChi = INFINITY, maxChi = -INFINITY;
   float threshold = start;
   int steps = 0, rc = 0;
   while ( maxChi < chi2 )
   {
      int n = 0;
      for ( int i = 0; i < N * K; ++i ) n += maxima[i] >= threshold;
      float chi = -1;
      if ( n > 0 )
      {
         float* prune = (float*)malloc(sizeof(float) * (size_t)n * (size_t)n);
         float* eig = (float*)malloc(sizeof(float) * (size_t)n);
         oracle_rmt_prune(matrix, maxima, N, K, threshold, idx, prune);
         oracle_rmt_eigenvalues(prune, n, eig);
         chi = oracle_rmt_chi_square(eig, n, min_eigen, min_pace, max_pace, bins);
         free(prune);
         free(eig);
         if ( chi == -2 )
         {
            rc = -2;
            break;
         }
      }
      if ( chi != -1 )
      {
         if ( chi < chi1 )
         {
            finalChi = chi;
            finalThreshold = threshold;
         }
         if ( finalChi < chi1 && chi > finalChi ) maxChi = chi;
      }
      if ( steps < cap )
      {
         log_t[steps] = threshold;
         log_size[steps] = n;
         log_chi[steps] = chi;
      }
      ++steps;
      threshold -= step;
      if ( threshold < stop )
      {
         rc = -1;
         break;
      }
   }
   free(maxima);
   free(idx);
   *final_threshold = finalThreshold;
   *steps_out = steps;
   return rc ? rc : steps;
}
