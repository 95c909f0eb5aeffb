/*
 * kinc_oracle.c — CPU oracle for the pairwise Similarity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the reference's
 * serial Similarity path (KINC v3.2.0, src/core/similarity_serial.cpp and
 * src/core/pairwise_*.cpp).  It is used by the tests as the checker of the
 * HIP path, and by bench.py as the timed CPU baseline.  Nothing in the
 * product (kinc_amd/) links, loads or calls it.
 *
 * Numeric environment (the reference's x86-64 SSE2 build): FP32 storage and
 * arithmetic, no FMA contraction (built with -ffp-contract=off), correctly
 * rounded division and sqrt, denormals preserved, sums in index order.
 * exp/log come from the shared kinc_math.h so that the oracle and the HIP
 * path use one implementation.  The cluster-initialisation RNG is the
 * per-fit POSIX LCG of src/opencl/outlier.cl:14-18 reset to 1 at every fit
 * (src/opencl/gmm.cl:432, src/opencl/kmeans.cl:62), the only per-pair
 * deterministic seed the reference has.
 *
 * Every reference quirk on this path is reproduced on purpose:
 *   - the compact X[0..N) array is paired with the full-length labels[0..S)
 *     array in outlier marking, label copy-back, Pearson and Spearman
 *     (pairwise_clustering.cpp:77-84,116-147; pairwise_pearson.cpp:26-31;
 *     pairwise_spearman.cpp:37-51);
 *   - the X workspace is allocated once per block and only its prefix is
 *     overwritten per pair, so Pearson reads stale entries of earlier pairs
 *     (similarity_serial.cpp:42,132);
 *   - the GMM k-means initialisation is dead code (pairwise_gmm.cpp:90-95);
 *   - Spearman accumulates in int32 with wrap-around
 *     (pairwise_spearman.cpp:73-82).
 *
 * Parity pins (tests/test_oracle.py): closed forms for the pair index and
 * BIC, LCG known answers, brute force for quartiles and for untied Spearman,
 * a textbook FP64 Pearson, K=1 GMM closed form, separated-cluster recovery,
 * Lloyd fixed points, and libm agreement for exp/log.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/kinc_sim.h"
#include "../kinc_amd/csrc/kinc_math.h"

typedef struct { float s[2]; } vec2;
typedef struct { float s[4]; } mat2;

/* 2*log(2*pi) in double, the D*log(2.0f*M_PI) term of
 * src/core/pairwise_gmm.cpp:46 with D = 2. */
static const double TWO_LOG_2PI = 3.6757541328186907;

/* ---------------------------------------------------------------- index */

/* src/core/pairwise_index.cpp:30-56 as a closed form. */
void oracle_pair_index(int64_t index, int32_t* x, int32_t* y)
{
   int64_t xx = (int64_t)((1.0 + sqrt(1.0 + 8.0 * (double)index)) / 2.0);
   while ( xx * (xx - 1) / 2 > index ) --xx;
   while ( (xx + 1) * xx / 2 <= index ) ++xx;
   *x = (int32_t)xx;
   *y = (int32_t)(index - xx * (xx - 1) / 2);
}

/* src/core/pairwise_index.cpp:30-56 exactly as written (linear search). */
void oracle_pair_index_loop(int64_t index, int32_t* x, int32_t* y)
{
   int64_t px = 1;
   int64_t pos = 0;
   while ( pos <= index )
   {
      ++px;
      pos = px * (px - 1) / 2;
   }
   --px;
   pos = px * (px - 1) / 2;
   *x = (int32_t)px;
   *y = (int32_t)(index - pos);
}

/* ------------------------------------------------------------------ rng */

/* src/opencl/outlier.cl:14-18 */
int oracle_lcg(uint64_t* state)
{
   *state = (*state) * 1103515245ull + 12345ull;
   return (int)((uint32_t)((*state) / 65536ull) % 32768u);
}

/* ---------------------------------------------------------------- fetch */

/* src/core/similarity_serial.cpp:108-141 */
static int fetch_pair(const float* g1, const float* g2, int S, float min_expr,
                      vec2* X, int8_t* labels)
{
   int n = 0;
   for ( int i = 0; i < S; ++i )
   {
      if ( isnan(g1[i]) || isnan(g2[i]) )
      {
         labels[i] = -9;
      }
      else if ( g1[i] < min_expr || g2[i] < min_expr )
      {
         labels[i] = -6;
      }
      else
      {
         X[n].s[0] = g1[i];
         X[n].s[1] = g2[i];
         n++;
         labels[i] = 0;
      }
   }
   return n;
}

/* ------------------------------------------------------------- outliers */

static int cmp_float(const void* a, const void* b)
{
   float fa = *(const float*)a, fb = *(const float*)b;
   return (fa < fb) ? -1 : (fb < fa) ? 1 : 0;
}

/* src/core/pairwise_clustering.cpp:110-148.  `work` holds >= N floats. */
void oracle_mark_outliers(const vec2* X, int N, int j, int8_t* labels,
                          int8_t cluster, int8_t marker, float* work)
{
   int n = 0;
   for ( int i = 0; i < N; i++ )
   {
      if ( labels[i] == cluster || labels[i] == marker )
      {
         work[n++] = X[i].s[j];
      }
   }
   if ( n == 0 )
   {
      return;
   }
   qsort(work, (size_t)n, sizeof(float), cmp_float);

   float Q1 = work[n * 1 / 4];
   float Q3 = work[n * 3 / 4];
   float T_min = Q1 - 1.5f * (Q3 - Q1);
   float T_max = Q3 + 1.5f * (Q3 - Q1);

   for ( int i = 0; i < N; ++i )
   {
      if ( labels[i] == cluster && (X[i].s[j] < T_min || T_max < X[i].s[j]) )
      {
         labels[i] = marker;
      }
   }
}

/* ------------------------------------------------------------------ GMM */

typedef struct
{
   float pi;
   vec2 mu;
   mat2 sigma;
   mat2 sigmaInv;
   float normalizer;
} gmm_component;

/* src/core/pairwise_linalg.cpp:163-173 + src/core/pairwise_gmm.cpp:31-47.
 * Returns 0 when the reference would throw (det <= 0). */
static int prepare_covariance(gmm_component* c)
{
   const float* A = c->sigma.s;
   float det = A[0] * A[3] - A[1] * A[2];
   c->sigmaInv.s[0] = +A[3] / det;
   c->sigmaInv.s[1] = -A[1] / det;
   c->sigmaInv.s[2] = -A[2] / det;
   c->sigmaInv.s[3] = +A[0] / det;
   if ( det <= 0 )
   {
      return 0;
   }
   c->normalizer = (float)(-0.5 * (TWO_LOG_2PI + (double)kinc_logf(det)));
   return 1;
}

/* src/core/pairwise_gmm.cpp:54-77 */
static float log_mv_norm(const gmm_component* c, vec2 x)
{
   float xm0 = x.s[0] - c->mu.s[0];
   float xm1 = x.s[1] - c->mu.s[1];
   const float* A = c->sigmaInv.s;
   float b0 = A[0] * xm0 + A[1] * xm1;
   float b1 = A[2] * xm0 + A[3] * xm1;
   float dot = xm0 * b0 + xm1 * b1;
   return c->normalizer - 0.5f * dot;
}

/* Statistics of one fit, for flop accounting: E-steps run, M-steps run
 * (a failing M-step counts), converged and failed flags. */
typedef struct
{
   int e_steps;
   int m_steps;
   int converged;
   int failed;
} fit_stats;

static uint16_t pack_iters(const fit_stats* st)
{
   uint16_t v = (uint16_t)(st->e_steps & 0x0fff);
   if ( st->converged ) v |= (uint16_t)(1u << 12);
   if ( st->failed ) v |= (uint16_t)(1u << 13);
   return v;
}

/* Test hook: which deliberate slip the entropy takes (0: none).  Only the
 * tests that check the entropy pin set it. */
static int g_plant = 0;

void oracle_set_plant(int which)
{
   g_plant = which;
}

/* src/core/pairwise_gmm.cpp:367-445 with the per-fit LCG.
 * labels: N entries written on success.  Returns 1 on success. */
int oracle_gmm_fit(const vec2* X, int N, int K, int8_t* labels,
                   float* logL_out, float* entropy_out, fit_stats* st)
{
   gmm_component comp[KINC_MAX_CLUSTER_SIZE];
   float logpi[KINC_MAX_CLUSTER_SIZE];
   float logGamma[KINC_MAX_CLUSTER_SIZE];
   uint64_t state = 1;
   fit_stats local = {0, 0, 0, 0};

   for ( int k = 0; k < K; ++k )
   {
      int i = oracle_lcg(&state) % N;
      comp[k].pi = 1.0f / K;
      comp[k].mu = X[i];
      comp[k].sigma.s[0] = 1; comp[k].sigma.s[1] = 0;
      comp[k].sigma.s[2] = 0; comp[k].sigma.s[3] = 1;
      memset(&comp[k].sigmaInv, 0, sizeof(mat2));
      comp[k].normalizer = 0;
      prepare_covariance(&comp[k]);
   }
   /* the k-means refinement (pairwise_gmm.cpp:84-136) never runs: its loop
    * condition tests diff > TOLERANCE with diff = 0 */

   float* loggamma = (float*)malloc(sizeof(float) * (size_t)K * (size_t)N);
   for ( int k = 0; k < K; ++k )
   {
      logpi[k] = kinc_logf(comp[k].pi);
   }

   const int MAX_ITERATIONS = 100;
   const float TOLERANCE = 1e-8f;
   float prevLogL = -INFINITY;
   float currentLogL = -INFINITY;
   int ok = 1;

   for ( int t = 0; t < MAX_ITERATIONS; ++t )
   {
      /* E step: calcLogMvNorm (pairwise_gmm.cpp:143-151) */
      for ( int k = 0; k < K; ++k )
      {
         for ( int i = 0; i < N; ++i )
         {
            loggamma[k * N + i] = log_mv_norm(&comp[k], X[i]);
         }
      }
      local.e_steps++;

      /* calcLogLikelihoodAndGammaNK (pairwise_gmm.cpp:158-187) */
      prevLogL = currentLogL;
      float logL = 0.0f;
      for ( int i = 0; i < N; ++i )
      {
         float maxArg = -INFINITY;
         for ( int k = 0; k < K; ++k )
         {
            const float logProbK = logpi[k] + loggamma[k * N + i];
            if ( logProbK > maxArg ) maxArg = logProbK;
         }
         float sum = 0.0f;
         for ( int k = 0; k < K; ++k )
         {
            const float logProbK = logpi[k] + loggamma[k * N + i];
            sum += kinc_expf(logProbK - maxArg);
         }
         const float logpx = maxArg + kinc_logf(sum);
         logL += logpx;
         for ( int k = 0; k < K; ++k )
         {
            loggamma[k * N + i] += -logpx;
         }
      }
      currentLogL = logL;

      if ( fabsf(currentLogL - prevLogL) < TOLERANCE )
      {
         local.converged = 1;
         break;
      }

      /* calcLogGammaK (pairwise_gmm.cpp:194-221) */
      for ( int k = 0; k < K; ++k )
      {
         const float* lg = &loggamma[k * N];
         float maxArg = -INFINITY;
         for ( int i = 0; i < N; ++i )
         {
            if ( lg[i] > maxArg ) maxArg = lg[i];
         }
         float sum = 0;
         for ( int i = 0; i < N; ++i )
         {
            sum += kinc_expf(lg[i] - maxArg);
         }
         logGamma[k] = maxArg + kinc_logf(sum);
      }

      /* calcLogGammaSum (pairwise_gmm.cpp:228-248) */
      float maxArg = -INFINITY;
      for ( int k = 0; k < K; ++k )
      {
         const float arg = logpi[k] + logGamma[k];
         if ( arg > maxArg ) maxArg = arg;
      }
      float sum = 0;
      for ( int k = 0; k < K; ++k )
      {
         const float arg = logpi[k] + logGamma[k];
         sum += kinc_expf(arg - maxArg);
      }
      const float logGammaSum = maxArg + kinc_logf(sum);

      /* performMStep (pairwise_gmm.cpp:255-316) */
      local.m_steps++;
      for ( int k = 0; k < K; ++k )
      {
         logpi[k] += logGamma[k] - logGammaSum;
         comp[k].pi = kinc_expf(logpi[k]);
      }
      for ( int k = 0; k < K; ++k )
      {
         for ( int i = 0; i < N; ++i )
         {
            loggamma[k * N + i] = kinc_expf(loggamma[k * N + i]);
         }
      }
      for ( int k = 0; k < K; ++k )
      {
         logGamma[k] = kinc_expf(logGamma[k]);
      }
      for ( int k = 0; k < K; ++k )
      {
         vec2* mu = &comp[k].mu;
         mu->s[0] = 0;
         mu->s[1] = 0;
         for ( int i = 0; i < N; ++i )
         {
            const float g = loggamma[k * N + i];
            mu->s[0] += g * X[i].s[0];
            mu->s[1] += g * X[i].s[1];
         }
         const float cm = 1.0f / logGamma[k];
         mu->s[0] *= cm;
         mu->s[1] *= cm;

         mat2* sigma = &comp[k].sigma;
         memset(sigma, 0, sizeof(mat2));
         for ( int i = 0; i < N; ++i )
         {
            const float xm0 = X[i].s[0] - mu->s[0];
            const float xm1 = X[i].s[1] - mu->s[1];
            const float g = loggamma[k * N + i];
            sigma->s[0] += g * (xm0 * xm0);
            sigma->s[1] += g * (xm0 * xm1);
            sigma->s[2] += g * (xm1 * xm0);
            sigma->s[3] += g * (xm1 * xm1);
         }
         const float cs = 1.0f / logGamma[k];
         sigma->s[0] *= cs;
         sigma->s[1] *= cs;
         sigma->s[2] *= cs;
         sigma->s[3] *= cs;

         if ( !prepare_covariance(&comp[k]) )
         {
            ok = 0;
            break;
         }
      }
      if ( !ok )
      {
         local.failed = 1;
         break;
      }
   }

   if ( ok )
   {
      *logL_out = currentLogL;

      /* calcLabels (pairwise_gmm.cpp:323-341) */
      for ( int i = 0; i < N; ++i )
      {
         int max_k = -1;
         float max_gamma = -INFINITY;
         for ( int k = 0; k < K; ++k )
         {
            if ( max_gamma < loggamma[k * N + i] )
            {
               max_k = k;
               max_gamma = loggamma[k * N + i];
            }
         }
         labels[i] = (int8_t)max_k;
      }

      /* calcEntropy (pairwise_gmm.cpp:348-360).  A label of -1 makes the
       * reference read before the array (undefined); it is taken as NaN.
       * g_plant switches in a deliberate slip for the tests that show the
       * entropy pin notices one; it is 0 everywhere else. */
      float E = 0;
      for ( int i = 0; i < N; ++i )
      {
         int k = labels[i];
         float v = (k >= 0) ? loggamma[k * N + i] : NAN;
         if ( g_plant == 1 && local.converged ) v = kinc_expf(v);   /* on gamma */
         if ( g_plant == 2 && k >= 0 ) v = v + logpi[k];            /* with log pi */
         E += (k >= 0) ? kinc_logf(v) : NAN;
      }
      *entropy_out = E;
   }

   free(loggamma);
   if ( st ) *st = local;
   return ok;
}

/* --------------------------------------------------------------- KMeans */

static float vector_diff_norm(vec2 a, vec2 b)
{
   /* src/core/pairwise_linalg.cpp:97-104 */
   float dist = 0;
   dist += (a.s[0] - b.s[0]) * (a.s[0] - b.s[0]);
   dist += (a.s[1] - b.s[1]) * (a.s[1] - b.s[1]);
   return sqrtf(dist);
}

/* src/core/pairwise_kmeans.cpp:12-127 with the per-fit LCG.
 * labels: receives the best init's labels (N entries) when an init improves
 * on -inf; left untouched otherwise, like the std::swap of the reference. */
int oracle_kmeans_fit(const vec2* X, int N, int K, int8_t* labels,
                      float* logL_out, fit_stats* st)
{
   const int NUM_INITS = 10;
   const int MAX_ITERATIONS = 300;
   vec2 means[KINC_MAX_CLUSTER_SIZE];
   uint64_t state = 1;
   float bestLogL = -INFINITY;
   fit_stats local = {0, 0, 0, 0};

   int8_t* y = (int8_t*)malloc((size_t)N);
   int8_t* y_next = (int8_t*)malloc((size_t)N);

   for ( int init = 0; init < NUM_INITS; ++init )
   {
      for ( int k = 0; k < K; ++k )
      {
         int i = oracle_lcg(&state) % N;
         means[k] = X[i];
      }
      memset(y, 0, (size_t)N);
      memset(y_next, 0, (size_t)N);

      for ( int t = 0; t < MAX_ITERATIONS; ++t )
      {
         local.e_steps++;
         for ( int i = 0; i < N; ++i )
         {
            int min_k = -1;
            float min_dist = 0;
            for ( int k = 0; k < K; ++k )
            {
               float dist = vector_diff_norm(X[i], means[k]);
               if ( min_k == -1 || dist < min_dist )
               {
                  min_k = k;
                  min_dist = dist;
               }
            }
            y_next[i] = (int8_t)min_k;
         }

         if ( memcmp(y, y_next, (size_t)N) == 0 )
         {
            break;
         }

         int8_t* tmp = y; y = y_next; y_next = tmp;

         for ( int k = 0; k < K; ++k )
         {
            int n_k = 0;
            means[k].s[0] = 0;
            means[k].s[1] = 0;
            for ( int i = 0; i < N; ++i )
            {
               if ( y[i] == k )
               {
                  means[k].s[0] += X[i].s[0];
                  means[k].s[1] += X[i].s[1];
                  n_k++;
               }
            }
            const float c = 1.0f / n_k;
            means[k].s[0] *= c;
            means[k].s[1] *= c;
         }
      }

      /* computeLogLikelihood (pairwise_kmeans.cpp:106-127) */
      float Ssum = 0;
      for ( int k = 0; k < K; ++k )
      {
         for ( int i = 0; i < N; ++i )
         {
            if ( y[i] != k ) continue;
            float dist = vector_diff_norm(X[i], means[k]);
            Ssum += dist * dist;
         }
      }
      float logL = -Ssum;

      if ( bestLogL < logL )
      {
         bestLogL = logL;
         memcpy(labels, y, (size_t)N);
      }
   }

   free(y);
   free(y_next);
   *logL_out = bestLogL;
   if ( st ) *st = local;
   return 1;
}

/* ------------------------------------------------------- model selection */

/* src/core/pairwise_clustering.cpp:155-160: log(int) is the double log. */
float oracle_bic(int K, float logL, int N, int D)
{
   int p = K * (1 + D + D * D);
   return (float)(log((double)N) * p - (double)(2 * logL));
}

/* src/core/pairwise_clustering.cpp:167-172 */
float oracle_icl(int K, float logL, int N, int D, float E)
{
   int p = K * (1 + D + D * D);
   return (float)(log((double)N) * p - (double)(2 * logL) - (double)(2 * E));
}

/* src/core/pairwise_clustering.cpp:23-103.  wlabels is the clustering
 * object's persistent work-label array (>= S entries). */
static int clustering_compute(const kinc_sim_params* p, const vec2* X, int N,
                              int8_t* labels, int8_t* wlabels, float* work,
                              uint16_t* iters)
{
   if ( p->pre_out )
   {
      oracle_mark_outliers(X, N, 0, labels, 0, -7, work);
      oracle_mark_outliers(X, N, 1, labels, 0, -7, work);
   }

   int bestK = 0;
   if ( N >= p->min_samples )
   {
      float bestValue = INFINITY;
      for ( int K = p->min_clusters; K <= p->max_clusters; ++K )
      {
         float logL = 0, E = 0;
         fit_stats st = {0, 0, 0, 0};
         int success;
         if ( p->clus_method == KINC_CLUS_GMM )
         {
            success = oracle_gmm_fit(X, N, K, wlabels, &logL, &E, &st);
         }
         else
         {
            success = oracle_kmeans_fit(X, N, K, wlabels, &logL, &st);
            E = 0;
         }
         if ( iters ) iters[K - 1] = pack_iters(&st);
         if ( !success ) continue;

         float value = (p->criterion == KINC_CRIT_BIC)
            ? oracle_bic(K, logL, N, 2)
            : oracle_icl(K, logL, N, 2, E);

         if ( value < bestValue )
         {
            bestK = K;
            bestValue = value;
            for ( int i = 0, j = 0; i < N; ++i )
            {
               if ( labels[i] >= 0 )
               {
                  labels[i] = wlabels[j];
                  ++j;
               }
            }
         }
      }
   }

   if ( bestK > 1 && p->post_out )
   {
      for ( int k = 0; k < bestK; ++k )
      {
         oracle_mark_outliers(X, N, 0, labels, (int8_t)k, -8, work);
         oracle_mark_outliers(X, N, 1, labels, (int8_t)k, -8, work);
      }
   }
   return bestK;
}

/* -------------------------------------------------------------- Pearson */

/* src/core/pairwise_pearson.cpp:12-52: full-length loop over labels with
 * data[i] read at the full index. */
/* kinc_sim_params.pearson_fp64 = 1 (an option the reference lacks): the
 * same sums in the same order, in double over exact products, and the
 * quotient in double, rounded to float at the end. */
static float pearson_fp64(const vec2* data, const int8_t* labels, int S,
                          int8_t cluster, int minSamples)
{
   int n = 0;
   double sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0;
   for ( int i = 0; i < S; ++i )
   {
      if ( labels[i] != cluster ) continue;
      const double x_i = data[i].s[0];
      const double y_i = data[i].s[1];
      sx += x_i;
      sy += y_i;
      sxx += x_i * x_i;
      syy += y_i * y_i;
      sxy += x_i * y_i;
      ++n;
   }
   if ( n < minSamples ) return NAN;
   return (float)((n * sxy - sx * sy) / sqrt((n * sxx - sx * sx) * (n * syy - sy * sy)));
}

float oracle_pearson(const vec2* data, const int8_t* labels, int S,
                     int8_t cluster, int minSamples)
{
   int n = 0;
   float sumx = 0, sumy = 0, sumx2 = 0, sumy2 = 0, sumxy = 0;
   for ( int i = 0; i < S; ++i )
   {
      if ( labels[i] == cluster )
      {
         float x_i = data[i].s[0];
         float y_i = data[i].s[1];
         sumx += x_i;
         sumy += y_i;
         sumx2 += x_i * x_i;
         sumy2 += y_i * y_i;
         sumxy += x_i * y_i;
         ++n;
      }
   }
   float result = NAN;
   if ( n >= minSamples )
   {
      float fn = (float)n;
      result = (fn * sumxy - sumx * sumy)
             / sqrtf((fn * sumx2 - sumx * sumx) * (fn * sumy2 - sumy * sumy));
   }
   return result;
}

/* ------------------------------------------------------------- Spearman */

static int next_power2(int n)
{
   int pow2 = 2;
   while ( pow2 < n ) pow2 *= 2;
   return pow2;
}

/* src/core/pairwise_spearman.cpp:109-166: the bitonic network, swapping the
 * payload with the key. */
void oracle_bitonic_sort(int size, float* key, float* payload)
{
   int bsize = size / 2;
   for ( int ob = 2; ob <= size; ob *= 2 )
   {
      for ( int ib = ob; ib >= 2; ib /= 2 )
      {
         int t = ib / 2;
         for ( int i = 0; i < bsize; ++i )
         {
            int dir = -((i / (ob / 2)) & 0x1);
            int a = (i / t) * ib + (i % t);
            int b = a + t;
            if ( ((key[a] > key[b]) && !dir) || ((key[a] < key[b]) && dir) )
            {
               float tk = key[a]; key[a] = key[b]; key[b] = tk;
               float tp = payload[a]; payload[a] = payload[b]; payload[b] = tp;
            }
         }
      }
   }
}

/* src/core/pairwise_spearman.cpp:27-86.  wx, wy, wr hold nextPow2(S) floats. */
float oracle_spearman(const vec2* data, const int8_t* labels, int S,
                      int8_t cluster, int minSamples, float* wx, float* wy, float* wr)
{
   int N_pow2 = next_power2(S);
   int n = 0;
   for ( int i = 0, j = 0; i < S; ++i )
   {
      if ( labels[i] >= 0 )
      {
         if ( labels[i] == cluster )
         {
            wx[n] = data[j].s[0];
            wy[n] = data[j].s[1];
            wr[n] = (float)(n + 1);
            ++n;
         }
         ++j;
      }
   }
   for ( int i = n; i < N_pow2; ++i )
   {
      wx[i] = INFINITY;
      wy[i] = INFINITY;
      wr[i] = 0;
   }

   float result = NAN;
   if ( n >= minSamples )
   {
      int n_pow2 = next_power2(n);
      oracle_bitonic_sort(n_pow2, wx, wy);
      oracle_bitonic_sort(n_pow2, wy, wr);

      /* int32 arithmetic with two's-complement wrap, as the reference's
       * x86-64 build computes it */
      uint32_t diff = 0;
      for ( int i = 0; i < n; ++i )
      {
         int tmp = (int)((float)(i + 1) - wr[i]);
         diff += (uint32_t)(tmp * tmp);
      }
      uint32_t un = (uint32_t)n;
      int32_t denom = (int32_t)(un * (un * un - 1u));
      result = (float)(1.0 - 6.0 * (double)(int32_t)diff / (double)denom);
   }
   return result;
}

/* ---------------------------------------------------------------- block */

/*
 * Similarity::Serial::execute (src/core/similarity_serial.cpp:33-101) for one
 * work block.  Output layout as kinc_sim_execute.
 */
int oracle_execute_block(const kinc_sim_params* p, const float* emx,
                         int32_t G, int32_t S, int64_t start, int64_t count,
                         int8_t* K_out, int8_t* labels_out, float* corr_out,
                         uint16_t* iters_out)
{
   if ( p->max_clusters < p->min_clusters || p->max_clusters > KINC_MAX_CLUSTER_SIZE
        || p->min_clusters < 1 )
   {
      return KINC_E_INVALID;
   }
   const int MC = p->max_clusters;
   const int NP2 = next_power2(S);

   vec2* X = (vec2*)calloc((size_t)S, sizeof(vec2));  /* per block, zeroed */
   int8_t* labels = (int8_t*)malloc((size_t)S);
   int8_t* wlabels = (int8_t*)calloc((size_t)S, 1);
   float* work = (float*)malloc(sizeof(float) * (size_t)S);
   float* wx = (float*)malloc(sizeof(float) * (size_t)NP2);
   float* wy = (float*)malloc(sizeof(float) * (size_t)NP2);
   float* wr = (float*)malloc(sizeof(float) * (size_t)NP2);

   int32_t x, y;
   oracle_pair_index(start, &x, &y);

   for ( int64_t q = 0; q < count; ++q )
   {
      const float* g1 = emx + (size_t)x * (size_t)S;
      const float* g2 = emx + (size_t)y * (size_t)S;
      int N = fetch_pair(g1, g2, S, p->min_expr, X, labels);

      uint16_t* it = iters_out ? iters_out + q * MC : NULL;
      if ( it ) memset(it, 0, sizeof(uint16_t) * (size_t)MC);

      int K = 1;
      if ( p->clus_method != KINC_CLUS_NONE )
      {
         K = clustering_compute(p, X, N, labels, wlabels, work, it);
      }

      float* corr = corr_out + q * MC;
      for ( int k = 0; k < MC; ++k ) corr[k] = NAN;
      for ( int k = 0; k < K; ++k )
      {
         corr[k] = (p->corr_method == KINC_CORR_PEARSON)
            ? (p->pearson_fp64 ? pearson_fp64(X, labels, S, (int8_t)k, p->min_samples)
                               : oracle_pearson(X, labels, S, (int8_t)k, p->min_samples))
            : oracle_spearman(X, labels, S, (int8_t)k, p->min_samples, wx, wy, wr);
      }

      K_out[q] = (int8_t)K;
      if ( labels_out ) memcpy(labels_out + q * S, labels, (size_t)S);

      /* Pairwise::Index::operator++ (src/core/pairwise_index.cpp:85-94) */
      if ( ++y >= x )
      {
         y = 0;
         ++x;
      }
   }
   (void)G;

   free(X); free(labels); free(wlabels); free(work);
   free(wx); free(wy); free(wr);
   return KINC_OK;
}

/*
 * An explicit pair list: pairs[2q], pairs[2q+1] = (x, y).  seg_lo[q] (NULL =
 * all 0) is the first pair of q's work block: every block starts with a
 * fresh, zeroed X workspace, exactly as separate Serial::execute calls do.
 */
int oracle_execute_pairs_range(const kinc_sim_params* p, const float* emx, int32_t G, int32_t S,
                               const int32_t* pairs, const int32_t* seg_lo, int64_t count,
                               int64_t first, int64_t n, int8_t* K_out, int8_t* labels_out,
                               float* corr_out, uint16_t* iters_out);

int oracle_execute_pairs(const kinc_sim_params* p, const float* emx, int32_t G, int32_t S,
                         const int32_t* pairs, const int32_t* seg_lo, int64_t count,
                         int8_t* K_out, int8_t* labels_out, float* corr_out, uint16_t* iters_out)
{
   return oracle_execute_pairs_range(p, emx, G, S, pairs, seg_lo, count, 0, count,
                                     K_out, labels_out, corr_out, iters_out);
}

/*
 * Pairs [first, first + n) of the call above, with outputs for those pairs
 * only.  The X workspace a pair reads depends on the fetches of the earlier
 * pairs of its work block alone (fetchPair writes it; clustering and
 * correlation only read it), so the pairs before `first` are fetched and not
 * clustered.  Disjoint ranges of one call can so run on separate threads and
 * give the results of the whole call.
 */
int oracle_execute_pairs_range(const kinc_sim_params* p, const float* emx, int32_t G, int32_t S,
                               const int32_t* pairs, const int32_t* seg_lo, int64_t count,
                               int64_t first, int64_t n, int8_t* K_out, int8_t* labels_out,
                               float* corr_out, uint16_t* iters_out)
{
   if ( first < 0 || n < 0 || first + n > count ) return KINC_E_INVALID;
   const int MC = p->max_clusters;
   const int NP2 = next_power2(S);
   vec2* X = (vec2*)calloc((size_t)S, sizeof(vec2));
   int8_t* labels = (int8_t*)malloc((size_t)S);
   int8_t* wlabels = (int8_t*)calloc((size_t)S, 1);
   float* work = (float*)malloc(sizeof(float) * (size_t)S);
   float* wx = (float*)malloc(sizeof(float) * (size_t)NP2);
   float* wy = (float*)malloc(sizeof(float) * (size_t)NP2);
   float* wr = (float*)malloc(sizeof(float) * (size_t)NP2);
   (void)G;
   /* the latest work-block start at or before `first` */
   int64_t q0 = first;
   if ( seg_lo && first < count ) q0 = seg_lo[first];
   else if ( !seg_lo ) q0 = 0;
   for ( int64_t q = q0; q < first + n; ++q )
   {
      const int32_t lo = seg_lo ? seg_lo[q] : 0;
      if ( q == lo || q == q0 ) memset(X, 0, sizeof(vec2) * (size_t)S);   /* a new work block */
      const float* g1 = emx + (size_t)pairs[2 * q] * (size_t)S;
      const float* g2 = emx + (size_t)pairs[2 * q + 1] * (size_t)S;
      int N = fetch_pair(g1, g2, S, p->min_expr, X, labels);
      if ( q < first ) continue;   /* workspace replay only */
      const int64_t o = q - first;
      uint16_t* it = iters_out ? iters_out + o * MC : NULL;
      if ( it ) memset(it, 0, sizeof(uint16_t) * (size_t)MC);
      int K = 1;
      if ( p->clus_method != KINC_CLUS_NONE )
      {
         K = clustering_compute(p, X, N, labels, wlabels, work, it);
      }
      float* corr = corr_out + o * MC;
      for ( int k = 0; k < MC; ++k ) corr[k] = NAN;
      for ( int k = 0; k < K; ++k )
      {
         corr[k] = (p->corr_method == KINC_CORR_PEARSON)
            ? (p->pearson_fp64 ? pearson_fp64(X, labels, S, (int8_t)k, p->min_samples)
                               : oracle_pearson(X, labels, S, (int8_t)k, p->min_samples))
            : oracle_spearman(X, labels, S, (int8_t)k, p->min_samples, wx, wy, wr);
      }
      K_out[o] = (int8_t)K;
      if ( labels_out ) memcpy(labels_out + o * S, labels, (size_t)S);
   }
   free(X); free(labels); free(wlabels); free(work);
   free(wx); free(wy); free(wr);
   return KINC_OK;
}

/* ------------------------------------------------- exported test helpers */

float oracle_expf(float x) { return kinc_expf(x); }
float oracle_logf(float x) { return kinc_logf(x); }

int oracle_gmm_fit_flat(const float* X, int N, int K, int8_t* labels,
                        float* logL, float* entropy, int32_t* stats4)
{
   fit_stats st;
   int ok = oracle_gmm_fit((const vec2*)X, N, K, labels, logL, entropy, &st);
   if ( stats4 )
   {
      stats4[0] = st.e_steps; stats4[1] = st.m_steps;
      stats4[2] = st.converged; stats4[3] = st.failed;
   }
   return ok;
}

int oracle_kmeans_fit_flat(const float* X, int N, int K, int8_t* labels,
                           float* logL, int32_t* stats4)
{
   fit_stats st;
   int ok = oracle_kmeans_fit((const vec2*)X, N, K, labels, logL, &st);
   if ( stats4 )
   {
      stats4[0] = st.e_steps; stats4[1] = st.m_steps;
      stats4[2] = st.converged; stats4[3] = st.failed;
   }
   return ok;
}

void oracle_mark_outliers_flat(const float* X, int N, int j, int8_t* labels,
                               int cluster, int marker)
{
   float* work = (float*)malloc(sizeof(float) * (size_t)(N > 0 ? N : 1));
   oracle_mark_outliers((const vec2*)X, N, j, labels, (int8_t)cluster, (int8_t)marker, work);
   free(work);
}

float oracle_pearson_flat(const float* data, const int8_t* labels, int S,
                          int cluster, int minSamples)
{
   return oracle_pearson((const vec2*)data, labels, S, (int8_t)cluster, minSamples);
}

float oracle_spearman_flat(const float* data, const int8_t* labels, int S,
                           int cluster, int minSamples)
{
   int np2 = next_power2(S);
   float* wx = (float*)malloc(sizeof(float) * (size_t)np2);
   float* wy = (float*)malloc(sizeof(float) * (size_t)np2);
   float* wr = (float*)malloc(sizeof(float) * (size_t)np2);
   float r = oracle_spearman((const vec2*)data, labels, S, (int8_t)cluster, minSamples, wx, wy, wr);
   free(wx); free(wy); free(wr);
   return r;
}

/* ------------------------------------------------- FP64 error reference */

/*
 * FP64 restatement of the correlation step for error statistics: the same
 * sample selection as the FP32 path (including the full-index reads of the
 * shared X workspace for Pearson), evaluated in double; Spearman with exact
 * int64 arithmetic (no wrap).  Called after oracle_execute_block on the same
 * block with the labels it produced.
 */
int oracle_corr_fp64_block(const kinc_sim_params* p, const float* emx,
                           int32_t G, int32_t S, int64_t start, int64_t count,
                           const int8_t* K_in, const int8_t* labels_in,
                           double* corr_out)
{
   const int MC = p->max_clusters;
   const int NP2 = next_power2(S);
   vec2* X = (vec2*)calloc((size_t)S, sizeof(vec2));
   int8_t* lab = (int8_t*)malloc((size_t)S);
   float* wx = (float*)malloc(sizeof(float) * (size_t)NP2);
   float* wy = (float*)malloc(sizeof(float) * (size_t)NP2);
   float* wr = (float*)malloc(sizeof(float) * (size_t)NP2);
   int32_t x, y;
   oracle_pair_index(start, &x, &y);
   (void)G;

   for ( int64_t q = 0; q < count; ++q )
   {
      const float* g1 = emx + (size_t)x * (size_t)S;
      const float* g2 = emx + (size_t)y * (size_t)S;
      fetch_pair(g1, g2, S, p->min_expr, X, lab);
      const int8_t* labels = labels_in + q * S;
      double* corr = corr_out + q * MC;
      for ( int k = 0; k < MC; ++k ) corr[k] = NAN;

      for ( int k = 0; k < K_in[q]; ++k )
      {
         if ( p->corr_method == KINC_CORR_PEARSON )
         {
            int n = 0;
            double sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0;
            for ( int i = 0; i < S; ++i )
            {
               if ( labels[i] != k ) continue;
               double a = X[i].s[0], b = X[i].s[1];
               sx += a; sy += b; sxx += a * a; syy += b * b; sxy += a * b;
               ++n;
            }
            if ( n >= p->min_samples )
            {
               corr[k] = (n * sxy - sx * sy) / sqrt((n * sxx - sx * sx) * (n * syy - sy * sy));
            }
         }
         else
         {
            int n = 0;
            for ( int i = 0, j = 0; i < S; ++i )
            {
               if ( labels[i] >= 0 )
               {
                  if ( labels[i] == k )
                  {
                     wx[n] = X[j].s[0];
                     wy[n] = X[j].s[1];
                     wr[n] = (float)(n + 1);
                     ++n;
                  }
                  ++j;
               }
            }
            for ( int i = n; i < NP2; ++i ) { wx[i] = INFINITY; wy[i] = INFINITY; wr[i] = 0; }
            if ( n >= p->min_samples )
            {
               int n2 = next_power2(n);
               oracle_bitonic_sort(n2, wx, wy);
               oracle_bitonic_sort(n2, wy, wr);
               int64_t diff = 0;
               for ( int i = 0; i < n; ++i )
               {
                  int64_t tmp = (int64_t)(i + 1) - (int64_t)wr[i];
                  diff += tmp * tmp;
               }
               double dn = (double)n;
               corr[k] = 1.0 - 6.0 * (double)diff / (dn * (dn * dn - 1.0));
            }
         }
      }
      if ( ++y >= x ) { y = 0; ++x; }
   }
   free(X); free(lab); free(wx); free(wy); free(wr);
   return KINC_OK;
}

/* Host evaluation of the shared FP32 functions over an array:
 * fn 0 exp, 1 log, 2 sqrt, 3 reciprocal. */
void oracle_math_array(int fn, const float* in, float* out, int64_t n)
{
   for ( int64_t i = 0; i < n; ++i )
   {
      const float x = in[i];
      switch ( fn )
      {
      case 0: out[i] = kinc_expf(x); break;
      case 1: out[i] = kinc_logf(x); break;
      case 2: out[i] = sqrtf(x); break;
      default: out[i] = 1.0f / x; break;
      }
   }
}
