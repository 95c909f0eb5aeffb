/*
 * sim_pair_impl.h — gfx950 device code of the pairwise Similarity path
 * (included by the per-KMAX kernel translation units).
 *
 * One wavefront (64 lanes, one workgroup) processes one gene pair at a time:
 *   fetch (NaN / minimum-expression mask, compaction)        similarity_serial.cpp:108-141
 *   pre-clustering outliers                                  pairwise_clustering.cpp:35-39,110-148
 *   for K = minclus..maxclus: GMM EM or K-means fit, BIC/ICL pairwise_clustering.cpp:23-103
 *   post-clustering outliers                                 pairwise_clustering.cpp:89-100
 *   Pearson or Spearman per cluster                          pairwise_pearson.cpp, pairwise_spearman.cpp
 *
 * Data layout: the expression matrix stays resident in HBM as FP32 [G][S];
 * each pair reads its two gene rows once with coalesced 256-byte wave loads.
 * Point i of the compacted pair (i < N) lives in lane i%64, register slot
 * i/64, so the per-point E- and M-step work is spread over the 64 lanes with
 * X and the K responsibilities held in VGPRs (P slots, 64*P >= N).
 *
 * Exact reduction order.  The reference sums every moment (log-likelihood,
 * Gamma_k, mu_k, Sigma_k, Pearson moments, K-means means) as a sequential FP32
 * chain in index order, and EM convergence (|dlogL| < 1e-8 on an FP32 logL)
 * means bitwise-equal consecutive logL, so any other order changes iteration
 * counts and labels.  The kernel therefore computes all per-point terms in
 * parallel, stages them in an LDS tile [chain][64 points], and lets one lane
 * per chain add its row in index order.  Max reductions, integer counts and
 * the Spearman rank-difference sum are order-independent and use wave
 * reductions.  All arithmetic is IEEE FP32 with contraction off; exp/log come
 * from kinc_math.h; BIC selection and the Spearman quotient use FP64 like
 * the reference's double-typed expressions.
 */
#ifndef KINC_SIM_PAIR_IMPL_H
#define KINC_SIM_PAIR_IMPL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kinc_math.h"
#include "sim_device.h"

#define WAVE 64
#define TILE 64

namespace {

/* ------------------------------------------------------------ wave utils */

__device__ __forceinline__ int lane_id()
{
   return __lane_id();
}

__device__ __forceinline__ uint64_t lanes_below()
{
   return (1ull << lane_id()) - 1ull;
}

__device__ __forceinline__ int popc(uint64_t m)
{
   return __popcll(m);
}

__device__ __forceinline__ uint64_t ballot(bool p)
{
   return __ballot(p);
}

__device__ __forceinline__ float bcast(float v, int src)
{
   return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

__device__ __forceinline__ int bcast_i(int v, int src)
{
   return __builtin_amdgcn_readlane(v, src);
}

/* NaN-skipping max in the reference's `if (v > max) max = v` sense. */
__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
   for ( int off = 32; off >= 1; off >>= 1 )
   {
      float o = __shfl_xor(v, off);
      v = (o > v) ? o : v;
   }
   return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
   for ( int off = 32; off >= 1; off >>= 1 )
   {
      v += (uint32_t)__shfl_xor((int)v, off);
   }
   return v;
}

__device__ __forceinline__ void wave_sync()
{
   __syncthreads();
}

__device__ __forceinline__ int next_pow2(int n)
{
   int p = 2;
   while ( p < n ) p <<= 1;
   return p;
}

/* src/opencl/outlier.cl:14-18 */
__device__ __forceinline__ int lcg(uint64_t& s)
{
   s = s * 1103515245ull + 12345ull;
   return (int)((uint32_t)(s / 65536ull) % 32768u);
}

/* Linear pair index -> (x, y), x > y (src/core/pairwise_index.cpp:30-56). */
__device__ __forceinline__ void pair_xy(int64_t k, int& x, int& y)
{
   double d = sqrt(1.0 + 8.0 * (double)k);
   int64_t xx = (int64_t)((1.0 + d) * 0.5);
   while ( xx * (xx - 1) / 2 > k ) --xx;
   while ( (xx + 1) * xx / 2 <= k ) ++xx;
   x = (int)xx;
   y = (int)(k - xx * (xx - 1) / 2);
}

/* ------------------------------------------------------------------ LDS */

struct Lds
{
   float2* X;      /* [S] */
   int8_t* lab;    /* [S] */
   int8_t* w;      /* [S] */
   float* u;       /* union region */
};

__device__ __forceinline__ Lds lds_view(const SimArgs& a)
{
   extern __shared__ __align__(16) char lds_raw[];
   Lds l;
   l.X = (float2*)(lds_raw + a.lds_x);
   l.lab = (int8_t*)(lds_raw + a.lds_lab);
   l.w = (int8_t*)(lds_raw + a.lds_w);
   l.u = (float*)(lds_raw + a.lds_u);
   return l;
}

/* ---------------------------------------------------------------- fetch */

/* similarity_serial.cpp:108-141: labels -9/-6/0 and compaction in sample order. */
__device__ int fetch_pair(const SimArgs& a, int gx, int gy, const Lds& L)
{
   const int S = a.S;
   const float* r1 = a.emx + (size_t)gx * (size_t)S;
   const float* r2 = a.emx + (size_t)gy * (size_t)S;
   const int lane = lane_id();
   int base = 0;
   for ( int s0 = 0; s0 < S; s0 += WAVE )
   {
      int s = s0 + lane;
      bool in = s < S;
      float g1 = in ? r1[s] : 0.0f;
      float g2 = in ? r2[s] : 0.0f;
      int8_t lab = 0;
      bool valid = false;
      if ( in )
      {
         if ( isnan(g1) || isnan(g2) ) lab = -9;
         else if ( g1 < a.min_expr || g2 < a.min_expr ) lab = -6;
         else valid = true;
         L.lab[s] = lab;
      }
      uint64_t m = ballot(valid);
      if ( valid )
      {
         L.X[base + popc(m & lanes_below())] = make_float2(g1, g2);
      }
      base += popc(m);
   }
   wave_sync();
   return base;
}

/* ---------------------------------------------------------- bitonic sort */

/*
 * The reference's bitonic network (pairwise_spearman.cpp:109-166): for
 * ob = 2..size, ib = ob..2, comparator i pairs a = (i/t)*ib + i%t with
 * b = a + t, direction from (i/(ob/2))&1, swap only when strictly out of
 * order.  Comparators of one stage touch disjoint elements, so running them
 * across lanes gives the same result as the serial loop.
 */
template <bool PAYLOAD>
__device__ void bitonic_sort(float* key, float* pay, int size)
{
   const int lane = lane_id();
   const int half = size >> 1;
   for ( int ob = 2; ob <= size; ob <<= 1 )
   {
      const int obh = ob >> 1;
      for ( int ib = ob; ib >= 2; ib >>= 1 )
      {
         const int t = ib >> 1;
         for ( int i = lane; i < half; i += WAVE )
         {
            int dir = (i / obh) & 1;
            int aa = (i / t) * ib + (i % t);
            int bb = aa + t;
            float ka = key[aa];
            float kb = key[bb];
            bool sw = dir ? (ka < kb) : (ka > kb);
            if ( sw )
            {
               key[aa] = kb;
               key[bb] = ka;
               if ( PAYLOAD )
               {
                  float pa = pay[aa];
                  pay[aa] = pay[bb];
                  pay[bb] = pa;
               }
            }
         }
         wave_sync();
      }
   }
}

/* ------------------------------------------------------------- outliers */

/* pairwise_clustering.cpp:110-148 */
__device__ void mark_outliers(const Lds& L, int N, int j, int8_t cluster, int8_t marker)
{
   const int lane = lane_id();
   float* key = L.u;
   int n = 0;
   for ( int i0 = 0; i0 < N; i0 += WAVE )
   {
      int i = i0 + lane;
      bool c = false;
      float v = 0.0f;
      if ( i < N )
      {
         int8_t l = L.lab[i];
         c = (l == cluster) || (l == marker);
         float2 xv = L.X[i];
         v = j ? xv.y : xv.x;
      }
      uint64_t m = ballot(c);
      if ( c ) key[n + popc(m & lanes_below())] = v;
      n += popc(m);
   }
   if ( n == 0 )
   {
      wave_sync();
      return;
   }
   int n2 = next_pow2(n);
   for ( int i = n + lane; i < n2; i += WAVE ) key[i] = INFINITY;
   wave_sync();
   bitonic_sort<false>(key, nullptr, n2);

   const float Q1 = key[n * 1 / 4];
   const float Q3 = key[n * 3 / 4];
   const float T_min = Q1 - 1.5f * (Q3 - Q1);
   const float T_max = Q3 + 1.5f * (Q3 - Q1);
   wave_sync();

   for ( int i = lane; i < N; i += WAVE )
   {
      float2 xv = L.X[i];
      float v = j ? xv.y : xv.x;
      if ( L.lab[i] == cluster && (v < T_min || T_max < v) )
      {
         L.lab[i] = marker;
      }
   }
   wave_sync();
}

/* pairwise_clustering.cpp:77-84: labels[i] >= 0 (i < N) take the fit's
 * labels in compact order. */
__device__ void copy_back(const Lds& L, int N)
{
   const int lane = lane_id();
   int jb = 0;
   for ( int i0 = 0; i0 < N; i0 += WAVE )
   {
      int i = i0 + lane;
      int8_t l = (i < N) ? L.lab[i] : (int8_t)-1;
      bool ge = l >= 0;
      uint64_t m = ballot(ge);
      if ( ge ) L.lab[i] = L.w[jb + popc(m & lanes_below())];
      jb += popc(m);
   }
   wave_sync();
}

/* --------------------------------------------------------- chain helpers */

/* Add row `lane` of the staged tile (len values) to acc in index order. */
__device__ __forceinline__ float chain_add(const float* tile, int stride, int nch, int len, float acc)
{
   const int lane = lane_id();
   if ( lane < nch )
   {
      const float* row = tile + lane * stride;
      if ( len == TILE )
      {
         const float4* r4 = (const float4*)row;
#pragma unroll
         for ( int j = 0; j < TILE / 4; ++j )
         {
            float4 v = r4[j];
            acc = acc + v.x;
            acc = acc + v.y;
            acc = acc + v.z;
            acc = acc + v.w;
         }
      }
      else
      {
         for ( int j = 0; j < len; ++j ) acc = acc + row[j];
      }
   }
   return acc;
}

/* ------------------------------------------------------------------ GMM */

struct FitOut
{
   float logL;
   float E;
   uint16_t iters;
   bool ok;
};

template <int KMAX>
struct GmmParams
{
   float mu0[KMAX], mu1[KMAX];
   float a00[KMAX], a01[KMAX], a10[KMAX], a11[KMAX];
   float nrm[KMAX];
   float logpi[KMAX];
};

/* pairwise_linalg.cpp:163-173 + pairwise_gmm.cpp:31-47; false when det <= 0. */
__device__ __forceinline__ bool prepare_cov(float s00, float s01, float s10, float s11,
                                            float& a00, float& a01, float& a10, float& a11, float& nrm)
{
   float det = s00 * s11 - s01 * s10;
   a00 = +s11 / det;
   a01 = -s01 / det;
   a10 = -s10 / det;
   a11 = +s00 / det;
   if ( det <= 0 ) return false;
   nrm = (float)(-0.5 * (KINC_TWO_LOG_2PI + (double)kinc_logf(det)));
   return true;
}

/* E step for all slots: lg[k][m] = log N(x|k) - log p(x), lpx[m] = log p(x)
 * (pairwise_gmm.cpp:54-77,143-187). */
template <int P, int KMAX>
__device__ __forceinline__ void gmm_estep(int N, int K, const float (&x0)[P], const float (&x1)[P],
                                          const GmmParams<KMAX>& g, float (&lg)[KMAX][P], float (&lpx)[P])
{
   const int lane = lane_id();
#pragma unroll
   for ( int m = 0; m < P; ++m )
   {
      if ( m * WAVE >= N ) break;
      if ( lane + m * WAVE < N )
      {
#pragma unroll
         for ( int k = 0; k < KMAX; ++k )
         {
            if ( k < K )
            {
               float xm0 = x0[m] - g.mu0[k];
               float xm1 = x1[m] - g.mu1[k];
               float b0 = g.a00[k] * xm0 + g.a01[k] * xm1;
               float b1 = g.a10[k] * xm0 + g.a11[k] * xm1;
               float dot = xm0 * b0 + xm1 * b1;
               lg[k][m] = g.nrm[k] - 0.5f * dot;
            }
         }
         float mx = -INFINITY;
#pragma unroll
         for ( int k = 0; k < KMAX; ++k )
         {
            if ( k < K )
            {
               float a = g.logpi[k] + lg[k][m];
               mx = (a > mx) ? a : mx;
            }
         }
         float sum = 0.0f;
#pragma unroll
         for ( int k = 0; k < KMAX; ++k )
         {
            if ( k < K ) sum = sum + kinc_expf((g.logpi[k] + lg[k][m]) - mx);
         }
         float lp = mx + kinc_logf(sum);
         lpx[m] = lp;
#pragma unroll
         for ( int k = 0; k < KMAX; ++k )
         {
            if ( k < K ) lg[k][m] = lg[k][m] + (-lp);
         }
      }
   }
}

/* GMM::fit (pairwise_gmm.cpp:367-445) with the per-fit LCG. Writes the
 * fit's labels for i < N into L.w. */
template <int P, int KMAX>
__device__ FitOut gmm_fit(const SimArgs& a, const Lds& L, int N, int K,
                          const float (&x0)[P], const float (&x1)[P], bool want_entropy)
{
   const int lane = lane_id();
   const int STR = a.tile_stride;
   float* tile = L.u;
   GmmParams<KMAX> g;
   FitOut out;
   out.ok = true;
   out.logL = 0.0f;
   out.E = 0.0f;

   uint64_t st = 1;
#pragma unroll
   for ( int k = 0; k < KMAX; ++k )
   {
      if ( k < K )
      {
         int i = lcg(st) % N;
         float2 c = L.X[i];
         g.mu0[k] = c.x;
         g.mu1[k] = c.y;
         prepare_cov(1.0f, 0.0f, 0.0f, 1.0f, g.a00[k], g.a01[k], g.a10[k], g.a11[k], g.nrm[k]);
         g.logpi[k] = kinc_logf(1.0f / K);
      }
   }

   float lg[KMAX][P];
   float lpx[P];
   float prevLogL = -INFINITY;
   float curLogL = -INFINITY;
   int e_steps = 0;
   bool converged = false;
   bool failed = false;

   for ( int t = 0; t < 100; ++t )
   {
      gmm_estep<P, KMAX>(N, K, x0, x1, g, lg, lpx);
      e_steps++;

      float gmax[KMAX];
#pragma unroll
      for ( int k = 0; k < KMAX; ++k )
      {
         if ( k < K )
         {
            float v = -INFINITY;
#pragma unroll
            for ( int m = 0; m < P; ++m )
            {
               if ( lane + m * WAVE < N ) v = (lg[k][m] > v) ? lg[k][m] : v;
            }
            gmax[k] = wave_max(v);
         }
      }

      /* pass A: chain 0 = logL, 1+k = Gamma_k sum, 1+K+2k / 2+K+2k = mu_k sums */
      const int nA = 1 + 3 * K;
      float acc = 0.0f;
#pragma unroll
      for ( int m = 0; m < P; ++m )
      {
         if ( m * WAVE >= N ) break;
         if ( lane + m * WAVE < N )
         {
            const int col = lane;
            tile[col] = lpx[m];
#pragma unroll
            for ( int k = 0; k < KMAX; ++k )
            {
               if ( k < K )
               {
                  tile[(1 + k) * STR + col] = kinc_expf(lg[k][m] - gmax[k]);
                  float gm = kinc_expf(lg[k][m]);
                  lg[k][m] = gm;
                  tile[(1 + K + 2 * k) * STR + col] = gm * x0[m];
                  tile[(2 + K + 2 * k) * STR + col] = gm * x1[m];
               }
            }
         }
         wave_sync();
         int len = N - m * WAVE;
         len = len < TILE ? len : TILE;
         acc = chain_add(tile, STR, nA, len, acc);
         wave_sync();
      }

      prevLogL = curLogL;
      curLogL = bcast(acc, 0);
      if ( fabsf(curLogL - prevLogL) < 1e-8f )
      {
         converged = true;
         break;
      }

      /* calcLogGammaK / calcLogGammaSum (pairwise_gmm.cpp:194-248) */
      float logGamma[KMAX];
#pragma unroll
      for ( int k = 0; k < KMAX; ++k )
      {
         if ( k < K ) logGamma[k] = gmax[k] + kinc_logf(bcast(acc, 1 + k));
      }
      float mx = -INFINITY;
#pragma unroll
      for ( int k = 0; k < KMAX; ++k )
      {
         if ( k < K )
         {
            float v = g.logpi[k] + logGamma[k];
            mx = (v > mx) ? v : mx;
         }
      }
      float sum = 0.0f;
#pragma unroll
      for ( int k = 0; k < KMAX; ++k )
      {
         if ( k < K ) sum = sum + kinc_expf((g.logpi[k] + logGamma[k]) - mx);
      }
      const float logGammaSum = mx + kinc_logf(sum);

      /* performMStep (pairwise_gmm.cpp:255-316) */
      float Gam[KMAX];
#pragma unroll
      for ( int k = 0; k < KMAX; ++k )
      {
         if ( k < K )
         {
            g.logpi[k] = g.logpi[k] + (logGamma[k] - logGammaSum);
            Gam[k] = kinc_expf(logGamma[k]);
            const float cm = 1.0f / Gam[k];
            g.mu0[k] = bcast(acc, 1 + K + 2 * k) * cm;
            g.mu1[k] = bcast(acc, 2 + K + 2 * k) * cm;
         }
      }

      /* pass B: chain 3k / 3k+1 / 3k+2 = Sigma_k entries 00, 01 (= 10), 11 */
      const int nB = 3 * K;
      float accB = 0.0f;
#pragma unroll
      for ( int m = 0; m < P; ++m )
      {
         if ( m * WAVE >= N ) break;
         if ( lane + m * WAVE < N )
         {
            const int col = lane;
#pragma unroll
            for ( int k = 0; k < KMAX; ++k )
            {
               if ( k < K )
               {
                  float xm0 = x0[m] - g.mu0[k];
                  float xm1 = x1[m] - g.mu1[k];
                  float gm = lg[k][m];
                  tile[(3 * k) * STR + col] = gm * (xm0 * xm0);
                  tile[(3 * k + 1) * STR + col] = gm * (xm0 * xm1);
                  tile[(3 * k + 2) * STR + col] = gm * (xm1 * xm1);
               }
            }
         }
         wave_sync();
         int len = N - m * WAVE;
         len = len < TILE ? len : TILE;
         accB = chain_add(tile, STR, nB, len, accB);
         wave_sync();
      }

#pragma unroll
      for ( int k = 0; k < KMAX; ++k )
      {
         if ( k < K && !failed )
         {
            const float cs = 1.0f / Gam[k];
            float s00 = bcast(accB, 3 * k) * cs;
            float s01 = bcast(accB, 3 * k + 1) * cs;
            float s11 = bcast(accB, 3 * k + 2) * cs;
            if ( !prepare_cov(s00, s01, s01, s11, g.a00[k], g.a01[k], g.a10[k], g.a11[k], g.nrm[k]) )
            {
               failed = true;
            }
         }
      }
      if ( failed ) break;
   }

   out.iters = (uint16_t)((e_steps & 0x0fff) | (converged ? (1u << 12) : 0u) | (failed ? (1u << 13) : 0u));
   if ( failed )
   {
      out.ok = false;
      return out;
   }
   out.logL = curLogL;

   /* A converged fit labels from the log-domain responsibilities of its last
    * E step (recomputed: the registers were exponentiated by pass A); a fit
    * that ran 100 iterations labels from the exponentiated ones. */
   if ( converged )
   {
      gmm_estep<P, KMAX>(N, K, x0, x1, g, lg, lpx);
   }

   /* calcLabels (pairwise_gmm.cpp:323-341) and the entropy terms (:348-360) */
   float ent[P];
#pragma unroll
   for ( int m = 0; m < P; ++m )
   {
      int i = lane + m * WAVE;
      ent[m] = 0.0f;
      if ( i < N )
      {
         int max_k = -1;
         float mg = -INFINITY;
         float sel = NAN;
#pragma unroll
         for ( int k = 0; k < KMAX; ++k )
         {
            if ( k < K && mg < lg[k][m] )
            {
               max_k = k;
               mg = lg[k][m];
               sel = lg[k][m];
            }
         }
         L.w[i] = (int8_t)max_k;
         ent[m] = (max_k >= 0) ? kinc_logf(sel) : NAN;
      }
   }
   wave_sync();

   if ( want_entropy )
   {
      float accE = 0.0f;
#pragma unroll
      for ( int m = 0; m < P; ++m )
      {
         if ( m * WAVE >= N ) break;
         if ( lane + m * WAVE < N ) tile[lane] = ent[m];
         wave_sync();
         int len = N - m * WAVE;
         len = len < TILE ? len : TILE;
         accE = chain_add(tile, STR, 1, len, accE);
         wave_sync();
      }
      out.E = bcast(accE, 0);
   }
   return out;
}

/* --------------------------------------------------------------- KMeans */

__device__ __forceinline__ float diff_norm(float ax, float ay, float bx, float by)
{
   /* pairwise_linalg.cpp:97-104 */
   float dist = 0.0f;
   dist = dist + (ax - bx) * (ax - bx);
   dist = dist + (ay - by) * (ay - by);
   return sqrtf(dist);
}

/* KMeans::fit (pairwise_kmeans.cpp:12-127) with the per-fit LCG. */
template <int P, int KMAX>
__device__ FitOut kmeans_fit(const SimArgs& a, const Lds& L, int N, int K,
                             const float (&x0)[P], const float (&x1)[P])
{
   const int lane = lane_id();
   const int STR = a.tile_stride;
   float* tile = L.u;
   FitOut out;
   out.ok = true;
   out.E = 0.0f;
   uint64_t st = 1;
   float bestLogL = -INFINITY;
   int iters = 0;
   float m0[KMAX], m1[KMAX];
   int8_t y[P];

   for ( int init = 0; init < 10; ++init )
   {
#pragma unroll
      for ( int k = 0; k < KMAX; ++k )
      {
         if ( k < K )
         {
            int i = lcg(st) % N;
            float2 c = L.X[i];
            m0[k] = c.x;
            m1[k] = c.y;
         }
      }
#pragma unroll
      for ( int m = 0; m < P; ++m ) y[m] = 0;

      for ( int t = 0; t < 300; ++t )
      {
         iters++;
         int8_t yn[P];
         bool diff = false;
#pragma unroll
         for ( int m = 0; m < P; ++m )
         {
            yn[m] = 0;
            if ( lane + m * WAVE < N )
            {
               int min_k = -1;
               float min_d = 0.0f;
#pragma unroll
               for ( int k = 0; k < KMAX; ++k )
               {
                  if ( k < K )
                  {
                     float d = diff_norm(x0[m], x1[m], m0[k], m1[k]);
                     if ( min_k == -1 || d < min_d )
                     {
                        min_k = k;
                        min_d = d;
                     }
                  }
               }
               yn[m] = (int8_t)min_k;
               diff = diff || (yn[m] != y[m]);
            }
         }
         if ( ballot(diff) == 0 ) break;
#pragma unroll
         for ( int m = 0; m < P; ++m ) y[m] = yn[m];

         /* means: chain 2k / 2k+1 over members in index order */
         const int nch = 2 * K;
         float acc = 0.0f;
#pragma unroll
         for ( int m = 0; m < P; ++m )
         {
            if ( m * WAVE >= N ) break;
            if ( lane + m * WAVE < N )
            {
#pragma unroll
               for ( int k = 0; k < KMAX; ++k )
               {
                  if ( k < K )
                  {
                     bool mem = y[m] == k;
                     tile[(2 * k) * STR + lane] = mem ? x0[m] : 0.0f;
                     tile[(2 * k + 1) * STR + lane] = mem ? x1[m] : 0.0f;
                  }
               }
            }
            wave_sync();
            int len = N - m * WAVE;
            len = len < TILE ? len : TILE;
            acc = chain_add(tile, STR, nch, len, acc);
            wave_sync();
         }
#pragma unroll
         for ( int k = 0; k < KMAX; ++k )
         {
            if ( k < K )
            {
               int nk = 0;
#pragma unroll
               for ( int m = 0; m < P; ++m )
               {
                  nk += popc(ballot(lane + m * WAVE < N && y[m] == k));
               }
               const float c = 1.0f / (float)nk;
               m0[k] = bcast(acc, 2 * k) * c;
               m1[k] = bcast(acc, 2 * k + 1) * c;
            }
         }
      }

      /* computeLogLikelihood (pairwise_kmeans.cpp:106-127): one chain over
       * the members in (k, i) order */
      float* seq = L.u;
      int kbase = 0;
#pragma unroll
      for ( int k = 0; k < KMAX; ++k )
      {
         if ( k < K )
         {
            int cnt = 0;
#pragma unroll
            for ( int m = 0; m < P; ++m )
            {
               bool mem = (lane + m * WAVE < N) && y[m] == k;
               uint64_t bm = ballot(mem);
               if ( mem )
               {
                  float d = diff_norm(x0[m], x1[m], m0[k], m1[k]);
                  seq[kbase + cnt + popc(bm & lanes_below())] = d * d;
               }
               cnt += popc(bm);
            }
            kbase += cnt;
         }
      }
      wave_sync();
      float Ssum = 0.0f;
      if ( lane == 0 )
      {
         for ( int j = 0; j < kbase; ++j ) Ssum = Ssum + seq[j];
      }
      wave_sync();
      const float logL = -bcast(Ssum, 0);

      if ( bestLogL < logL )
      {
         bestLogL = logL;
#pragma unroll
         for ( int m = 0; m < P; ++m )
         {
            int i = lane + m * WAVE;
            if ( i < N ) L.w[i] = y[m];
         }
         wave_sync();
      }
   }
   out.logL = bestLogL;
   out.iters = (uint16_t)(iters & 0x0fff);
   return out;
}

/* ------------------------------------------------- stale X (Pearson) */

/* Last block-relative pair d < qp with blockN[d] > v, or -1. */
__device__ int find_prev_greater(const SimArgs& a, int qp, int v)
{
   const int lane = lane_id();
   if ( qp <= 0 ) return -1;
   const int c = (qp - 1) >> 6;
   {
      int idx = c * WAVE + lane;
      bool ok = idx < qp && a.blockN[idx] > v;
      uint64_t m = ballot(ok);
      if ( m ) return c * WAVE + 63 - __clzll(m);
   }
   for ( int cb = c - 1; cb >= 0; cb -= WAVE )
   {
      int cc = cb - lane;
      bool ok = cc >= 0 && a.chunkMax[cc] > v;
      uint64_t m = ballot(ok);
      if ( m )
      {
         int ch = cb - (__ffsll((unsigned long long)m) - 1);
         int idx = ch * WAVE + lane;
         uint64_t m2 = ballot(a.blockN[idx] > v);
         return ch * WAVE + 63 - __clzll(m2);
      }
   }
   return -1;
}

/*
 * The reference's X workspace is allocated once per block and each pair only
 * overwrites its first N entries (similarity_serial.cpp:42,132), while
 * Pearson reads X at full sample indices (pairwise_pearson.cpp:26-31).  Entry
 * i >= N of pair q therefore holds entry i of the latest earlier pair of the
 * block with more than i valid samples, or zero.  Rebuild that view in L.X.
 */
__device__ void fill_stale(const SimArgs& a, const Lds& L, int q, int N)
{
   const int S = a.S;
   const int lane = lane_id();
   int v = N;
   int qp = q;
   while ( v < S )
   {
      int d = find_prev_greater(a, qp, v);
      if ( d < 0 )
      {
         for ( int i = v + lane; i < S; i += WAVE ) L.X[i] = make_float2(0.0f, 0.0f);
         break;
      }
      const int hi = a.blockN[d];
      int gx, gy;
      pair_xy(a.block_start + d, gx, gy);
      const float* r1 = a.emx + (size_t)gx * (size_t)S;
      const float* r2 = a.emx + (size_t)gy * (size_t)S;
      int base = 0;
      for ( int s0 = 0; s0 < S && base < hi; s0 += WAVE )
      {
         int s = s0 + lane;
         bool in = s < S;
         float g1 = in ? r1[s] : 0.0f;
         float g2 = in ? r2[s] : 0.0f;
         bool valid = in && !isnan(g1) && !isnan(g2) && !(g1 < a.min_expr || g2 < a.min_expr);
         uint64_t m = ballot(valid);
         int r = base + popc(m & lanes_below());
         if ( valid && r >= v && r < hi ) L.X[r] = make_float2(g1, g2);
         base += popc(m);
      }
      v = hi;
      qp = d;
   }
   wave_sync();
}

/* -------------------------------------------------------------- Pearson */

/* pairwise_pearson.cpp:12-52 for clusters 0..K-1 at once: chain 5k+t holds
 * sum x, sum y, sum x^2, sum y^2, sum xy of cluster k. */
template <int KMAX>
__device__ void pearson_all(const SimArgs& a, const Lds& L, int K, float* corr)
{
   const int S = a.S;
   const int lane = lane_id();
   const int nch = 5 * K;
   float acc = 0.0f;
   if ( lane < nch )
   {
      const int k = lane / 5;
      const int t = lane - 5 * k;
      for ( int i = 0; i < S; ++i )
      {
         int8_t l = L.lab[i];
         float2 v = L.X[i];
         float term = (t == 0) ? v.x
                    : (t == 1) ? v.y
                    : (t == 2) ? v.x * v.x
                    : (t == 3) ? v.y * v.y
                    : v.x * v.y;
         acc = (l == k) ? acc + term : acc;
      }
   }
#pragma unroll
   for ( int k = 0; k < KMAX; ++k )
   {
      if ( k < K )
      {
         int n = 0;
         for ( int i0 = 0; i0 < S; i0 += WAVE )
         {
            int i = i0 + lane;
            n += popc(ballot(i < S && L.lab[i] == k));
         }
         float r = NAN;
         if ( n >= a.min_samples )
         {
            const float fn = (float)n;
            float sx = bcast(acc, 5 * k), sy = bcast(acc, 5 * k + 1);
            float sxx = bcast(acc, 5 * k + 2), syy = bcast(acc, 5 * k + 3);
            float sxy = bcast(acc, 5 * k + 4);
            r = (fn * sxy - sx * sy) / sqrtf((fn * sxx - sx * sx) * (fn * syy - sy * sy));
         }
         corr[k] = r;
      }
   }
}

/* ------------------------------------------------------------- Spearman */

/* pairwise_spearman.cpp:27-86 */
__device__ float spearman_k(const SimArgs& a, const Lds& L, int k)
{
   const int S = a.S;
   const int lane = lane_id();
   float* kx = L.u;
   float* ky = L.u + a.np2;
   float* kr = L.u + 2 * a.np2;
   int n = 0;
   int jb = 0;
   for ( int i0 = 0; i0 < S; i0 += WAVE )
   {
      int i = i0 + lane;
      int8_t l = (i < S) ? L.lab[i] : (int8_t)-1;
      bool ge = l >= 0;
      bool in = l == k;
      uint64_t mge = ballot(ge);
      uint64_t min = ballot(in);
      if ( in )
      {
         int j = jb + popc(mge & lanes_below());
         int pos = n + popc(min & lanes_below());
         float2 v = L.X[j];
         kx[pos] = v.x;
         ky[pos] = v.y;
         kr[pos] = (float)(pos + 1);
      }
      jb += popc(mge);
      n += popc(min);
   }
   if ( n < a.min_samples )
   {
      wave_sync();
      return NAN;
   }
   const int n2 = next_pow2(n);
   for ( int i = n + lane; i < n2; i += WAVE )
   {
      kx[i] = INFINITY;
      ky[i] = INFINITY;
      kr[i] = 0.0f;
   }
   wave_sync();
   bitonic_sort<true>(kx, ky, n2);
   bitonic_sort<true>(ky, kr, n2);

   uint32_t d = 0;
   for ( int i = lane; i < n; i += WAVE )
   {
      int tmp = (int)((float)(i + 1) - kr[i]);
      d += (uint32_t)(tmp * tmp);
   }
   d = wave_sum_u32(d);
   wave_sync();
   const uint32_t un = (uint32_t)n;
   const int32_t denom = (int32_t)(un * (un * un - 1u));
   return (float)(1.0 - 6.0 * (double)(int32_t)d / (double)denom);
}

/* -------------------------------------------------------------- driver */

template <int P, int KMAX>
__device__ void process_pair(const SimArgs& a, int q)
{
   const Lds L = lds_view(a);
   const int lane = lane_id();
   const int S = a.S;
   const int MC = a.max_k;

   int gx, gy;
   pair_xy(a.block_start + q, gx, gy);
   const int N = fetch_pair(a, gx, gy, L);

   /* iteration record of the fit with k+1 clusters lives in lane k */
   uint16_t itv = 0;

   int K = 1;
   if ( a.clus != 0 )
   {
      if ( a.pre_out )
      {
         mark_outliers(L, N, 0, 0, -7);
         mark_outliers(L, N, 1, 0, -7);
      }
      int bestK = 0;
      if ( N >= a.min_samples )
      {
         float x0[P], x1[P];
#pragma unroll
         for ( int m = 0; m < P; ++m )
         {
            int i = lane + m * WAVE;
            float2 v = (i < N) ? L.X[i] : make_float2(0.0f, 0.0f);
            x0[m] = v.x;
            x1[m] = v.y;
         }
         float bestValue = INFINITY;
         for ( int Kc = a.min_k; Kc <= a.max_k; ++Kc )
         {
            FitOut f = (a.clus == 1)
               ? gmm_fit<P, KMAX>(a, L, N, Kc, x0, x1, a.crit == 1)
               : kmeans_fit<P, KMAX>(a, L, N, Kc, x0, x1);
            if ( lane == Kc - 1 ) itv = f.iters;
            if ( !f.ok ) continue;

            /* computeBIC / computeICL (pairwise_clustering.cpp:155-172) */
            const int p = Kc * 7;
            double v = a.logN[N] * (double)p - (double)(2.0f * f.logL);
            if ( a.crit == 1 ) v = v - (double)(2.0f * f.E);
            const float value = (float)v;
            if ( value < bestValue )
            {
               bestK = Kc;
               bestValue = value;
               copy_back(L, N);
            }
         }
      }
      if ( bestK > 1 && a.post_out )
      {
         for ( int k = 0; k < bestK; ++k )
         {
            mark_outliers(L, N, 0, (int8_t)k, -8);
            mark_outliers(L, N, 1, (int8_t)k, -8);
         }
      }
      K = bestK;
   }

   float corr[KMAX];
#pragma unroll
   for ( int k = 0; k < KMAX; ++k ) corr[k] = NAN;
   if ( K > 0 )
   {
      if ( a.corr == 0 )
      {
         if ( N < S ) fill_stale(a, L, q, N);
         pearson_all<KMAX>(a, L, K, corr);
      }
      else
      {
#pragma unroll
         for ( int k = 0; k < KMAX; ++k )
         {
            if ( k < K ) corr[k] = spearman_k(a, L, k);
         }
      }
   }

   if ( lane == 0 ) a.K_out[q] = (int8_t)K;
   float* co = a.corr_out + (size_t)q * MC;
#pragma unroll
   for ( int k = 0; k < KMAX; ++k )
   {
      if ( k < MC && lane == k ) co[k] = corr[k];
   }
   if ( a.iters_out && lane < MC ) a.iters_out[(size_t)q * MC + lane] = itv;
   if ( a.labels_out )
   {
      int8_t* lo = a.labels_out + (size_t)q * S;
      for ( int s = lane; s < S; s += WAVE ) lo[s] = L.lab[s];
   }
   wave_sync();
}

} // namespace

/* ------------------------------------------------------------- kernels */

template <int P, int KMAX>
__global__ __launch_bounds__(64) void k_pairs_plain(SimArgs a)
{
   const int q = a.first + (int)blockIdx.x;
   if ( q >= a.first + a.launch_count ) return;
   process_pair<P, KMAX>(a, q);
}

template <int P, int KMAX>
__global__ __launch_bounds__(64) void k_pairs_persistent(SimArgs a)
{
   const int len = *a.list_len;
   for ( ;; )
   {
      int idx = 0;
      if ( lane_id() == 0 ) idx = atomicAdd(a.list_head, 1);
      idx = bcast_i(idx, 0);
      if ( idx >= len ) break;
      process_pair<P, KMAX>(a, a.list[idx]);
   }
}

#endif
