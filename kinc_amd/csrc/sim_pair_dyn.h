/*
 * sim_pair_dyn.h — fits and correlations with a cluster count known only at
 * run time, for max_clusters 9..64 (the reference's limit,
 * src/core/similarity_input.cpp:180-189, Pairwise::Index::MAX_CLUSTER_SIZE
 * = 64, pairwise_index.h:36).  Included by sim_pair_impl.h; used by the
 * memory-store instance built with KMAX = 64.
 *
 * The per-point state stays O(1) whatever K is: X and log p(x) of the last
 * E step in the wave's scratch slice; every log responsibility is
 * recomputed from X, log p(x) and that E step's parameters with the same
 * operations in the same order as the fixed-K fits, so the bits are those
 * of the reference's order of operations.  Component parameters live in LDS
 * (the E step's set and the one the M step builds), and the 1+3K (pass A)
 * and 3K (pass B) sequential chains run in groups of at most 64 rows, one
 * row per lane.
 */
#ifndef KINC_SIM_PAIR_DYN_H
#define KINC_SIM_PAIR_DYN_H

/* chain-row groups: pass A has at most 1 + 3*64 = 193 rows, K-means 2*64 */
#define DYN_MAX_GROUPS 4

/* LDS carve-up of the KMAX = 64 instance (offsets from the parameter base):
 * E-step parameters [64][8], M-step parameters [64][8], per-component
 * maxima [64]; the chain tile has 64 rows */
#define DYN_PRM_NEW (64 * 8)
#define DYN_GMAX (128 * 8)

/* log N(x | component k) from parameters in LDS (a broadcast read) */
__device__ __forceinline__ float dyn_logpdf(const float* prm, int k, float x0, float x1)
{
   const float4 A = *(const float4*)(prm + 8 * k);
   const float4 B = *(const float4*)(prm + 8 * k + 4);
   return gmm_logpdf(x0, x1, A, B);
}

/* value of chain row R held by the lanes' group accumulators */
__device__ __forceinline__ float dyn_gather(const float (&acc)[DYN_MAX_GROUPS], int R)
{
   float v = 0.0f;
#pragma unroll
   for ( int g = 0; g < DYN_MAX_GROUPS; ++g )
   {
      const float t = __shfl(acc[g], R & 63);
      v = (R >> 6) == g ? t : v;
   }
   return v;
}

/* per-point log p(x) of the last E step, [slot][lane] in the scratch slice */
__device__ __forceinline__ float& dyn_lpx(const Lds& L, int m)
{
   return L.lgbuf[((size_t)m << 6) + lane_id()];
}

__device__ __forceinline__ float dyn_x0(const Lds& L, int N, int m)
{
   const int i = lane_id() + m * WAVE;
   return (i < N) ? L.X[i].x : 0.0f;
}

__device__ __forceinline__ float dyn_x1(const Lds& L, int N, int m)
{
   const int i = lane_id() + m * WAVE;
   return (i < N) ? L.X[i].y : 0.0f;
}

/* the chain length of slot m: min(N - 64m, 64), opaque to hoisting */
__device__ __forceinline__ int dyn_len(int N, int m)
{
   int len = N - m * WAVE;
   len = len < TILE ? len : TILE;
   asm volatile("" : "+s"(len));
   return len;
}

/* GMM::fit (pairwise_gmm.cpp:367-445) for a run-time K <= 64, the same
 * arithmetic as gmm_fit; labels of the fit for i < N go to L.w. */
__device__ __noinline__ FitOut gmm_fit_dyn(const Lds& L, int N, int K, bool want_entropy)
{
   /* wave-uniform (arguments of a call arrive in vector registers) */
   N = __builtin_amdgcn_readfirstlane(N);
   K = __builtin_amdgcn_readfirstlane(K);
   const int lane = lane_id();
   constexpr int STR = KINC_TILE_STRIDE;
   const int nslots = (N + WAVE - 1) / WAVE;
   float* tile = L.tile;
   float* prm = L.prm;                  /* parameters of the E step */
   float* prmN = L.prm + DYN_PRM_NEW;   /* parameters the M step builds */
   float* gmax = L.prm + DYN_GMAX;
   const int rowsA = 1 + 3 * K;
   const int rowsB = 3 * K;
   const int groupsA = (rowsA + 63) >> 6;
   const int groupsB = (rowsB + 63) >> 6;
   FitOut out;
   out.ok = true;
   out.logL = 0.0f;
   out.E = 0.0f;

   /* initialisation (pairwise_gmm.cpp:372-392): K draws of rand() % N */
   {
      uint64_t rs = 1;
      int my = 0;
      for ( int k = 0; k < K; ++k )
      {
         const int i = lcg(rs) % N;
         my = (lane == k) ? i : my;
      }
      if ( lane < K )
      {
         const float2 c = L.X[my];
         float a00, a01, a10, a11, nrm;
         prepare_cov(1.0f, 0.0f, 0.0f, 1.0f, a00, a01, a10, a11, nrm);
         float* p = prm + 8 * lane;
         *(float4*)p = make_float4(c.x, c.y, a00, a01);
         *(float4*)(p + 4) = make_float4(a10, a11, nrm, kinc_logf(1.0f / K));
      }
      wave_sync(true);
   }

   float prevLogL = -INFINITY;
   float curLogL = -INFINITY;
   int e_steps = 0;
   bool converged = false;
   bool failed = false;

   for ( int t = 0; t < 100; ++t )
   {
      /* E step: log p(x) per point (pairwise_gmm.cpp:143-187) */
      for ( int m = 0; m < nslots; ++m )
      {
         const float x0 = dyn_x0(L, N, m);
         const float x1 = dyn_x1(L, N, m);
         float mx = -INFINITY;
         for ( int k = 0; k < K; ++k )
         {
            const float v = prm[8 * k + PRM_LPI] + dyn_logpdf(prm, k, x0, x1);
            mx = fmaxf(mx, v);
         }
         float sum = 0.0f;
         for ( int k = 0; k < K; ++k )
         {
            sum = sum + kinc_expf_le0((prm[8 * k + PRM_LPI] + dyn_logpdf(prm, k, x0, x1)) - mx);
         }
         dyn_lpx(L, m) = mx + kinc_logf_ge1(sum);
      }
      e_steps++;
      wave_sync(true);

      /* each component's largest log responsibility over the points */
      for ( int k = 0; k < K; ++k )
      {
         float v = -INFINITY;
         for ( int m = 0; m < nslots; ++m )
         {
            const float l = dyn_logpdf(prm, k, dyn_x0(L, N, m), dyn_x1(L, N, m)) + (-dyn_lpx(L, m));
            v = (lane + m * WAVE < N) ? fmaxf(v, l) : v;
         }
         v = wave_max(v);
         if ( lane == 0 ) gmax[k] = v;
      }
      wave_sync(true);

      /* pass A: row 0 = logL, rows 1+3k, 2+3k, 3+3k = Gamma_k, mu_k sums */
      float accA[DYN_MAX_GROUPS] = {0.0f, 0.0f, 0.0f, 0.0f};
      for ( int m = 0; m < nslots; ++m )
      {
         const float lp = dyn_lpx(L, m);
         const float xa = dyn_x0(L, N, m);
         const float xb = dyn_x1(L, N, m);
         const int len = dyn_len(N, m);
#pragma unroll
         for ( int g = 0; g < DYN_MAX_GROUPS; ++g )
         {
            if ( g >= groupsA ) break;
            const int r0 = 64 * g;
            if ( g == 0 ) tile[lane] = lp;
            const int klo = r0 >= 3 ? (r0 - 3) / 3 : 0;
            int khi = (r0 + 63 - 1) / 3;
            khi = khi < K - 1 ? khi : K - 1;
            for ( int k = klo; k <= khi; ++k )
            {
               const float l = dyn_logpdf(prm, k, xa, xb) + (-lp);
               const float gm = kinc_expf(l);
               const int r = 1 + 3 * k - r0;
               if ( r >= 0 && r < 64 ) tile[r * STR + lane] = kinc_expf_le0(l - gmax[k]);
               if ( r + 1 >= 0 && r + 1 < 64 ) tile[(r + 1) * STR + lane] = gm * xa;
               if ( r + 2 >= 0 && r + 2 < 64 ) tile[(r + 2) * STR + lane] = gm * xb;
            }
            wave_sync(true);
            const int nr = rowsA - r0 < 64 ? rowsA - r0 : 64;
            accA[g] = chain_add(tile, STR, nr, len, accA[g]);
            wave_sync(true);
         }
      }

      prevLogL = curLogL;
      curLogL = bcast(accA[0], 0);
      if ( fabsf(curLogL - prevLogL) < 1e-8f )
      {
         converged = true;
         break;
      }

      /* M step (pairwise_gmm.cpp:194-316): component k on lane k */
      const int kk = lane < K ? lane : 0;
      const float gsum = dyn_gather(accA, 1 + 3 * kk);
      const float mxs = dyn_gather(accA, 2 + 3 * kk);
      const float mys = dyn_gather(accA, 3 + 3 * kk);
      const float logGamma = gmax[kk] + kinc_logf_ge1(gsum);
      const float lpk = prm[8 * kk + PRM_LPI];
      const float v = lpk + logGamma;
      float mx = -INFINITY;
      for ( int k = 0; k < K; ++k )
      {
         const float vk = bcast(v, k);
         mx = (vk > mx) ? vk : mx;
      }
      const float ek = kinc_expf_le0(v - mx);
      float sum = 0.0f;
      for ( int k = 0; k < K; ++k ) sum = sum + bcast(ek, k);
      const float logGammaSum = mx + kinc_logf_ge1(sum);
      const float Gam = kinc_expf(logGamma);
      const float cm = 1.0f / Gam;
      const float mu0 = mxs * cm;
      const float mu1 = mys * cm;
      if ( lane < K )
      {
         float* p = prmN + 8 * lane;
         p[PRM_MU0] = mu0;
         p[PRM_MU1] = mu1;
         p[PRM_LPI] = lpk + (logGamma - logGammaSum);
      }
      wave_sync(true);

      /* pass B: rows 3k, 3k+1, 3k+2 = Sigma_k entries 00, 01 (= 10), 11, with
       * gamma from this E step's parameters and the new means */
      float accB[DYN_MAX_GROUPS] = {0.0f, 0.0f, 0.0f, 0.0f};
      for ( int m = 0; m < nslots; ++m )
      {
         const float lp = dyn_lpx(L, m);
         const float xa = dyn_x0(L, N, m);
         const float xb = dyn_x1(L, N, m);
         const int len = dyn_len(N, m);
#pragma unroll
         for ( int g = 0; g < DYN_MAX_GROUPS; ++g )
         {
            if ( g >= groupsB ) break;
            const int r0 = 64 * g;
            const int klo = r0 / 3;
            int khi = (r0 + 63) / 3;
            khi = khi < K - 1 ? khi : K - 1;
            for ( int k = klo; k <= khi; ++k )
            {
               const float gm = kinc_expf(dyn_logpdf(prm, k, xa, xb) + (-lp));
               const float xm0 = xa - prmN[8 * k + PRM_MU0];
               const float xm1 = xb - prmN[8 * k + PRM_MU1];
               const int r = 3 * k - r0;
               if ( r >= 0 && r < 64 ) tile[r * STR + lane] = gm * (xm0 * xm0);
               if ( r + 1 >= 0 && r + 1 < 64 ) tile[(r + 1) * STR + lane] = gm * (xm0 * xm1);
               if ( r + 2 >= 0 && r + 2 < 64 ) tile[(r + 2) * STR + lane] = gm * (xm1 * xm1);
            }
            wave_sync(true);
            const int nr = rowsB - r0 < 64 ? rowsB - r0 : 64;
            accB[g] = chain_add(tile, STR, nr, len, accB[g]);
            wave_sync(true);
         }
      }

      const float s00 = dyn_gather(accB, 3 * kk) * cm;
      const float s01 = dyn_gather(accB, 3 * kk + 1) * cm;
      const float s11 = dyn_gather(accB, 3 * kk + 2) * cm;
      float a00, a01, a10, a11, nrm = 0.0f;
      const bool ok = prepare_cov(s00, s01, s01, s11, a00, a01, a10, a11, nrm);
      failed = ballot(lane < K && !ok) != 0;
      if ( failed ) break;
      if ( lane < K )
      {
         float* p = prmN + 8 * lane;
         *(float4*)(p + 0) = make_float4(mu0, mu1, a00, a01);
         p[PRM_A10] = a10;
         p[PRM_A11] = a11;
         p[PRM_NRM] = nrm;
      }
      wave_sync(true);
      /* the new parameters become the next E step's; after the 100th M step
       * the labels read the responsibilities of the last E step, so its
       * parameters stay */
      if ( t < 99 && lane < K )
      {
         *(float4*)(prm + 8 * lane) = *(const float4*)(prmN + 8 * lane);
         *(float4*)(prm + 8 * lane + 4) = *(const float4*)(prmN + 8 * lane + 4);
      }
      wave_sync(true);
   }

   out.iters = (uint16_t)((e_steps & 0x0fff) | (converged ? (1u << 12) : 0u) | (failed ? (1u << 13) : 0u));
   if ( failed )
   {
      out.ok = false;
      return out;
   }
   out.logL = curLogL;

   /* calcLabels (pairwise_gmm.cpp:323-341): a converged fit compares the
    * log responsibilities of its last E step, a 100-iteration fit their
    * exponentials; the entropy terms (:348-360) go where log p(x) was */
   for ( int m = 0; m < nslots; ++m )
   {
      const int i = lane + m * WAVE;
      const float lp = dyn_lpx(L, m);
      const float xa = dyn_x0(L, N, m);
      const float xb = dyn_x1(L, N, m);
      int max_k = -1;
      float mg = -INFINITY;
      for ( int k = 0; k < K; ++k )
      {
         const float l = dyn_logpdf(prm, k, xa, xb) + (-lp);
         const float val = converged ? l : kinc_expf(l);
         if ( mg < val )
         {
            max_k = k;
            mg = val;
         }
      }
      if ( i < N ) L.w[i] = (int8_t)max_k;
      if ( want_entropy ) dyn_lpx(L, m) = (max_k >= 0) ? kinc_logf(mg) : NAN;
   }
   wave_sync(true);

   if ( want_entropy )
   {
      float accE = 0.0f;
      for ( int m = 0; m < nslots; ++m )
      {
         tile[lane] = dyn_lpx(L, m);
         wave_sync(true);
         accE = chain_add(tile, STR, 1, dyn_len(N, m), accE);
         wave_sync(true);
      }
      out.E = bcast(accE, 0);
   }
   return out;
}

/* nearest mean in the reference's terms (pairwise_kmeans.cpp:37-54), run-time
 * K: nearest_mean_q's rule with the squared distances recomputed */
__device__ __forceinline__ int dyn_nearest(const float* means, int K, float xa, float xb)
{
   int kq = 0;
   float qmin = diff_sq(xa, xb, means[0], means[1]);
   float before = NAN;   /* smallest q before kq, as in nearest_mean_q */
   for ( int k = 1; k < K; ++k )
   {
      const float q = diff_sq(xa, xb, means[2 * k], means[2 * k + 1]);
      if ( q < qmin )
      {
         before = qmin;
         kq = k;
         qmin = q;
      }
   }
   const bool tie = before <= qmin * (1.0f + 0x1p-20f);
   if ( ballot(tie) != 0 )
   {
      /* the reference's comparison of square roots */
      int min_k = -1;
      float min_d = 0.0f;
      for ( int k = 0; k < K; ++k )
      {
         const float d = sqrtf(diff_sq(xa, xb, means[2 * k], means[2 * k + 1]));
         if ( min_k == -1 || d < min_d )
         {
            min_k = k;
            min_d = d;
         }
      }
      return min_k;
   }
   return kq;
}

/* KMeans::fit (pairwise_kmeans.cpp:12-127) for a run-time K <= 64, the same
 * arithmetic as kmeans_fit; the best init's labels go to L.w. */
__device__ __noinline__ FitOut kmeans_fit_dyn(const SimArgs& a, const Lds& L, int N, int K)
{
   N = __builtin_amdgcn_readfirstlane(N);
   K = __builtin_amdgcn_readfirstlane(K);
   const int lane = lane_id();
   constexpr int STR = KINC_TILE_STRIDE;
   const int nslots = (N + WAVE - 1) / WAVE;
   float* tile = L.tile;
   float* means = L.prm;
   const int rows = 2 * K;
   const int groups = (rows + 63) >> 6;
   FitOut out;
   out.ok = true;
   out.E = 0.0f;
   uint64_t rs = 1;
   float bestLogL = -INFINITY;
   int iters = 0;
   auto Y = [&](int m) -> int8_t& { return L.ybuf[((size_t)m << 6) + lane]; };

   for ( int init = 0; init < 10; ++init )
   {
      int my = 0;
      for ( int k = 0; k < K; ++k )
      {
         const int i = lcg(rs) % N;
         my = (lane == k) ? i : my;
      }
      wave_sync(true);
      if ( lane < K ) *(float2*)(means + 2 * lane) = L.X[my];
      for ( int m = 0; m < nslots; ++m ) Y(m) = 0;
      wave_sync(true);

      for ( int t = 0; t < 300; ++t )
      {
         iters++;
         bool diff = false;
         for ( int m = 0; m < nslots; ++m )
         {
            const int min_k = dyn_nearest(means, K, dyn_x0(L, N, m), dyn_x1(L, N, m));
            /* new labels over the old ones, as in kmeans_fit */
            const int8_t old = Y(m);
            Y(m) = (int8_t)min_k;
            diff = diff || ((lane + m * WAVE < N) && (int8_t)min_k != old);
         }
         if ( ballot(diff) == 0 ) break;
         wave_sync(true);

         /* means: rows 2k, 2k+1 over the members of cluster k in index
          * order, members packed to the front of each row per slot and the
          * rows zero-filled to a common length (adding +0 to a sum that
          * started at +0 changes nothing) */
         float acc[DYN_MAX_GROUPS] = {0.0f, 0.0f, 0.0f, 0.0f};
         int nk = 0;   /* lane k: members of cluster k */
         for ( int m = 0; m < nslots; ++m )
         {
            const float xa = dyn_x0(L, N, m);
            const float xb = dyn_x1(L, N, m);
            const int ym = (lane + m * WAVE < N) ? (int)Y(m) : -1;
#pragma unroll
            for ( int g = 0; g < DYN_MAX_GROUPS; ++g )
            {
               if ( g >= groups ) break;
               const int k0 = 32 * g;
               const int k1 = k0 + 32 < K ? k0 + 32 : K;
               /* the group's rows zeroed by all lanes, then each member
                * written over the front (the wave's LDS stores land in
                * order): every row holds its members, then zeros */
               for ( int r = 0; r < 2 * (k1 - k0); ++r ) tile[r * STR + lane] = 0.0f;
               int maxc = 0;
               for ( int k = k0; k < k1; ++k )
               {
                  const uint64_t bm = ballot(ym == k);
                  const int ck = popc(bm);
                  nk += (lane == k) ? ck : 0;
                  maxc = ck > maxc ? ck : maxc;
                  if ( ym == k )
                  {
                     const int rank = popc(bm & lanes_below());
                     tile[(2 * (k - k0)) * STR + rank] = xa;
                     tile[(2 * (k - k0) + 1) * STR + rank] = xb;
                  }
               }
               const int len = (maxc + 3) & ~3;
               wave_sync(true);
               acc[g] = chain_add4(tile, STR, 2 * (k1 - k0), len, acc[g]);
               wave_sync(true);
            }
         }
         const int kk = lane < K ? lane : 0;
         const float sx = dyn_gather(acc, 2 * kk);
         const float sy = dyn_gather(acc, 2 * kk + 1);
         const float c = 1.0f / (float)nk;
         if ( lane < K ) *(float2*)(means + 2 * lane) = make_float2(sx * c, sy * c);
         wave_sync(true);
      }

      /* computeLogLikelihood (pairwise_kmeans.cpp:106-127): one chain over
       * the members in (k, i) order */
      float* seq = L.u;
      int kbase = 0;
      for ( int k = 0; k < K; ++k )
      {
         const float2 mk = *(const float2*)(means + 2 * k);
         int cnt = 0;
         for ( int m = 0; m < nslots; ++m )
         {
            const bool mem = (lane + m * WAVE < N) && Y(m) == k;
            const uint64_t bm = ballot(mem);
            if ( mem )
            {
               const float d = diff_norm(dyn_x0(L, N, m), dyn_x1(L, N, m), mk.x, mk.y);
               seq[kbase + cnt + popc(bm & lanes_below())] = d * d;
            }
            cnt += popc(bm);
         }
         kbase += cnt;
      }
      wave_sync(true);
      float Ssum = 0.0f;
      if ( lane < KINC_MIN_ACTIVE_LANES )
      {
         for ( int j = 0; j < kbase; ++j ) Ssum = Ssum + seq[j];
      }
      const float logL = -bcast(Ssum, 0);
      wave_sync(true);

      if ( bestLogL < logL )
      {
         bestLogL = logL;
         for ( int m = 0; m < nslots; ++m )
         {
            const int i = lane + m * WAVE;
            if ( i < N ) L.w[i] = Y(m);
         }
         wave_sync(true);
      }
   }
   out.logL = bestLogL;
   out.iters = (uint16_t)(iters & 0x0fff);
   return out;
}

/* Pearson (pairwise_pearson.cpp:12-52) of clusters k0 .. k0+kc-1 (kc <= 12):
 * chain 5j+t holds sum x, sum y, sum x^2, sum y^2, sum xy of cluster k0+j;
 * results go straight to co[k]. */
__device__ __noinline__ void pearson_group(const SimArgs& a, const Lds& L, int k0, int kc, float* co)
{
   k0 = __builtin_amdgcn_readfirstlane(k0);
   kc = __builtin_amdgcn_readfirstlane(kc);
   const int S = a.S;
   const int lane = lane_id();
   const int nch = 5 * kc;
   float acc = 0.0f;
   double dacc = 0.0;
   if ( lane < nch || lane < KINC_MIN_ACTIVE_LANES )
   {
      const int k = k0 + lane / 5;
      const int t = lane - 5 * (lane / 5);
      for ( int i = 0; i < S; ++i )
      {
         const int8_t l = L.lab[i];
         const float2 xv = L.X[i];
         const float x = xv.x, y = xv.y;
         if ( !a.pearson_fp64 )
         {
            const float term = (t == 0) ? x : (t == 1) ? y : (t == 2) ? x * x : (t == 3) ? y * y : x * y;
            acc = (l == k) ? acc + term : acc;
         }
         else
         {
            const float f1 = (t == 1 || t == 3) ? y : x;
            const float f2 = (t < 2) ? 1.0f : (t == 2) ? x : y;
            const double term = (double)f1 * (double)f2;
            dacc = (l == k) ? dacc + term : dacc;
         }
      }
   }
   for ( int j = 0; j < kc; ++j )
   {
      const int k = k0 + j;
      int n = 0;
      for ( int i0 = 0; i0 < S; i0 += WAVE )
      {
         const int i = i0 + lane;
         n += popc(ballot(i < S && L.lab[i] == k));
      }
      float r = NAN;
      if ( n >= a.min_samples && !a.pearson_fp64 )
      {
         const float fn = (float)n;
         const float sx = bcast(acc, 5 * j), sy = bcast(acc, 5 * j + 1);
         const float sxx = bcast(acc, 5 * j + 2), syy = bcast(acc, 5 * j + 3);
         const float sxy = bcast(acc, 5 * j + 4);
         r = (fn * sxy - sx * sy) / sqrtf((fn * sxx - sx * sx) * (fn * syy - sy * sy));
      }
      else if ( n >= a.min_samples )
      {
         const double dn = (double)n;
         const double sx = bcast_d(dacc, 5 * j), sy = bcast_d(dacc, 5 * j + 1);
         const double sxx = bcast_d(dacc, 5 * j + 2), syy = bcast_d(dacc, 5 * j + 3);
         const double sxy = bcast_d(dacc, 5 * j + 4);
         r = (float)((dn * sxy - sx * sy) / sqrt((dn * sxx - sx * sx) * (dn * syy - sy * sy)));
      }
      if ( lane == 0 ) co[k] = r;
   }
   wave_sync(true);
}

#endif
