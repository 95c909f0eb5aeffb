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
#ifdef KINC_CENSUS
#include "../../include/kinc_selftest.h"
#endif

#define WAVE 64
#define TILE 64

/* One wave per workgroup; the second bound of a pair kernel asks for at
 * least W waves per SIMD, which caps its VGPR budget (512 / W).  W is a
 * template parameter of the kernel, so instances built for different W are
 * different kernels. */
#ifndef KINC_PAIR_WAVES
#define KINC_PAIR_WAVES 4
#endif

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

__device__ __forceinline__ double bcast_d(double v, int src)
{
   const uint64_t b = (uint64_t)__double_as_longlong(v);
   const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, src);
   const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), src);
   return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ __forceinline__ int bcast_i(int v, int src)
{
   return __builtin_amdgcn_readlane(v, src);
}

/*
 * NaN-skipping max in the reference's `if (v > max) max = v` sense, of
 * values that are never NaN (per-lane maxima taken with fmaxf, which skips
 * NaN, from -inf), so the order of comparison does not matter, and whose
 * zero sign no caller reads: quads by DPP quad permutations, then half rows
 * and rows by mirrors, then across rows with row_bcast:15 and row_bcast:31
 * (lane 63 ends with the maximum), read back to every lane.  Each step is
 * one v_max_f32 with a DPP source instead of a DPP move, a compare and a
 * select, and no step goes through LDS.
 */
__device__ __forceinline__ float wave_max(float v)
{
   /* each step one v_max_f32 with a DPP source, in place (lanes of rows
    * outside the row mask are not written and keep v); the s_nop gives the
    * two wait states a DPP read needs after the VALU write of its source */
   asm volatile("s_nop 1\n\tv_max_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(v));
   asm volatile("s_nop 1\n\tv_max_f32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf" : "+v"(v));
   asm volatile("s_nop 1\n\tv_max_f32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf" : "+v"(v));
   asm volatile("s_nop 1\n\tv_max_f32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf" : "+v"(v));
   asm volatile("s_nop 1\n\tv_max_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf" : "+v"(v));
   asm volatile("s_nop 1\n\tv_max_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf" : "+v"(v));
   return bcast(v, 63);
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

/*
 * Make this wave's earlier writes to its working set visible to its later
 * reads by other lanes.  Each workgroup is one wave.  LDS operations of a
 * wave complete in execution order, so for an LDS working set ordering the
 * compiler is enough (a wavefront-scope fence); a working set in global
 * memory (memory-store instance) takes the full workgroup barrier with its
 * memory-counter waits.
 */
__device__ __forceinline__ void wave_sync(bool global_mem)
{
   if ( global_mem )
   {
      __syncthreads();
   }
   else
   {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
   }
}

/*
 * Phase census (instrumented build only, KINC_CENSUS defined): per-wave
 * counters in static LDS, added to by lane 0 with LDS atomics (no return, so
 * no wait), flushed to global memory when the wave exits.  A lap reads the
 * shader clock (s_memtime, which waits for the wave's outstanding LDS and
 * scalar loads) and charges the cycles since the previous lap to a phase.
 * Without KINC_CENSUS every macro expands to nothing.
 */
#ifdef KINC_CENSUS
__shared__ unsigned long long g_census[KINC_CENSUS_COUNTERS];

__device__ __forceinline__ void census_add(int c, uint64_t v)
{
   if ( lane_id() == 0 ) atomicAdd(&g_census[c], (unsigned long long)v);
}
#define CENSUS_T(t) uint64_t t = __builtin_amdgcn_s_memtime()
#define CENSUS_RESET(t) (t = __builtin_amdgcn_s_memtime())
#define CENSUS_LAP(ph, t)                                                     \
   do                                                                         \
   {                                                                          \
      const uint64_t n_ = __builtin_amdgcn_s_memtime();                       \
      census_add((ph), n_ - (t));                                             \
      (t) = n_;                                                               \
   } while ( 0 )
#define CENSUS_COUNT(c, v) census_add((c), (uint64_t)(v))
#else
#define CENSUS_T(t) ((void)0)
#define CENSUS_RESET(t) ((void)0)
#define CENSUS_LAP(ph, t) ((void)0)
#define CENSUS_COUNT(c, v) ((void)0)
#endif

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
   float2* X;      /* [S] compact X, then the stale-read view */
   int8_t* lab;    /* [S] full-index labels */
   int8_t* w;      /* [S] work labels of the current fit */
   float* prm;     /* [KMAX][8] component parameters / K-means means (LDS) */
   float* tile;    /* chain-term tile (LDS) */
   float* u;       /* sort buffers / sequence scratch */
   float* lgbuf;   /* memory-store variant: per-point responsibilities */
   int8_t* ybuf;   /* memory-store variant: per-point K-means labels */
};

/*
 * Working set of the pair a wave is processing.  Register-slot instances
 * keep everything in LDS (tile and sort buffers share one region); the
 * memory-store instance (P = 0, any sample count) keeps the per-pair arrays
 * in the wave's slice of a device scratch buffer and only the parameters and
 * the chain tile in LDS.
 */
template <bool MEM>
__device__ __forceinline__ Lds lds_view(const SimArgs& a)
{
   extern __shared__ __align__(16) char lds_raw[];
   Lds l;
   l.prm = (float*)(lds_raw + a.lds_prm);
   l.tile = (float*)(lds_raw + a.lds_tile);
   if ( !MEM )
   {
      l.X = (float2*)(lds_raw + a.lds_x);
      l.lab = (int8_t*)(lds_raw + a.lds_lab);
      l.w = (int8_t*)(lds_raw + a.lds_w);
      l.u = (float*)(lds_raw + a.lds_u);
      l.lgbuf = nullptr;
      l.ybuf = nullptr;
   }
   else
   {
      char* base = a.scratch + (size_t)blockIdx.x * (size_t)a.scratch_stride;
      l.X = (float2*)(base + a.lds_x);
      l.lab = (int8_t*)(base + a.lds_lab);
      l.w = (int8_t*)(base + a.lds_w);
      l.u = (float*)(base + a.lds_u);
      l.lgbuf = (float*)(base + a.scr_lg);
      l.ybuf = (int8_t*)(base + a.scr_y);
   }
   return l;
}

/* ---------------------------------------------------------------- fetch */

/* similarity_serial.cpp:108-141: labels -9/-6/0 and compaction in sample order. */
__device__ __forceinline__ int fetch_pair(const SimArgs& a, int gx, int gy, const Lds& L)
{
   const int S = a.S;
   const float* r1 = a.emx + (size_t)gx * (size_t)S;
   const float* r2 = a.emx + (size_t)gy * (size_t)S;
   const int lane = lane_id();
   int base = 0;
   /* four slots' loads are issued before any of them is used */
   for ( int s0 = 0; s0 < S; s0 += 4 * WAVE )
   {
      float g1[4], g2[4];
#pragma unroll
      for ( int j = 0; j < 4; ++j )
      {
         const int s = s0 + j * WAVE + lane;
         const bool in = s < S;
         g1[j] = in ? r1[s] : 0.0f;
         g2[j] = in ? r2[s] : 0.0f;
      }
#pragma unroll
      for ( int j = 0; j < 4; ++j )
      {
         const int s = s0 + j * WAVE + lane;
         const bool in = s < S;
         int8_t lab = 0;
         bool valid = false;
         if ( in )
         {
            if ( isnan(g1[j]) || isnan(g2[j]) ) lab = -9;
            else if ( g1[j] < a.min_expr || g2[j] < a.min_expr ) lab = -6;
            else valid = true;
            L.lab[s] = lab;
         }
         uint64_t m = ballot(valid);
         if ( valid )
         {
            L.X[base + popc(m & lanes_below())] = make_float2(g1[j], g2[j]);
         }
         base += popc(m);
      }
   }
   wave_sync(L.lgbuf != nullptr);
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
__device__ __forceinline__ void bitonic_sort(float* key, float* pay, int size, bool global_mem)
{
   const int lane = lane_id();
   const int half = size >> 1;
   /* size, ob, ib and t are powers of two: the divisions are shifts */
   for ( int lob = 1; (1 << lob) <= size; ++lob )
   {
      const int lobh = lob - 1;                /* log2(ob / 2) */
      for ( int lib = lob; lib >= 1; --lib )
      {
         const int t = 1 << (lib - 1);
         for ( int i = lane; i < half; i += WAVE )
         {
            int dir = (i >> lobh) & 1;
            int aa = ((i >> (lib - 1)) << lib) + (i & (t - 1));
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
         wave_sync(global_mem);
      }
   }
}

/* ------------------------------------------------------------- outliers */

/* pairwise_clustering.cpp:110-148 */
__device__ __forceinline__ void mark_outliers(const Lds& L, int N, int j, int8_t cluster, int8_t marker)
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
      wave_sync(L.lgbuf != nullptr);
      return;
   }
   int n2 = next_pow2(n);
   for ( int i = n + lane; i < n2; i += WAVE ) key[i] = INFINITY;
   wave_sync(L.lgbuf != nullptr);
   bitonic_sort<false>(key, nullptr, n2, L.lgbuf != nullptr);

   const float Q1 = key[n * 1 / 4];
   const float Q3 = key[n * 3 / 4];
   const float T_min = Q1 - 1.5f * (Q3 - Q1);
   const float T_max = Q3 + 1.5f * (Q3 - Q1);
   wave_sync(L.lgbuf != nullptr);

   for ( int i = lane; i < N; i += WAVE )
   {
      float2 xv = L.X[i];
      float v = j ? xv.y : xv.x;
      if ( L.lab[i] == cluster && (v < T_min || T_max < v) )
      {
         L.lab[i] = marker;
      }
   }
   wave_sync(L.lgbuf != nullptr);
}

/* pairwise_clustering.cpp:77-84: labels[i] >= 0 (i < N) take the fit's
 * labels in compact order. */
__device__ __forceinline__ void copy_back(const Lds& L, int N, const int8_t* w)
{
   const int lane = lane_id();
   int jb = 0;
   for ( int i0 = 0; i0 < N; i0 += WAVE )
   {
      int i = i0 + lane;
      int8_t l = (i < N) ? L.lab[i] : (int8_t)-1;
      bool ge = l >= 0;
      uint64_t m = ballot(ge);
      if ( ge ) L.lab[i] = w[jb + popc(m & lanes_below())];
      jb += popc(m);
   }
   wave_sync(L.lgbuf != nullptr);
}

/* --------------------------------------------------------- chain helpers */

/*
 * A wave64 VALU instruction with 8 or fewer active lanes issues about 5x
 * slower on gfx950 than one with 9 or more (measured: 0.19 vs 0.98 wave
 * instructions per CU cycle for a v_add_f32 stream).  Loops that only a few
 * lanes need therefore run on at least this many lanes, the extra ones
 * computing values that are never read.
 */
#define KINC_MIN_ACTIVE_LANES 16

/* Add the first len (<= TILE) values of one tile row to acc in index order:
 * groups of four float4 reads are issued together (all inside the row's
 * TILE values), and only the groups' values below len are added. */
__device__ __forceinline__ float chain_row_var(const float* row, int len, float acc)
{
   const float4* r4 = (const float4*)row;
   const int n4 = len >> 2;
   for ( int g = 0; g < n4; g += 4 )
   {
      float4 v[4];
#pragma unroll
      for ( int u = 0; u < 4; ++u ) v[u] = r4[g + u];
#pragma unroll
      for ( int u = 0; u < 4; ++u )
      {
         if ( g + u < n4 )
         {
            acc = acc + v[u].x;
            acc = acc + v[u].y;
            acc = acc + v[u].z;
            acc = acc + v[u].w;
         }
      }
   }
   for ( int j = n4 << 2; j < len; ++j ) acc = acc + row[j];
   return acc;
}

/* Add row `lane` of the staged tile (len values) to acc in index order, for
 * lanes < nch; lanes nch..15 repeat the last row and their sums are unused.
 * A full row's reads run AHEAD groups of 16-byte reads ahead of the adds
 * (the schedule is pinned: left to itself the compiler waits on every
 * read), so the chain runs at the dependent-add rate instead of one LDS
 * round trip per four adds; two reads ahead was the fastest where the
 * registers allow (a deeper window costs more registers than it saves,
 * measured), one where the caller keeps more values live (the lockstep
 * 4- and 5-cluster fits). */
template <int AHEAD = 2>
__device__ __forceinline__ float chain_add(const float* tile, int stride, int nch, int len, float acc)
{
   const int lane = lane_id();
   if ( lane < nch || lane < KINC_MIN_ACTIVE_LANES )
   {
      CENSUS_COUNT(KINC_CENSUS_CHAIN_ADDS, len);
      CENSUS_COUNT(KINC_CENSUS_CHAIN_LANES, (uint64_t)len * (uint64_t)__builtin_popcountll(__builtin_amdgcn_read_exec()));
      CENSUS_COUNT(KINC_CENSUS_CHAIN_CALLS, 1);
      const float* row = tile + (lane < nch ? lane : nch - 1) * stride;
      if ( len == TILE )
      {
         constexpr int G = 4 * AHEAD;            /* values per group */
         const float4* r4 = (const float4*)row;
         float4 b[AHEAD];
#pragma unroll
         for ( int u = 0; u < AHEAD; ++u ) b[u] = r4[u];
         /* scheduling groups: 0x100 = LDS reads, 0x002 = VALU */
         __builtin_amdgcn_sched_group_barrier(0x100, AHEAD, 0);
#pragma unroll
         for ( int g = 0; g < TILE / G; ++g )
         {
            float4 nb[AHEAD];
            if ( g + 1 < TILE / G )
            {
#pragma unroll
               for ( int u = 0; u < AHEAD; ++u ) nb[u] = r4[AHEAD * (g + 1) + u];
               __builtin_amdgcn_sched_group_barrier(0x100, AHEAD, 0);
            }
#pragma unroll
            for ( int u = 0; u < AHEAD; ++u )
            {
               acc = acc + b[u].x;
               acc = acc + b[u].y;
               acc = acc + b[u].z;
               acc = acc + b[u].w;
            }
            __builtin_amdgcn_sched_group_barrier(0x002, G, 0);
            if ( g + 1 < TILE / G )
            {
#pragma unroll
               for ( int u = 0; u < AHEAD; ++u ) b[u] = nb[u];
            }
         }
      }
      else
      {
         /* the partial last tile: whole groups of four, then the rest, in
          * the same index order */
         const float4* r4 = (const float4*)row;
         const int n4 = len >> 2;
         for ( int j = 0; j < n4; ++j )
         {
            float4 v = r4[j];
            acc = acc + v.x;
            acc = acc + v.y;
            acc = acc + v.z;
            acc = acc + v.w;
         }
         for ( int j = n4 << 2; j < len; ++j ) acc = acc + row[j];
      }
   }
   return acc;
}

/* chain_add for a length that is a multiple of 4 (at most the tile width). */
__device__ __forceinline__ float chain_add4(const float* tile, int stride, int nch, int len4, float acc)
{
   const int lane = lane_id();
   if ( lane < nch || lane < KINC_MIN_ACTIVE_LANES )
   {
      acc = chain_row_var(tile + (lane < nch ? lane : nch - 1) * stride, len4, acc);
   }
   return acc;
}

/* ---------------------------------------------------------- point stores */

/*
 * Per-point state of one fit.  RegStore keeps point i (lane i%64, slot i/64)
 * in VGPRs for P compile-time slots; MemStore keeps it in the wave's scratch
 * slice, [slot][lane] so that every access is one coalesced 256-byte wave
 * access, for any number of slots.  Both expose the same accessors and the
 * fits below are written once against them.
 */
template <int P, int K>
struct RegStore
{
   static constexpr int PMAX = P;
   float x0[P], x1[P];
   float lg[K][P];
   float lpx[P];
   int8_t y[P];

   __device__ __forceinline__ void init(const Lds& L, int N)
   {
      const int lane = lane_id();
#pragma unroll
      for ( int m = 0; m < P; ++m )
      {
         const int i = lane + m * WAVE;
         const float2 v = (i < N) ? L.X[i] : make_float2(0.0f, 0.0f);
         x0[m] = v.x;
         x1[m] = v.y;
      }
   }
   __device__ __forceinline__ float X0(int m) const { return x0[m]; }
   __device__ __forceinline__ float X1(int m) const { return x1[m]; }
   __device__ __forceinline__ float& LG(int k, int m) { return lg[k][m]; }
   __device__ __forceinline__ float& LPX(int m) { return lpx[m]; }
   __device__ __forceinline__ int8_t& Y(int m) { return y[m]; }
};

template <int K>
struct MemStore
{
   static constexpr int PMAX = 0;
   const float2* X;
   float* lgb;
   int8_t* yb;
   int ms;
   int n;

   __device__ __forceinline__ void init(const Lds& L, int N, int slots)
   {
      X = L.X;
      lgb = L.lgbuf;
      yb = L.ybuf;
      ms = slots;
      n = N;
   }
   __device__ __forceinline__ float X0(int m) const
   {
      const int i = lane_id() + m * WAVE;
      return (i < n) ? X[i].x : 0.0f;
   }
   __device__ __forceinline__ float X1(int m) const
   {
      const int i = lane_id() + m * WAVE;
      return (i < n) ? X[i].y : 0.0f;
   }
   __device__ __forceinline__ float& LG(int k, int m) { return lgb[((size_t)(k * ms + m) << 6) + lane_id()]; }
   __device__ __forceinline__ float& LPX(int m) { return lgb[((size_t)(K * ms + m) << 6) + lane_id()]; }
   __device__ __forceinline__ int8_t& Y(int m) { return yb[((size_t)m << 6) + lane_id()]; }
};

/* loop bound over slots: the compile-time slot count of a register store
 * (fully unrolled), the runtime count otherwise */
template <class ST>
__device__ __forceinline__ constexpr int slot_bound(int nslots)
{
   return ST::PMAX > 0 ? ST::PMAX : nslots;
}

/* ------------------------------------------------------------------ GMM */

struct FitOut
{
   float logL;
   float E;
   uint16_t iters;
   bool ok;
};

/* Component parameters live in LDS, 8 floats each:
 * mu0 mu1 a00 a01 | a10 a11 normalizer logpi (a = Sigma^-1). */
#define PRM_MU0 0
#define PRM_MU1 1
#define PRM_A00 2
#define PRM_A01 3
#define PRM_A10 4
#define PRM_A11 5
#define PRM_NRM 6
#define PRM_LPI 7

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

/* uniform value v_k of lane k, selected into lane k's register */
template <int K>
__device__ __forceinline__ float lane_pick(const float (&v)[K])
{
   const int lane = lane_id();
   float r = 0.0f;
#pragma unroll
   for ( int k = 0; k < K; ++k ) r = (lane == k) ? v[k] : r;
   return r;
}

/* E step for the slots in use: LG(k,m) = log N(x|k) - log p(x),
 * LPX(m) = log p(x) (pairwise_gmm.cpp:54-77,143-187).  Lanes past the
 * valid points compute harmless values that are never read. */
/* log N(x | component) (pairwise_gmm.cpp:54-77) from the component's
 * parameters A = (mu0, mu1, a00, a01), B = (a10, a11, normalizer, logpi). */
__device__ __forceinline__ float gmm_logpdf(float x0, float x1, const float4& A, const float4& B)
{
   float xm0 = x0 - A.x;
   float xm1 = x1 - A.y;
   float b0 = A.z * xm0 + A.w * xm1;
   float b1 = B.x * xm0 + B.y * xm1;
   float dot = xm0 * b0 + xm1 * b1;
   return B.z - 0.5f * dot;
}

template <class ST, int K>
__device__ __forceinline__ void gmm_estep(int nslots, ST& st, const float* prm, float* vmax = nullptr,
                                          bool store_lg = true)
{
   if constexpr ( ST::PMAX == 0 )
   {
      /* memory store: slot by slot, so each responsibility is written once
       * (the same arithmetic as below, in the same order per element); with
       * vmax, also each lane's maximum of every component's log
       * responsibility over its valid points; without store_lg only log p(x)
       * is stored and the next pass recomputes the responsibilities */
      if ( vmax )
      {
#pragma unroll
         for ( int k = 0; k < K; ++k ) vmax[k] = -INFINITY;
      }
      float4 A[K], B[K];
#pragma unroll
      for ( int k = 0; k < K; ++k )
      {
         A[k] = *(const float4*)(prm + 8 * k);
         B[k] = *(const float4*)(prm + 8 * k + 4);
      }
      for ( int m = 0; m < nslots; ++m )
      {
         const float x0 = st.X0(m);
         const float x1 = st.X1(m);
         float lgm[K];
#pragma unroll
         for ( int k = 0; k < K; ++k ) lgm[k] = gmm_logpdf(x0, x1, A[k], B[k]);
         float mx = -INFINITY;
#pragma unroll
         for ( int k = 0; k < K; ++k ) mx = fmaxf(mx, B[k].w + lgm[k]);
         float sum = 0.0f;
#pragma unroll
         for ( int k = 0; k < K; ++k ) sum = sum + kinc_expf_le0((B[k].w + lgm[k]) - mx);
         const float lp = mx + kinc_logf_ge1(sum);   /* sum >= 1 or NaN */
         st.LPX(m) = lp;
         const bool valid = lane_id() + m * WAVE < st.n;
#pragma unroll
         for ( int k = 0; k < K; ++k )
         {
            const float l = lgm[k] + (-lp);
            if ( store_lg ) st.LG(k, m) = l;
            if ( vmax ) vmax[k] = valid ? fmaxf(vmax[k], l) : vmax[k];
         }
      }
      return;
   }
   float logpi[K];
#pragma unroll
   for ( int k = 0; k < K; ++k )
   {
      const float4 A = *(const float4*)(prm + 8 * k);
      const float4 B = *(const float4*)(prm + 8 * k + 4);
      logpi[k] = B.w;
#pragma unroll
      for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
      {
         if ( m < nslots )
         {
            float xm0 = st.X0(m) - A.x;
            float xm1 = st.X1(m) - A.y;
            float b0 = A.z * xm0 + A.w * xm1;
            float b1 = B.x * xm0 + B.y * xm1;
            float dot = xm0 * b0 + xm1 * b1;
            st.LG(k, m) = B.z - 0.5f * dot;
         }
      }
   }
#pragma unroll
   for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
   {
      if ( m < nslots )
      {
         float lgm[K];
#pragma unroll
         for ( int k = 0; k < K; ++k ) lgm[k] = st.LG(k, m);
         /* fmaxf(mx, v) is (v > mx) ? v : mx here: mx is never NaN (a NaN v
          * leaves it alone either way), and when the two differ only in the
          * sign of a zero maximum, every use below (v - mx, mx + log(sum))
          * gives the same value */
         float mx = -INFINITY;
#pragma unroll
         for ( int k = 0; k < K; ++k )
         {
            float v = logpi[k] + lgm[k];
            mx = fmaxf(mx, v);
         }
         float sum = 0.0f;
#pragma unroll
         for ( int k = 0; k < K; ++k ) sum = sum + kinc_expf_le0((logpi[k] + lgm[k]) - mx);
         float lp = mx + kinc_logf_ge1(sum);   /* sum >= 1 or NaN */
         st.LPX(m) = lp;
#pragma unroll
         for ( int k = 0; k < K; ++k ) st.LG(k, m) = lgm[k] + (-lp);
      }
   }
}

/* GMM::fit (pairwise_gmm.cpp:367-445) with the per-fit LCG.  Writes the
 * fit's labels for i < N into L.w.  Per-component scalar work of the M step
 * runs with component k on lane k. */
template <class ST, int K>
__device__ __forceinline__ FitOut gmm_fit(const SimArgs& a, const Lds& L, int N, ST& st,
                                          bool want_entropy)
{
   const int lane = lane_id();
   constexpr int STR = KINC_TILE_STRIDE;
   const int nslots = (N + WAVE - 1) / WAVE;
   /* a register store holds X for the whole fit, so the chain tile lies over
    * X and the shared region after it (written only after the initial draws
    * have read X; X[0, N) is written back before the fit returns, as in
    * gmm_fit_joint); the memory store's X is in its scratch slice and the
    * tile in LDS */
   float* tile = (ST::PMAX > 0) ? (float*)L.X : L.tile;
   auto restore_x = [&]() {
      if constexpr ( ST::PMAX > 0 )
      {
#pragma unroll
         for ( int m = 0; m < ST::PMAX; ++m )
         {
            const int i = lane + m * WAVE;
            if ( m < nslots && i < N ) L.X[i] = make_float2(st.X0(m), st.X1(m));
         }
         wave_sync(false);
      }
   };
   float* prm = L.prm;
   FitOut out;
   out.ok = true;
   out.logL = 0.0f;
   out.E = 0.0f;
   CENSUS_T(ct);
   CENSUS_COUNT(KINC_CENSUS_FITS, 1);

   /* initialisation (pairwise_gmm.cpp:372-392): K draws of rand() % N */
   {
      uint64_t rs = 1;
      int my = 0;
#pragma unroll
      for ( int k = 0; k < K; ++k )
      {
         int i = lcg(rs) % N;
         my = (lane == k) ? i : my;
      }
      if ( lane < K )
      {
         float2 c = L.X[my];
         float a00, a01, a10, a11, nrm;
         prepare_cov(1.0f, 0.0f, 0.0f, 1.0f, a00, a01, a10, a11, nrm);
         float* p = prm + 8 * lane;
         *(float4*)p = make_float4(c.x, c.y, a00, a01);
         *(float4*)(p + 4) = make_float4(a10, a11, nrm, kinc_logf(1.0f / K));
      }
      wave_sync(L.lgbuf != nullptr);
   }
   CENSUS_LAP(KINC_CENSUS_FIT_INIT, ct);

   float prevLogL = -INFINITY;
   float curLogL = -INFINITY;
   int e_steps = 0;
   bool converged = false;
   bool failed = false;

   for ( int t = 0; t < 100; ++t )
   {
      float vmax[K];
      gmm_estep<ST, K>(nslots, st, prm, vmax, ST::PMAX != 0);
      e_steps++;
      CENSUS_LAP(KINC_CENSUS_ESTEP, ct);
      CENSUS_COUNT(KINC_CENSUS_ITERS, 1);
      CENSUS_COUNT(KINC_CENSUS_ESTEP_SLOTS, nslots);
      CENSUS_COUNT(KINC_CENSUS_EXPS, K * nslots);
      CENSUS_COUNT(KINC_CENSUS_LOGS, nslots);

      /* max over the points of each component's log responsibility
       * (pairwise_gmm.cpp:200-207); only the last slot holds lanes past N.
       * fmaxf skips a NaN as the reference's `>` does, and the sign of a
       * zero maximum does not reach l - gmax or gmax + log(sum) */
      float gmax[K];
#pragma unroll
      for ( int k = 0; k < K; ++k )
      {
         if constexpr ( ST::PMAX == 0 )
         {
            gmax[k] = wave_max(vmax[k]);   /* per-lane maxima from the E step */
            continue;
         }
         float v = -INFINITY;
#pragma unroll
         for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
         {
            if ( m < nslots - 1 )
            {
               v = fmaxf(v, st.LG(k, m));
            }
            else if ( m == nslots - 1 )
            {
               const float l = st.LG(k, m);
               v = (lane + m * WAVE < N) ? fmaxf(v, l) : v;
            }
         }
         gmax[k] = wave_max(v);
      }
      CENSUS_LAP(KINC_CENSUS_GMAX, ct);

      /* pass A: chain 0 = logL, 1+k = Gamma_k sum, 1+K+2k / 2+K+2k = mu_k sums.
       * The memory store recomputes the log responsibilities from X and the
       * parameters of this E step (identical arithmetic) instead of reading
       * them back. */
      float acc = 0.0f;
      float4 PA[ST::PMAX == 0 ? K : 1], PB[ST::PMAX == 0 ? K : 1];
      if constexpr ( ST::PMAX == 0 )
      {
#pragma unroll
         for ( int k = 0; k < K; ++k )
         {
            PA[k] = *(const float4*)(prm + 8 * k);
            PB[k] = *(const float4*)(prm + 8 * k + 4);
         }
      }
#pragma unroll
      for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
      {
         if ( m < nslots )
         {
            const float lp = st.LPX(m);
            tile[lane] = lp;
            const float xa = st.X0(m);
            const float xb = st.X1(m);
#pragma unroll
            for ( int k = 0; k < K; ++k )
            {
               float l;
               if constexpr ( ST::PMAX == 0 ) l = gmm_logpdf(xa, xb, PA[k], PB[k]) + (-lp);
               else l = st.LG(k, m);
               /* l <= gmax[k] on the valid points (gmax is their maximum) */
               tile[(1 + k) * STR + lane] = kinc_expf_le0(l - gmax[k]);
               const float gm = kinc_expf(l);
               /* the memory store keeps gamma only where the labels of a
                * 100-iteration fit will read it; pass B recomputes it */
               if ( ST::PMAX != 0 || t == 99 ) st.LG(k, m) = gm;
               tile[(1 + K + 2 * k) * STR + lane] = gm * xa;
               tile[(2 + K + 2 * k) * STR + lane] = gm * xb;
            }
            wave_sync(L.lgbuf != nullptr);
            int len = N - m * WAVE;
            len = len < TILE ? len : TILE;
            /* opaque to loop-invariant hoisting: the partial-tile values of
             * every unrolled slot would otherwise live across the EM loop */
            asm volatile("" : "+s"(len));
            CENSUS_LAP(KINC_CENSUS_A_TERMS, ct);
            acc = chain_add(tile, STR, 1 + 3 * K, len, acc);
            wave_sync(L.lgbuf != nullptr);
            CENSUS_LAP(KINC_CENSUS_A_CHAIN, ct);
            CENSUS_COUNT(KINC_CENSUS_A_SLOTS, 1);
            CENSUS_COUNT(KINC_CENSUS_EXPS, 2 * K);
         }
      }

      prevLogL = curLogL;
      curLogL = bcast(acc, 0);
      if ( fabsf(curLogL - prevLogL) < 1e-8f )
      {
         converged = true;
         break;
      }

      /* calcLogGammaK / calcLogGammaSum / performMStep (pairwise_gmm.cpp:194-316):
       * component k on lane k */
      const float gsum = __shfl(acc, 1 + lane);
      const float mxs = __shfl(acc, 1 + K + 2 * lane);
      const float mys = __shfl(acc, 2 + K + 2 * lane);
      /* gsum >= 1 (its largest term is exp(0)) or NaN; likewise sum below,
       * and v - mx <= 0 or NaN on lanes k < K */
      const float logGamma = lane_pick<K>(gmax) + kinc_logf_ge1(gsum);
      const float lpk = prm[8 * (lane < K ? lane : 0) + PRM_LPI];
      const float v = lpk + logGamma;
      float mx = -INFINITY;
#pragma unroll
      for ( int k = 0; k < K; ++k )
      {
         const float vk = bcast(v, k);
         mx = (vk > mx) ? vk : mx;
      }
      float sum = 0.0f;
      /* exp(v_k - mx) on lane k, all at once; summed in k order */
      const float ek = kinc_expf_le0(v - mx);
#pragma unroll
      for ( int k = 0; k < K; ++k ) sum = sum + bcast(ek, k);
      const float logGammaSum = mx + kinc_logf_ge1(sum);
      CENSUS_COUNT(KINC_CENSUS_EXPS, 2);
      CENSUS_COUNT(KINC_CENSUS_LOGS, 2);

      const float Gam = kinc_expf(logGamma);
      const float cm = 1.0f / Gam;
      const float mu0 = mxs * cm;
      const float mu1 = mys * cm;
      wave_sync(L.lgbuf != nullptr);
      if ( lane < K )
      {
         float* p = prm + 8 * lane;
         p[PRM_MU0] = mu0;
         p[PRM_MU1] = mu1;
         p[PRM_LPI] = lpk + (logGamma - logGammaSum);
      }
      wave_sync(L.lgbuf != nullptr);

      CENSUS_LAP(KINC_CENSUS_MSTEP, ct);

      /* pass B: chain 3k / 3k+1 / 3k+2 = Sigma_k entries 00, 01 (= 10), 11 */
      float accB = 0.0f;
      float muk0[K], muk1[K];
#pragma unroll
      for ( int k = 0; k < K; ++k )
      {
         const float2 mu = *(const float2*)(prm + 8 * k);
         muk0[k] = mu.x;
         muk1[k] = mu.y;
      }
#pragma unroll
      for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
      {
         if ( m < nslots )
         {
            const float xa = st.X0(m);
            const float xb = st.X1(m);
            float lp = 0.0f;
            if constexpr ( ST::PMAX == 0 ) lp = st.LPX(m);
#pragma unroll
            for ( int k = 0; k < K; ++k )
            {
               const float xm0 = xa - muk0[k];
               const float xm1 = xb - muk1[k];
               /* gamma as pass A computed it, from this E step's parameters */
               float gm;
               if constexpr ( ST::PMAX == 0 ) gm = kinc_expf(gmm_logpdf(xa, xb, PA[k], PB[k]) + (-lp));
               else gm = st.LG(k, m);
               tile[(3 * k) * STR + lane] = gm * (xm0 * xm0);
               tile[(3 * k + 1) * STR + lane] = gm * (xm0 * xm1);
               tile[(3 * k + 2) * STR + lane] = gm * (xm1 * xm1);
            }
            wave_sync(L.lgbuf != nullptr);
            int len = N - m * WAVE;
            len = len < TILE ? len : TILE;
            /* opaque to loop-invariant hoisting: the partial-tile values of
             * every unrolled slot would otherwise live across the EM loop */
            asm volatile("" : "+s"(len));
            CENSUS_LAP(KINC_CENSUS_B_TERMS, ct);
            accB = chain_add(tile, STR, 3 * K, len, accB);
            wave_sync(L.lgbuf != nullptr);
            CENSUS_LAP(KINC_CENSUS_B_CHAIN, ct);
            CENSUS_COUNT(KINC_CENSUS_B_SLOTS, 1);
         }
      }

      const float s00 = __shfl(accB, 3 * lane) * cm;
      const float s01 = __shfl(accB, 3 * lane + 1) * cm;
      const float s11 = __shfl(accB, 3 * lane + 2) * cm;
      float a00, a01, a10, a11, nrm = 0.0f;
      const bool ok = prepare_cov(s00, s01, s01, s11, a00, a01, a10, a11, nrm);
      failed = ballot(lane < K && !ok) != 0;
      if ( failed ) break;
      if ( lane < K )
      {
         float* p = prm + 8 * lane;
         *(float4*)(p + 0) = make_float4(mu0, mu1, a00, a01);
         p[PRM_A10] = a10;
         p[PRM_A11] = a11;
         p[PRM_NRM] = nrm;
      }
      wave_sync(L.lgbuf != nullptr);
      CENSUS_LAP(KINC_CENSUS_COV, ct);
   }

   out.iters = (uint16_t)((e_steps & 0x0fff) | (converged ? (1u << 12) : 0u) | (failed ? (1u << 13) : 0u));
   if ( failed )
   {
      wave_sync(L.lgbuf != nullptr);
      restore_x();
      out.ok = false;
      return out;
   }
   out.logL = curLogL;

   /* A converged fit labels from the log-domain responsibilities of its last
    * E step (recomputed: the registers were exponentiated by pass A); a fit
    * that ran 100 iterations labels from the exponentiated ones. */
   if ( converged )
   {
      gmm_estep<ST, K>(nslots, st, prm);
      CENSUS_COUNT(KINC_CENSUS_ESTEP_SLOTS, nslots);
      CENSUS_COUNT(KINC_CENSUS_EXPS, K * nslots);
      CENSUS_COUNT(KINC_CENSUS_LOGS, nslots);
   }
   if ( want_entropy ) CENSUS_COUNT(KINC_CENSUS_LOGS, nslots);

   /* calcLabels (pairwise_gmm.cpp:323-341); the entropy terms (:348-360)
    * go to LPX, which is free after the last E step */
#pragma unroll
   for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
   {
      const int i = lane + m * WAVE;
      if ( m < nslots )
      {
         int max_k = -1;
         float mg = -INFINITY;
#pragma unroll
         for ( int k = 0; k < K; ++k )
         {
            const float l = st.LG(k, m);
            if ( mg < l )
            {
               max_k = k;
               mg = l;
            }
         }
         if ( i < N ) L.w[i] = (int8_t)max_k;
         if ( want_entropy ) st.LPX(m) = (max_k >= 0) ? kinc_logf(mg) : NAN;
      }
   }
   wave_sync(L.lgbuf != nullptr);
   CENSUS_LAP(KINC_CENSUS_LABELS, ct);

   if ( want_entropy )
   {
      float accE = 0.0f;
#pragma unroll
      for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
      {
         if ( m < nslots )
         {
            tile[lane] = st.LPX(m);
            wave_sync(L.lgbuf != nullptr);
            int len = N - m * WAVE;
            len = len < TILE ? len : TILE;
            /* opaque to loop-invariant hoisting: the partial-tile values of
             * every unrolled slot would otherwise live across the EM loop */
            asm volatile("" : "+s"(len));
            accE = chain_add(tile, STR, 1, len, accE);
            wave_sync(L.lgbuf != nullptr);
         }
      }
      out.E = bcast(accE, 0);
      CENSUS_LAP(KINC_CENSUS_ENTROPY, ct);
   }
   restore_x();
   return out;
}

/* ------------------------------------------------------ GMM, two fits */

/*
 * Two GMM fits of one pair (KA and KB components, KA + KB <= 9) run in
 * lockstep: the same EM arithmetic per fit as gmm_fit, in the same order per
 * value, but one chain tile holds both fits' rows, so each pass's sequential
 * chain sums (64 dependent adds per slot, the latency-bound part of an
 * iteration) and the per-component M-step and covariance work serve both
 * fits at once.  A fit that converges or fails stops (its E step and terms
 * are skipped, its parameters frozen) while the other runs on.  Component k
 * of the first fit is parameter set k; component k of the second is set
 * KA + k.  Labels of the first fit go to L.w, of the second to wb.
 *
 * The chain tile (17 pass-A rows for the fits with 2 and 3 clusters, 29 for
 * those with 4 and 5) lies over the compact X and the shared region after
 * it, which the host places next to each other: X is in registers for the
 * whole fit and the tile is only written after the initial draws have read
 * X, and X[0, N) is written back from the registers before the fit returns
 * (X[N, S) is rewritten by every later reader that needs it, fill_stale).
 * wb and the entropy chain's row lie in the shared region past X, which the
 * tile no longer uses once the EM loop has ended.
 */
template <int P, int KT>
struct JointStore
{
   static constexpr int PMAX = P;
   /* With nine components the log p(x) values do not fit the register
    * budget beside the responsibilities and X: they live in the wave's
    * private memory (the address escapes below, so the compiler keeps the
    * array there), written by the E step and read by pass A one slot ahead
    * of their use, so the load latency hides behind a slot's work.  Left to
    * the register allocator they were spilled and reloaded right before the
    * use, each reload a full memory round trip. */
   static constexpr bool LPX_MEM = KT > 5;
   float x0[P], x1[P];
   float lg[KT][P];
   float lpx[2][P];      /* fewer than nine components */
   float* lpm;           /* nine: [2][P] in the fit's private array */

   __device__ __forceinline__ void init(const Lds& L, int N)
   {
      const int lane = lane_id();
#pragma unroll
      for ( int m = 0; m < P; ++m )
      {
         const int i = lane + m * WAVE;
         const float2 v = (i < N) ? L.X[i] : make_float2(0.0f, 0.0f);
         x0[m] = v.x;
         x1[m] = v.y;
      }
   }
};

/* one fit's view of a JointStore: components OFF.., log p(x) set F */
template <int P, int KT, int OFF, int F>
struct JointView
{
   static constexpr int PMAX = P;
   JointStore<P, KT>& js;
   __device__ __forceinline__ float X0(int m) const { return js.x0[m]; }
   __device__ __forceinline__ float X1(int m) const { return js.x1[m]; }
   __device__ __forceinline__ float& LG(int k, int m) { return js.lg[OFF + k][m]; }
   __device__ __forceinline__ float& LPX(int m)
   {
      if constexpr ( JointStore<P, KT>::LPX_MEM ) return js.lpm[F * P + m];
      else return js.lpx[F][m];
   }
};

/* calcLabels and the entropy terms of one finished fit (as in gmm_fit) */
template <int P, int KT, int OFF, int F, int K>
__device__ __forceinline__ float joint_labels(const Lds& L, int N, JointView<P, KT, OFF, F> v,
                                              const float* prm, bool converged, bool want_entropy,
                                              int8_t* w)
{
   const int lane = lane_id();
   constexpr int STR = KINC_TILE_STRIDE;
   const int nslots = (N + WAVE - 1) / WAVE;
   if ( converged )
   {
      gmm_estep<JointView<P, KT, OFF, F>, K>(nslots, v, prm + 8 * OFF);
      CENSUS_COUNT(KINC_CENSUS_ESTEP_SLOTS, nslots);
      CENSUS_COUNT(KINC_CENSUS_EXPS, K * nslots);
      CENSUS_COUNT(KINC_CENSUS_LOGS, nslots);
   }
   if ( want_entropy ) CENSUS_COUNT(KINC_CENSUS_LOGS, nslots);
#pragma unroll
   for ( int m = 0; m < P; ++m )
   {
      const int i = lane + m * WAVE;
      if ( m < nslots )
      {
         int max_k = -1;
         float mg = -INFINITY;
#pragma unroll
         for ( int k = 0; k < K; ++k )
         {
            const float l = v.LG(k, m);
            if ( mg < l )
            {
               max_k = k;
               mg = l;
            }
         }
         if ( i < N ) w[i] = (int8_t)max_k;
         if ( want_entropy ) v.LPX(m) = (max_k >= 0) ? kinc_logf(mg) : NAN;
      }
   }
   wave_sync(false);
   float accE = 0.0f;
   if ( want_entropy )
   {
      float* tile = L.tile;
#pragma unroll
      for ( int m = 0; m < P; ++m )
      {
         if ( m < nslots )
         {
            tile[lane] = v.LPX(m);
            wave_sync(false);
            int len = N - m * WAVE;
            len = len < TILE ? len : TILE;
            asm volatile("" : "+s"(len));
            accE = chain_add(tile, STR, 1, len, accE);
            wave_sync(false);
         }
      }
   }
   return bcast(accE, 0);
}

/* pass-A terms of one fit into tile rows R.. (lp = LPX(m), K gamma-sum
 * terms, 2K mean terms) and the exponentiated responsibilities into LG */
template <class V, int K>
__device__ __forceinline__ void joint_terms_a(V& v, int m, float* tile, int R, const float* gmax,
                                              float lp)
{
   constexpr int STR = KINC_TILE_STRIDE;
   const int lane = lane_id();
   tile[R * STR + lane] = lp;
   const float xa = v.X0(m);
   const float xb = v.X1(m);
#pragma unroll
   for ( int k = 0; k < K; ++k )
   {
      const float l = v.LG(k, m);
      tile[(R + 1 + k) * STR + lane] = kinc_expf_le0(l - gmax[k]);
      const float gm = kinc_expf(l);
      v.LG(k, m) = gm;
      tile[(R + 1 + K + 2 * k) * STR + lane] = gm * xa;
      tile[(R + 2 + K + 2 * k) * STR + lane] = gm * xb;
   }
}

/* pass-B terms of one fit (components OFF..OFF+K-1 of the joint tile) */
template <class V, int K, int OFF>
__device__ __forceinline__ void joint_terms_b(V& v, int m, float* tile, const float* muk0, const float* muk1)
{
   constexpr int STR = KINC_TILE_STRIDE;
   const int lane = lane_id();
   const float xa = v.X0(m);
   const float xb = v.X1(m);
#pragma unroll
   for ( int k = 0; k < K; ++k )
   {
      const float xm0 = xa - muk0[OFF + k];
      const float xm1 = xb - muk1[OFF + k];
      const float gm = v.LG(k, m);
      tile[(3 * (OFF + k)) * STR + lane] = gm * (xm0 * xm0);
      tile[(3 * (OFF + k) + 1) * STR + lane] = gm * (xm0 * xm1);
      tile[(3 * (OFF + k) + 2) * STR + lane] = gm * (xm1 * xm1);
   }
}

/* per-component maximum of one fit's log responsibilities (as gmm_fit) */
template <class V, int K>
__device__ __forceinline__ void joint_gmax(V& v, int N, int nslots, float* gmax)
{
   const int lane = lane_id();
#pragma unroll
   for ( int k = 0; k < K; ++k )
   {
      float x = -INFINITY;
#pragma unroll
      for ( int m = 0; m < V::PMAX; ++m )
      {
         if ( m < nslots - 1 )
         {
            x = fmaxf(x, v.LG(k, m));
         }
         else if ( m == nslots - 1 )
         {
            const float l = v.LG(k, m);
            x = (lane + m * WAVE < N) ? fmaxf(x, l) : x;
         }
      }
      gmax[k] = wave_max(x);
   }
}

template <int P, int KA, int KB>
__device__ __forceinline__ void gmm_fit_joint(const Lds& L, int N, bool want_entropy, FitOut& fa,
                                              FitOut& fb, int8_t* wb)
{
   constexpr int KT = KA + KB;
   static_assert(KT <= 9, "two fits share at most nine components");
   constexpr int STR = KINC_TILE_STRIDE;
   constexpr int RB = 1 + 3 * KA;          /* first pass-A row of the second fit */
   constexpr int NCA = RB + 1 + 3 * KB;    /* pass-A chains */
   const int lane = lane_id();
   const int nslots = (N + WAVE - 1) / WAVE;
   /* the tile over X and the shared region after it */
   float* tile = (float*)L.X;
   float* prm = L.prm;
   CENSUS_T(ct);
   CENSUS_COUNT(KINC_CENSUS_FITS, 2);

   JointStore<P, KT> js;
   /* the private array of log p(x) (nine components): its address escapes
    * here, so it stays in memory and apart from the register-held fields */
   float lpx_mem[2 * P];
   js.lpm = lpx_mem;
   if constexpr ( JointStore<P, KT>::LPX_MEM ) asm volatile("" : : "v"(lpx_mem) : "memory");
   js.init(L, N);
   JointView<P, KT, 0, 0> va{js};
   JointView<P, KT, KA, 1> vb{js};

   /* initialisation (pairwise_gmm.cpp:372-392), each fit with its own LCG */
   {
      uint64_t ra = 1, rb = 1;
      int my = 0;
#pragma unroll
      for ( int k = 0; k < KA; ++k )
      {
         const int i = lcg(ra) % N;
         my = (lane == k) ? i : my;
      }
#pragma unroll
      for ( int k = 0; k < KB; ++k )
      {
         const int i = lcg(rb) % N;
         my = (lane == KA + k) ? i : my;
      }
      if ( lane < KT )
      {
         float2 c = L.X[my];
         float a00, a01, a10, a11, nrm;
         prepare_cov(1.0f, 0.0f, 0.0f, 1.0f, a00, a01, a10, a11, nrm);
         float* p = prm + 8 * lane;
         *(float4*)p = make_float4(c.x, c.y, a00, a01);
         *(float4*)(p + 4) = make_float4(a10, a11, nrm, kinc_logf(1.0f / (lane < KA ? KA : KB)));
      }
      wave_sync(false);
   }
   CENSUS_LAP(KINC_CENSUS_FIT_INIT, ct);

   float prevA = -INFINITY, curA = -INFINITY, prevB = -INFINITY, curB = -INFINITY;
   int eA = 0, eB = 0;
   bool actA = true, actB = true;
   bool convA = false, convB = false, failA = false, failB = false;

   for ( int t = 0; t < 100; ++t )
   {
      if ( actA )
      {
         gmm_estep<JointView<P, KT, 0, 0>, KA>(nslots, va, prm);
         eA++;
         CENSUS_COUNT(KINC_CENSUS_ESTEP_SLOTS, nslots);
         CENSUS_COUNT(KINC_CENSUS_EXPS, KA * nslots);
         CENSUS_COUNT(KINC_CENSUS_LOGS, nslots);
      }
      if ( actB )
      {
         gmm_estep<JointView<P, KT, KA, 1>, KB>(nslots, vb, prm + 8 * KA);
         eB++;
         CENSUS_COUNT(KINC_CENSUS_ESTEP_SLOTS, nslots);
         CENSUS_COUNT(KINC_CENSUS_EXPS, KB * nslots);
         CENSUS_COUNT(KINC_CENSUS_LOGS, nslots);
      }
      CENSUS_LAP(KINC_CENSUS_ESTEP, ct);
      CENSUS_COUNT(KINC_CENSUS_ITERS, (actA ? 1 : 0) + (actB ? 1 : 0));

      /* log p(x) of the next pass-A slot, loaded one slot ahead (the first
       * before the maxima; nslots >= 1; the scheduling barriers keep each
       * load where it is issued) */
      constexpr bool LPX_MEM = JointStore<P, KT>::LPX_MEM;
      float lpnA = va.LPX(0), lpnB = vb.LPX(0);
      if constexpr ( LPX_MEM ) __builtin_amdgcn_sched_barrier(0);

      float gmax[KT];
#pragma unroll
      for ( int k = 0; k < KT; ++k ) gmax[k] = 0.0f;
      if ( actA ) joint_gmax<JointView<P, KT, 0, 0>, KA>(va, N, nslots, gmax);
      if ( actB ) joint_gmax<JointView<P, KT, KA, 1>, KB>(vb, N, nslots, gmax + KA);
      CENSUS_LAP(KINC_CENSUS_GMAX, ct);

      /* pass A: rows 0..RB-1 the first fit, RB.. the second (rows of a fit
       * that has stopped hold stale values whose sums are not read) */
      float acc = 0.0f;
#pragma unroll
      for ( int m = 0; m < P; ++m )
      {
         if ( m < nslots )
         {
            const float lpA = lpnA, lpB = lpnB;
            if ( m + 1 < nslots )   /* only slots the E step wrote */
            {
               lpnA = va.LPX(m + 1);
               lpnB = vb.LPX(m + 1);
               if constexpr ( LPX_MEM ) __builtin_amdgcn_sched_barrier(0);
            }
            if ( actA ) joint_terms_a<JointView<P, KT, 0, 0>, KA>(va, m, tile, 0, gmax, lpA);
            if ( actB ) joint_terms_a<JointView<P, KT, KA, 1>, KB>(vb, m, tile, RB, gmax + KA, lpB);
            CENSUS_COUNT(KINC_CENSUS_A_SLOTS, (actA ? 1 : 0) + (actB ? 1 : 0));
            CENSUS_COUNT(KINC_CENSUS_EXPS, (actA ? 2 * KA : 0) + (actB ? 2 * KB : 0));
            wave_sync(false);
            int len = N - m * WAVE;
            len = len < TILE ? len : TILE;
            /* opaque to loop-invariant hoisting, as in gmm_fit */
            asm volatile("" : "+s"(len));
            CENSUS_LAP(KINC_CENSUS_A_TERMS, ct);
            acc = chain_add<(NCA > 17 ? 1 : 2)>(tile, STR, NCA, len, acc);
            wave_sync(false);
            CENSUS_LAP(KINC_CENSUS_A_CHAIN, ct);
         }
      }

      if ( actA )
      {
         prevA = curA;
         curA = bcast(acc, 0);
         if ( fabsf(curA - prevA) < 1e-8f )
         {
            convA = true;
            actA = false;
         }
      }
      if ( actB )
      {
         prevB = curB;
         curB = bcast(acc, RB);
         if ( fabsf(curB - prevB) < 1e-8f )
         {
            convB = true;
            actB = false;
         }
      }
      if ( !actA && !actB ) break;

      /* M step (pairwise_gmm.cpp:194-316): component k of the first fit on
       * lane k, of the second on lane KA + k; each fit's log-sum-exp over its
       * own lanes, in component order */
      const bool inA = lane < KA;
      const int rg = inA ? 1 + lane : RB + 1 + (lane - KA);
      const int rm = inA ? 1 + KA + 2 * lane : RB + 1 + KB + 2 * (lane - KA);
      const float gsum = __shfl(acc, rg);
      const float mxs = __shfl(acc, rm);
      const float mys = __shfl(acc, rm + 1);
      const float logGamma = lane_pick<KT>(gmax) + kinc_logf_ge1(gsum);
      const float lpk = prm[8 * (lane < KT ? lane : 0) + PRM_LPI];
      const float v = lpk + logGamma;
      float mxA = -INFINITY, mxB = -INFINITY;
#pragma unroll
      for ( int k = 0; k < KA; ++k )
      {
         const float vk = bcast(v, k);
         mxA = (vk > mxA) ? vk : mxA;
      }
#pragma unroll
      for ( int k = 0; k < KB; ++k )
      {
         const float vk = bcast(v, KA + k);
         mxB = (vk > mxB) ? vk : mxB;
      }
      const float mx = inA ? mxA : mxB;
      const float ek = kinc_expf_le0(v - mx);
      float sumA = 0.0f, sumB = 0.0f;
#pragma unroll
      for ( int k = 0; k < KA; ++k ) sumA = sumA + bcast(ek, k);
#pragma unroll
      for ( int k = 0; k < KB; ++k ) sumB = sumB + bcast(ek, KA + k);
      const float logGammaSum = mx + kinc_logf_ge1(inA ? sumA : sumB);
      CENSUS_COUNT(KINC_CENSUS_EXPS, 2);
      CENSUS_COUNT(KINC_CENSUS_LOGS, 2);

      const float Gam = kinc_expf(logGamma);
      const float cm = 1.0f / Gam;
      const float mu0 = mxs * cm;
      const float mu1 = mys * cm;
      const bool mine = lane < KT && (inA ? actA : actB);
      wave_sync(false);
      if ( mine )
      {
         float* p = prm + 8 * lane;
         p[PRM_MU0] = mu0;
         p[PRM_MU1] = mu1;
         p[PRM_LPI] = lpk + (logGamma - logGammaSum);
      }
      wave_sync(false);
      CENSUS_LAP(KINC_CENSUS_MSTEP, ct);

      /* pass B: rows 3c, 3c+1, 3c+2 = Sigma entries of parameter set c */
      float accB = 0.0f;
      float muk0[KT], muk1[KT];
#pragma unroll
      for ( int k = 0; k < KT; ++k )
      {
         const float2 mu = *(const float2*)(prm + 8 * k);
         muk0[k] = mu.x;
         muk1[k] = mu.y;
      }
#pragma unroll
      for ( int m = 0; m < P; ++m )
      {
         if ( m < nslots )
         {
            if ( actA ) joint_terms_b<JointView<P, KT, 0, 0>, KA, 0>(va, m, tile, muk0, muk1);
            if ( actB ) joint_terms_b<JointView<P, KT, KA, 1>, KB, KA>(vb, m, tile, muk0, muk1);
            CENSUS_COUNT(KINC_CENSUS_B_SLOTS, (actA ? 1 : 0) + (actB ? 1 : 0));
            wave_sync(false);
            int len = N - m * WAVE;
            len = len < TILE ? len : TILE;
            asm volatile("" : "+s"(len));
            CENSUS_LAP(KINC_CENSUS_B_TERMS, ct);
            accB = chain_add<(NCA > 17 ? 1 : 2)>(tile, STR, 3 * KT, len, accB);
            wave_sync(false);
            CENSUS_LAP(KINC_CENSUS_B_CHAIN, ct);
         }
      }

      const float s00 = __shfl(accB, 3 * lane) * cm;
      const float s01 = __shfl(accB, 3 * lane + 1) * cm;
      const float s11 = __shfl(accB, 3 * lane + 2) * cm;
      float a00, a01, a10, a11, nrm = 0.0f;
      const bool ok = prepare_cov(s00, s01, s01, s11, a00, a01, a10, a11, nrm);
      if ( actA && ballot(inA && !ok) != 0 )
      {
         failA = true;
         actA = false;
      }
      if ( actB && ballot(!inA && lane < KT && !ok) != 0 )
      {
         failB = true;
         actB = false;
      }
      if ( lane < KT && (inA ? actA : actB) )
      {
         float* p = prm + 8 * lane;
         *(float4*)(p + 0) = make_float4(mu0, mu1, a00, a01);
         p[PRM_A10] = a10;
         p[PRM_A11] = a11;
         p[PRM_NRM] = nrm;
      }
      wave_sync(false);
      CENSUS_LAP(KINC_CENSUS_COV, ct);
      if ( !actA && !actB ) break;
   }

   fa.iters = (uint16_t)((eA & 0x0fff) | (convA ? (1u << 12) : 0u) | (failA ? (1u << 13) : 0u));
   fb.iters = (uint16_t)((eB & 0x0fff) | (convB ? (1u << 12) : 0u) | (failB ? (1u << 13) : 0u));
   fa.ok = !failA;
   fb.ok = !failB;
   fa.logL = failA ? 0.0f : curA;
   fb.logL = failB ? 0.0f : curB;
   fa.E = 0.0f;
   fb.E = 0.0f;
   if ( !failA ) fa.E = joint_labels<P, KT, 0, 0, KA>(L, N, va, prm, convA, want_entropy, L.w);
   if ( !failB ) fb.E = joint_labels<P, KT, KA, 1, KB>(L, N, vb, prm, convB, want_entropy, wb);
   /* the tile covered X: write the pair's points back */
#pragma unroll
   for ( int m = 0; m < P; ++m )
   {
      const int i = lane + m * WAVE;
      if ( m < nslots && i < N ) L.X[i] = make_float2(js.x0[m], js.x1[m]);
   }
   wave_sync(false);
   CENSUS_LAP(KINC_CENSUS_LABELS, ct);
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

/* squared distance exactly as diff_norm forms it before the square root */
__device__ __forceinline__ float diff_sq(float ax, float ay, float bx, float by)
{
   float dist = 0.0f;
   dist = dist + (ax - bx) * (ax - bx);
   dist = dist + (ay - by) * (ay - by);
   return dist;
}

/*
 * Nearest mean of one point in the reference's terms (pairwise_kmeans.cpp:
 * 37-54): the first k whose sqrt(q_k) is smallest, q_k the squared distance,
 * where a NaN distance at k = 0 wins and later NaNs never do.  sqrt is
 * monotone and correctly rounded, so the first k with the smallest q_k (same
 * comparisons on q) is that k unless an earlier k has a q_k so close above
 * the minimum that both square roots round to the same float, which needs
 * q_k < q_min * (1 + 2^-20).  Such a near tie is flagged (some earlier q_k <=
 * q_min * (1 + 2^-20), a safe margin also in the subnormal range) and the
 * caller then decides with the square roots.  The smallest q_k before the
 * chosen k is the running minimum at the moment the choice last moved, so
 * the flag needs one comparison, not one per earlier k.
 */
template <int K>
__device__ __forceinline__ int nearest_mean_q(const float (&q)[K], bool& near_tie)
{
   int kq = 0;
   float qmin = q[0];
   float before = NAN;   /* smallest q_k with k < kq; NaN (no flag) while kq = 0 */
#pragma unroll
   for ( int k = 1; k < K; ++k )
   {
      if ( q[k] < qmin )
      {
         before = qmin;
         kq = k;
         qmin = q[k];
      }
   }
   near_tie = before <= qmin * (1.0f + 0x1p-20f);
   return kq;
}

/* KMeans::fit (pairwise_kmeans.cpp:12-127) with the per-fit LCG.  Means live
 * in LDS (2 floats per cluster); the best init's labels go to L.w. */
template <class ST, int K>
__device__ __forceinline__ FitOut kmeans_fit(const SimArgs& a, const Lds& L, int N, ST& st)
{
   const int lane = lane_id();
   constexpr int STR = KINC_TILE_STRIDE;
   const int nslots = (N + WAVE - 1) / WAVE;
   float* tile = L.tile;
   float* means = L.prm;
   FitOut out;
   out.ok = true;
   out.E = 0.0f;
   uint64_t rs = 1;
   float bestLogL = -INFINITY;
   int iters = 0;

   for ( int init = 0; init < 10; ++init )
   {
      int my = 0;
#pragma unroll
      for ( int k = 0; k < K; ++k )
      {
         int i = lcg(rs) % N;
         my = (lane == k) ? i : my;
      }
      wave_sync(L.lgbuf != nullptr);
      if ( lane < K ) *(float2*)(means + 2 * lane) = L.X[my];
      wave_sync(L.lgbuf != nullptr);
#pragma unroll
      for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
      {
         if ( m < nslots ) st.Y(m) = 0;
      }

      for ( int t = 0; t < 300; ++t )
      {
         iters++;
         float m0[K], m1[K];
#pragma unroll
         for ( int k = 0; k < K; ++k )
         {
            const float2 mk = *(const float2*)(means + 2 * k);
            m0[k] = mk.x;
            m1[k] = mk.y;
         }
         bool diff = false;
#pragma unroll
         for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
         {
            if ( m < nslots )
            {
               const float xa = st.X0(m);
               const float xb = st.X1(m);
               float q[K];
#pragma unroll
               for ( int k = 0; k < K; ++k ) q[k] = diff_sq(xa, xb, m0[k], m1[k]);
               bool near_tie;
               int min_k = nearest_mean_q<K>(q, near_tie);
               if ( ballot(near_tie) != 0 )
               {
                  /* the reference's comparison of square roots */
                  min_k = -1;
                  float min_d = 0.0f;
#pragma unroll
                  for ( int k = 0; k < K; ++k )
                  {
                     const float d = sqrtf(q[k]);
                     if ( min_k == -1 || d < min_d )
                     {
                        min_k = k;
                        min_d = d;
                     }
                  }
               }
               /* the new label goes straight over the old one: if none
                * changed, the loop stops with the labels it had (the
                * reference's y == y_next); else they are the reference's
                * swapped-in y_next */
               const int8_t old = st.Y(m);
               st.Y(m) = (int8_t)min_k;
               diff = diff || ((lane + m * WAVE < N) && (int8_t)min_k != old);
            }
         }
         if ( ballot(diff) == 0 ) break;

         /* means: chain 2k / 2k+1 over the members of cluster k in index
          * order.  Each slot's members of k are packed to the front of rows
          * 2k and 2k+1 and the rows are zero-filled to a common length (a
          * multiple of 4), so a chain adds only its members plus zeros; the
          * running sums start at +0 and never become -0, so the zeros leave
          * them unchanged. */
         float acc = 0.0f;
         int nkc[K];
#pragma unroll
         for ( int k = 0; k < K; ++k ) nkc[k] = 0;
#pragma unroll
         for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
         {
            if ( m < nslots )
            {
               const float xa = st.X0(m);
               const float xb = st.X1(m);
               const int ym = (lane + m * WAVE < N) ? (int)st.Y(m) : -1;
               int ck[K];
               int rank = 0, maxc = 0;
#pragma unroll
               for ( int k = 0; k < K; ++k )
               {
                  const uint64_t bm = ballot(ym == k);
                  ck[k] = popc(bm);
                  nkc[k] += ck[k];
                  maxc = ck[k] > maxc ? ck[k] : maxc;
                  rank = (ym == k) ? popc(bm & lanes_below()) : rank;
               }
               const int len = (maxc + 3) & ~3;
               /* every row zeroed by all lanes, then the members written
                * over the front (the wave's LDS stores land in order), so
                * each row holds its members and zeros up to len */
#pragma unroll
               for ( int k = 0; k < K; ++k )
               {
                  tile[(2 * k) * STR + lane] = 0.0f;
                  tile[(2 * k + 1) * STR + lane] = 0.0f;
               }
               if ( ym >= 0 )
               {
                  tile[(2 * ym) * STR + rank] = xa;
                  tile[(2 * ym + 1) * STR + rank] = xb;
               }
               wave_sync(L.lgbuf != nullptr);
               acc = chain_add4(tile, STR, 2 * K, len, acc);
               wave_sync(L.lgbuf != nullptr);
            }
         }
         int nk = 0;
#pragma unroll
         for ( int k = 0; k < K; ++k ) nk = (lane == k) ? nkc[k] : nk;
         const float sx = __shfl(acc, 2 * lane);
         const float sy = __shfl(acc, 2 * lane + 1);
         const float c = 1.0f / (float)nk;
         if ( lane < K ) *(float2*)(means + 2 * lane) = make_float2(sx * c, sy * c);
         wave_sync(L.lgbuf != nullptr);
      }

      /* computeLogLikelihood (pairwise_kmeans.cpp:106-127): one chain over
       * the members in (k, i) order.  The terms are staged in order in the
       * shared region (the mean chains' tile is free here) and added
       * whenever the next slot's members would not fit a.ll_cap values, so
       * the region need not hold all N of them; the adds run in the same
       * order either way. */
      float* seq = L.u;
      const int cap = a.ll_cap;
      int pos = 0;
      float Ssum = 0.0f;
      auto add_staged = [&]() {
         wave_sync(L.lgbuf != nullptr);
         if ( lane < KINC_MIN_ACTIVE_LANES )   /* identical sums; lane 0's is used */
         {
            /* four values per LDS read, added in index order */
            const float4* s4 = (const float4*)seq;
            const int n4 = pos >> 2;
            for ( int j = 0; j < n4; ++j )
            {
               const float4 v = s4[j];
               Ssum = Ssum + v.x;
               Ssum = Ssum + v.y;
               Ssum = Ssum + v.z;
               Ssum = Ssum + v.w;
            }
            for ( int j = n4 << 2; j < pos; ++j ) Ssum = Ssum + seq[j];
         }
         wave_sync(L.lgbuf != nullptr);
         pos = 0;
      };
#pragma unroll
      for ( int k = 0; k < K; ++k )
      {
         const float2 mk = *(const float2*)(means + 2 * k);
#pragma unroll
         for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
         {
            if ( m < nslots )
            {
               const bool mem = (lane + m * WAVE < N) && st.Y(m) == k;
               const uint64_t bm = ballot(mem);
               const int c = popc(bm);
               if ( pos + c > cap ) add_staged();   /* cap >= 64 >= c */
               if ( mem )
               {
                  const float d = diff_norm(st.X0(m), st.X1(m), mk.x, mk.y);
                  seq[pos + popc(bm & lanes_below())] = d * d;
               }
               pos += c;
            }
         }
      }
      add_staged();
      const float logL = -bcast(Ssum, 0);

      if ( bestLogL < logL )
      {
         bestLogL = logL;
#pragma unroll
         for ( int m = 0; m < slot_bound<ST>(nslots); ++m )
         {
            const int i = lane + m * WAVE;
            if ( m < nslots && i < N ) L.w[i] = st.Y(m);
         }
         wave_sync(L.lgbuf != nullptr);
      }
   }
   out.logL = bestLogL;
   out.iters = (uint16_t)(iters & 0x0fff);
   return out;
}

/* ------------------------------------------------- stale X (Pearson) */

/* (x, y) of block-relative pair q */
__device__ __forceinline__ void pair_genes(const SimArgs& a, int q, int& x, int& y)
{
   if ( a.pairs )
   {
      /* a pair outside x > y >= 0, x < G reads gene 0 for both rows so that
       * no access leaves the matrix; its results are meaningless, and the
       * host entry point rejects such pairs before launching */
      const int2 v = a.pairs[q];
      const bool ok = 0 <= v.y && v.y < v.x && v.x < a.G;
      x = ok ? v.x : 0;
      y = ok ? v.y : 0;
   }
   else
   {
      pair_xy(a.block_start + q, x, y);
   }
}

/* Last block-relative pair d with lo <= d < qp and blockN[d] > v, or -1.
 * Chunks of 64 pairs lying wholly in [lo, qp) are skipped by their maxima. */
__device__ __forceinline__ int prev_greater(const int32_t* blockN, const int32_t* chunkMax,
                                            int qp, int v, int lo)
{
   const int lane = lane_id();
   if ( qp <= lo ) return -1;
   const int c = (qp - 1) >> 6;
   {
      int idx = c * WAVE + lane;
      bool ok = idx >= lo && idx < qp && blockN[idx] > v;
      uint64_t m = ballot(ok);
      if ( m ) return c * WAVE + 63 - __clzll(m);
   }
   const int cfull = (lo + WAVE - 1) >> 6;   /* first chunk wholly >= lo */
   for ( int cb = c - 1; cb >= cfull; cb -= WAVE )
   {
      int cc = cb - lane;
      bool ok = cc >= cfull && chunkMax[cc] > v;
      uint64_t m = ballot(ok);
      if ( m )
      {
         int ch = cb - (__ffsll((unsigned long long)m) - 1);
         int idx = ch * WAVE + lane;
         uint64_t m2 = ballot(blockN[idx] > v);
         return ch * WAVE + 63 - __clzll(m2);
      }
   }
   const int cpart = lo >> 6;                /* chunk holding lo, if partial */
   if ( cpart < cfull && cpart < c )
   {
      int idx = cpart * WAVE + lane;
      bool ok = idx >= lo && blockN[idx] > v;
      uint64_t m = ballot(ok);
      if ( m ) return cpart * WAVE + 63 - __clzll(m);
   }
   return -1;
}

__device__ __forceinline__ int find_prev_greater(const SimArgs& a, int qp, int v, int lo)
{
   return prev_greater(a.blockN, a.chunkMax, qp, v, lo);
}

/*
 * The reference's X workspace is allocated once per block and each pair only
 * overwrites its first N entries (similarity_serial.cpp:42,132), while
 * Pearson reads X at full sample indices (pairwise_pearson.cpp:26-31).  Entry
 * i >= N of pair q therefore holds entry i of the latest earlier pair of the
 * block with more than i valid samples, or zero.  Rebuild that view in L.X.
 */
__device__ __forceinline__ void fill_stale(const SimArgs& a, const Lds& L, int q, int N)
{
   const int S = a.S;
   const int lane = lane_id();
   if ( a.xbuf != nullptr )
   {
      /* the loop below with its searches precomputed: entries [v, N_d)
       * come from the compact X of d (stored by the count kernel), and the
       * search that follows, from d with threshold N_d, is prev[d] */
      const float2* xb = (const float2*)a.xbuf;
      int v = N;
      int d = a.prev[q];
      while ( v < S )
      {
         if ( d < 0 )
         {
            for ( int i = v + lane; i < S; i += WAVE ) L.X[i] = make_float2(0.0f, 0.0f);
            break;
         }
         const int hi = a.blockN[d];
         const int dn = a.prev[d];
         const float2* src = xb + (size_t)d * (size_t)S;
         for ( int i = v + lane; i < hi; i += WAVE ) L.X[i] = src[i];
         v = hi;
         d = dn;
      }
      wave_sync(L.lgbuf != nullptr);
      return;
   }
   /* clamped into [0, q] like the link kernel's, so no read leaves the block */
   const int lo = a.seg_lo ? min(max(a.seg_lo[q], 0), q) : 0;
   int v = N;
   int qp = q;
   while ( v < S )
   {
      int d = find_prev_greater(a, qp, v, lo);
      if ( d < 0 )
      {
         for ( int i = v + lane; i < S; i += WAVE ) L.X[i] = make_float2(0.0f, 0.0f);
         break;
      }
      const int hi = a.blockN[d];
      int gx, gy;
      pair_genes(a, d, gx, gy);
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
   wave_sync(L.lgbuf != nullptr);
}

/* -------------------------------------------------------------- Pearson */

/* pairwise_pearson.cpp:12-52 for clusters 0..K-1 at once: chain 5k+t holds
 * sum x, sum y, sum x^2, sum y^2, sum xy of cluster k.  With pearson_fp64
 * (an option the reference lacks) the same sums run in double over exact
 * products and the quotient is evaluated in double, then rounded to float. */
template <int KMAX>
__device__ __forceinline__ void pearson_all(const SimArgs& a, const Lds& L, int K, float* corr)
{
   const int S = a.S;
   const int lane = lane_id();
   const int nch = 5 * K;
   float acc = 0.0f;
   double dacc = 0.0;   /* pearson_fp64 mode */
   /* lanes nch..15 (cluster index >= K, which no label takes) add nothing */
   if ( lane < nch || lane < KINC_MIN_ACTIVE_LANES )
   {
      const int k = lane / 5;
      const int t = lane - 5 * k;
      /* four samples per pass: one 4-byte read of their labels and two
       * 16-byte reads of their points (both arrays start 16-byte aligned),
       * added in index order */
      const int S4 = S & ~3;
      if ( !a.pearson_fp64 )
      {
         auto add = [&](int8_t l, float x, float y) {
            float term = (t == 0) ? x
                       : (t == 1) ? y
                       : (t == 2) ? x * x
                       : (t == 3) ? y * y
                       : x * y;
            acc = (l == k) ? acc + term : acc;
         };
         for ( int i = 0; i < S4; i += 4 )
         {
            const uint32_t l4 = *(const uint32_t*)(L.lab + i);
            const float4 va = *(const float4*)(L.X + i);
            const float4 vb = *(const float4*)(L.X + i + 2);
            add((int8_t)l4, va.x, va.y);
            add((int8_t)(l4 >> 8), va.z, va.w);
            add((int8_t)(l4 >> 16), vb.x, vb.y);
            add((int8_t)(l4 >> 24), vb.z, vb.w);
         }
         for ( int i = S4; i < S; ++i ) add(L.lab[i], L.X[i].x, L.X[i].y);
      }
      else
      {
         /* term = f1 * f2, exact in double (two 24-bit significands) */
         auto add = [&](int8_t l, float x, float y) {
            const float f1 = (t == 1 || t == 3) ? y : x;
            const float f2 = (t < 2) ? 1.0f : (t == 2) ? x : y;
            const double term = (double)f1 * (double)f2;
            dacc = (l == k) ? dacc + term : dacc;
         };
         for ( int i = 0; i < S4; i += 4 )
         {
            const uint32_t l4 = *(const uint32_t*)(L.lab + i);
            const float4 va = *(const float4*)(L.X + i);
            const float4 vb = *(const float4*)(L.X + i + 2);
            add((int8_t)l4, va.x, va.y);
            add((int8_t)(l4 >> 8), va.z, va.w);
            add((int8_t)(l4 >> 16), vb.x, vb.y);
            add((int8_t)(l4 >> 24), vb.z, vb.w);
         }
         for ( int i = S4; i < S; ++i ) add(L.lab[i], L.X[i].x, L.X[i].y);
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
         if ( n >= a.min_samples && !a.pearson_fp64 )
         {
            const float fn = (float)n;
            float sx = bcast(acc, 5 * k), sy = bcast(acc, 5 * k + 1);
            float sxx = bcast(acc, 5 * k + 2), syy = bcast(acc, 5 * k + 3);
            float sxy = bcast(acc, 5 * k + 4);
            r = (fn * sxy - sx * sy) / sqrtf((fn * sxx - sx * sx) * (fn * syy - sy * sy));
         }
         else if ( n >= a.min_samples )
         {
            const double dn = (double)n;
            const double sx = bcast_d(dacc, 5 * k), sy = bcast_d(dacc, 5 * k + 1);
            const double sxx = bcast_d(dacc, 5 * k + 2), syy = bcast_d(dacc, 5 * k + 3);
            const double sxy = bcast_d(dacc, 5 * k + 4);
            r = (float)((dn * sxy - sx * sy) / sqrt((dn * sxx - sx * sx) * (dn * syy - sy * sy)));
         }
         corr[k] = r;
      }
   }
}

/* ------------------------------------------------------------- Spearman */

/* pairwise_spearman.cpp:27-86 */
__device__ __forceinline__ float spearman_k(const SimArgs& a, const Lds& L, int k)
{
   const int S = a.S;
   const int lane = lane_id();
   /* x, y in two buffers; the rank payload of the second sort reuses the
    * x buffer, which the first sort leaves free */
   float* kx = L.u;
   float* ky = L.u + a.np2;
   float* kr = kx;
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
      }
      jb += popc(mge);
      n += popc(min);
   }
   if ( n < a.min_samples )
   {
      wave_sync(L.lgbuf != nullptr);
      return NAN;
   }
   const int n2 = next_pow2(n);
   for ( int i = n + lane; i < n2; i += WAVE )
   {
      kx[i] = INFINITY;
      ky[i] = INFINITY;
   }
   wave_sync(L.lgbuf != nullptr);
   bitonic_sort<true>(kx, ky, n2, L.lgbuf != nullptr);
   /* ranks n+1 for the extracted samples in sample order, 0 for the padding
    * (pairwise_spearman.cpp:45,57) */
   for ( int i = lane; i < n2; i += WAVE ) kr[i] = (i < n) ? (float)(i + 1) : 0.0f;
   wave_sync(L.lgbuf != nullptr);
   bitonic_sort<true>(ky, kr, n2, L.lgbuf != nullptr);

   uint32_t d = 0;
   for ( int i = lane; i < n; i += WAVE )
   {
      int tmp = (int)((float)(i + 1) - kr[i]);
      d += (uint32_t)(tmp * tmp);
   }
   d = wave_sum_u32(d);
   wave_sync(L.lgbuf != nullptr);
   const uint32_t un = (uint32_t)n;
   const int32_t denom = (int32_t)(un * (un * un - 1u));
   return (float)(1.0 - 6.0 * (double)(int32_t)d / (double)denom);
}

#include "sim_pair_dyn.h"

/* -------------------------------------------------------------- driver */

/* One clustering fit with K components; P > 0 selects the register store
 * with P slots, P = 0 the memory store. */
template <int P, int K>
__device__ __forceinline__ FitOut fit_one(const SimArgs& a, const Lds& L, int N)
{
   if constexpr ( P > 0 )
   {
      RegStore<P, K> st;
      st.init(L, N);
      return (a.clus == 1) ? gmm_fit<RegStore<P, K>, K>(a, L, N, st, a.crit == 1)
                           : kmeans_fit<RegStore<P, K>, K>(a, L, N, st);
   }
   else
   {
      MemStore<K> st;
      st.init(L, N, a.mem_slots);
      return (a.clus == 1) ? gmm_fit<MemStore<K>, K>(a, L, N, st, a.crit == 1)
                           : kmeans_fit<MemStore<K>, K>(a, L, N, st);
   }
}

template <int P, int KMAX>
__device__ __forceinline__ FitOut fit_dispatch(const SimArgs& a, const Lds& L, int N, int Kc)
{
   switch ( Kc )
   {
   case 1: return fit_one<P, 1>(a, L, N);
   case 2: return fit_one<P, 2>(a, L, N);
   case 3: return fit_one<P, 3>(a, L, N);
   case 4: return fit_one<P, 4>(a, L, N);
   case 5: return fit_one<P, 5>(a, L, N);
   default: break;
   }
   if constexpr ( KMAX > 5 )
   {
      switch ( Kc )
      {
      case 6: return fit_one<P, 6>(a, L, N);
      case 7: return fit_one<P, 7>(a, L, N);
      case 8: return fit_one<P, 8>(a, L, N);
      default: break;
      }
   }
   FitOut f;
   f.ok = false;
   f.iters = 0;
   f.logL = 0.0f;
   f.E = 0.0f;
   return f;
}

/* The fits with 4 and 5 clusters run in lockstep (a.joint45) on the
 * register-slot instances whose VGPR budget holds their nine components'
 * responsibilities (9 P registers) besides X and log p(x) (4 P): at most 7
 * slots in the 128 VGPRs of 4 waves per SIMD, 10 in the 168 of 3. */
template <int P>
__device__ __forceinline__ constexpr bool joint45_built()
{
   return P > 0 && ((KINC_PAIR_WAVES == 4 && P <= 7) || (KINC_PAIR_WAVES == 3 && P <= 10));
}

template <int P, int KMAX>
__device__ __forceinline__ void process_pair(const SimArgs& a, int q)
{
   const Lds L = lds_view<P == 0>(a);
   const int lane = lane_id();
   const int S = a.S;
   const int MC = a.max_k;

   CENSUS_T(cp);
   int gx, gy;
   pair_genes(a, q, gx, gy);
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
      CENSUS_LAP(KINC_CENSUS_FETCH, cp);
      CENSUS_COUNT(KINC_CENSUS_PAIRS, 1);
      if ( N >= a.min_samples )
      {
         float bestValue = INFINITY;
         /* register-slot GMM: the fits with 2 and 3 clusters run together,
          * and (a.joint45, at most 7 slots) so do those with 4 and 5; the
          * second fit's labels wait in the tile region past its first row
          * until its turn in the selection */
         FitOut fs;
         int pend = 0;   /* cluster count of the joint fit whose result waits in fs */
         int8_t* const ws = (int8_t*)(L.tile + KINC_TILE_STRIDE);
         for ( int Kc = a.min_k; Kc <= a.max_k; ++Kc )
         {
            FitOut f;
            const int8_t* wsrc = L.w;
            if constexpr ( KMAX > 8 )
            {
               /* run-time K (memory store): max_clusters 9..64 */
               f = (a.clus == 1) ? gmm_fit_dyn(L, N, Kc, a.crit == 1) : kmeans_fit_dyn(a, L, N, Kc);
            }
            else if constexpr ( P > 0 )
            {
               if ( Kc == pend )
               {
                  f = fs;
                  wsrc = ws;
                  pend = 0;
               }
               else if ( a.clus == 1 && Kc == 2 && a.max_k >= 3 )
               {
                  gmm_fit_joint<P, 2, 3>(L, N, a.crit == 1, f, fs, ws);
                  pend = 3;
               }
               else if ( joint45_built<P>() && a.joint45 && a.clus == 1 && Kc == 4 && a.max_k >= 5 )
               {
                  if constexpr ( joint45_built<P>() ) gmm_fit_joint<P, 4, 5>(L, N, a.crit == 1, f, fs, ws);
                  pend = 5;
               }
               else
               {
                  f = fit_dispatch<P, KMAX>(a, L, N, Kc);
               }
            }
            else
            {
               f = fit_dispatch<P, KMAX>(a, L, N, Kc);
            }
            if ( a.clus == 1 ) CENSUS_RESET(cp);
            else CENSUS_LAP(KINC_CENSUS_OTHER_FIT, cp);
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
               copy_back(L, N, wsrc);
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
      CENSUS_LAP(KINC_CENSUS_SELECT, cp);
   }

   if constexpr ( KMAX > 8 )
   {
      /* up to 64 correlations: kept in LDS (the parameter area, free after
       * the fits), written out below */
      float* cl = L.prm;
      if ( K > 0 )
      {
         if ( a.corr == 0 )
         {
            if ( N < S ) fill_stale(a, L, q, N);
            for ( int k0 = 0; k0 < K; k0 += 12 )
            {
               pearson_group(a, L, k0, K - k0 < 12 ? K - k0 : 12, cl);
            }
         }
         else
         {
            for ( int k = 0; k < K; ++k )
            {
               const float r = spearman_k(a, L, k);
               if ( lane == 0 ) cl[k] = r;
            }
         }
      }
      wave_sync(true);
      if ( lane == 0 ) a.K_out[q] = (int8_t)K;
      float* co = a.corr_out + (size_t)q * MC;
      if ( lane < MC ) co[lane] = lane < K ? cl[lane] : NAN;
      if ( a.iters_out && lane < MC ) a.iters_out[(size_t)q * MC + lane] = itv;
      if ( a.labels_out )
      {
         int8_t* lo = a.labels_out + (size_t)q * S;
         for ( int s = lane; s < S; s += WAVE ) lo[s] = L.lab[s];
      }
      wave_sync(true);
      return;
   }
   else
   {
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
   wave_sync(L.lgbuf != nullptr);
   CENSUS_LAP(KINC_CENSUS_CORR, cp);
   }
}

} // namespace

/* ------------------------------------------------------------- kernels */

/*
 * The pair kernel of both execution forms.  Without a list (a.list == null)
 * the workgroups stride over the launch's pairs: the plain form's one
 * workgroup per pair, or resident waves over the whole block (persistent
 * form without clustering).  With a list, resident waves pop block-relative
 * pair ids from a device queue until it is empty, so every wave reaches the
 * exit test.
 */
#ifdef KINC_CENSUS
#define KINC_PAIR_KERNEL k_pairs_census
#else
#define KINC_PAIR_KERNEL k_pairs
#endif
template <int P, int KMAX, int W>
__global__ __launch_bounds__(64, W) void KINC_PAIR_KERNEL(SimArgs a)
{
   /* one loop for both: the next pair is the workgroup's next stride (no
    * list) or the next id popped from the bin's queue, so the per-pair code
    * is inlined once */
#ifdef KINC_CENSUS
   if ( lane_id() < KINC_CENSUS_COUNTERS ) g_census[lane_id()] = 0;
   wave_sync(false);
#endif
   CENSUS_T(kt);
   const bool queued = a.list != nullptr;
   const int len = queued ? *a.list_len : 0;
   const int end = a.first + a.launch_count;
   int q = a.first + (int)blockIdx.x;
   for ( ;; )
   {
      int pair;
      if ( !queued )
      {
         if ( q >= end ) break;
         pair = q;
         q += (int)gridDim.x;
      }
      else
      {
         CENSUS_T(tq);
         int idx = 0;
         if ( lane_id() == 0 ) idx = atomicAdd(a.list_head, 1);
         idx = bcast_i(idx, 0);
         CENSUS_LAP(KINC_CENSUS_QUEUE, tq);
         if ( idx >= len ) break;
         pair = a.list[idx];
      }
      process_pair<P, KMAX>(a, pair);
   }
#ifdef KINC_CENSUS
   CENSUS_LAP(KINC_CENSUS_WAVE, kt);
   wave_sync(false);
   if ( lane_id() < KINC_CENSUS_COUNTERS ) atomicAdd(&a.census[lane_id()], g_census[lane_id()]);
#endif
}

#endif
