/*
 * sim_util_kernels.hip — block preparation kernels of the pairwise
 * Similarity path: valid-sample counts and compact X, chunk maxima and
 * previous-greater links for the stale-read view, Pearson without
 * clustering with one pair per lane, sample-count binning for the
 * persistent form, and a zero fill.
 */
#include "sim_pair_impl.h"

/* Valid-sample count per pair of a block (one wave per pair, 4 per group),
 * and with xbuf non-null the pair's compact X (fetchPair's X,
 * similarity_serial.cpp:108-141) at xbuf[w][0..N), for the stale-read view
 * of later pairs. */
__global__ __launch_bounds__(256) void k_count_valid(const float* emx, int S, int G,
                                                     int64_t block_start, const int2* pairs,
                                                     int count, float min_expr, int32_t* outN,
                                                     float* xbuf)
{
   const int w = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
   if ( w >= count ) return;
   int gx, gy;
   if ( pairs )
   {
      const int2 v = pairs[w];
      const bool ok = 0 <= v.y && v.y < v.x && v.x < G;
      gx = ok ? v.x : 0;
      gy = ok ? v.y : 0;
   }
   else
   {
      pair_xy(block_start + w, gx, gy);
   }
   const float* r1 = emx + (size_t)gx * (size_t)S;
   const float* r2 = emx + (size_t)gy * (size_t)S;
   float2* xo = xbuf ? (float2*)xbuf + (size_t)w * (size_t)S : nullptr;
   const int lane = lane_id();
   int n = 0;
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
         const bool valid = s < S && !isnan(g1[j]) && !isnan(g2[j]) &&
                            !(g1[j] < min_expr || g2[j] < min_expr);
         const uint64_t m = ballot(valid);
         if ( xo && valid ) xo[n + popc(m & lanes_below())] = make_float2(g1[j], g2[j]);
         n += popc(m);
      }
   }
   if ( lane == 0 ) outN[w] = n;
}

/* prev[q]: the latest earlier pair of q's segment with more valid samples
 * than q, or -1 (one wave per pair, 4 per group; after k_chunk_max). */
__global__ __launch_bounds__(256) void k_prev_greater(const int32_t* N, const int32_t* cmax,
                                                      const int32_t* seg_lo, int count,
                                                      int32_t* prev)
{
   const int q = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
   if ( q >= count ) return;
   /* a segment start outside [0, q] is clamped into it, so that no read
    * leaves the block's arrays (the host entry point rejects such input;
    * the device entry point passes it through) */
   const int lo = seg_lo ? min(max(seg_lo[q], 0), q) : 0;
   const int d = prev_greater(N, cmax, q, N[q], lo);
   if ( lane_id() == 0 ) prev[q] = d;
}

/* max of N over 64-pair chunks (one wave per chunk, 4 per group) */
__global__ __launch_bounds__(256) void k_chunk_max(const int32_t* N, int count, int32_t* cmax)
{
   const int c = (int)(blockIdx.x * 4 + (threadIdx.x >> 6));
   if ( c * WAVE >= count ) return;
   const int i = c * WAVE + lane_id();
   int m = i < count ? N[i] : -1;
   for ( int off = WAVE / 2; off > 0; off >>= 1 )
   {
      const int o = __shfl_xor(m, off);
      m = o > m ? o : m;
   }
   if ( lane_id() == 0 ) cmax[c] = m;
}

/*
 * Bin the pairs of a block by register slots P(N) for the persistent form.
 * Bin b holds pairs whose N needs the kernel instance with slot count
 * bins.P[b]; pairs below min_samples (no clustering to do) go to the
 * smallest bin that still fits their N.  counters[nb] must be zero.
 */
__global__ __launch_bounds__(256) void k_bin_pairs(const int32_t* N, int count, BinBounds bins,
                                                   int nb, int32_t* counters, int32_t* lists)
{
   const int q = (int)(blockIdx.x * blockDim.x + threadIdx.x);
   if ( q >= count ) return;
   const int n = N[q];
   int b = nb - 1;
   for ( int i = 0; i < nb && i < KINC_MAX_BINS; ++i )
   {
      if ( n <= 64 * bins.P[i] )
      {
         b = i;
         break;
      }
   }
   const int pos = atomicAdd(&counters[b], 1);
   lists[(size_t)b * count + pos] = q;
}

/*
 * Pearson without clustering, one pair per lane (pairwise_pearson.cpp:12-52
 * with K = 1, fetchPair's labels, similarity_serial.cpp:108-141).  Lane l of
 * workgroup b owns block-relative pair q = first + 64 b + l and walks its
 * samples in index order, keeping the five moment sums in registers.  The
 * value Pearson reads at a valid index i is the block workspace entry i:
 * the pair's own compact X below N, then the chain of earlier pairs with
 * more valid samples (prev links), all read from the compact-X buffer that
 * k_count_valid wrote; 0 past the chain, as in a fresh workspace.  Eight
 * indices' loads are issued before any of them is used.
 */
__global__ __launch_bounds__(64) void k_pearson_lanes(SimArgs a)
{
   const int lane = lane_id();
   const int q = a.first + (int)blockIdx.x * WAVE + lane;
   const bool act = q < a.first + a.launch_count;
   const int S = a.S;
   const int MC = a.max_k;
   int gx = 0, gy = 0;
   if ( act ) pair_genes(a, q, gx, gy);
   const float* r1 = a.emx + (size_t)gx * (size_t)S;
   const float* r2 = a.emx + (size_t)gy * (size_t)S;
   const float2* xb = (const float2*)a.xbuf;
   int8_t* lab_out = (act && a.labels_out) ? a.labels_out + (size_t)q * (size_t)S : nullptr;
   const float me = a.min_expr;

   /* workspace entries [.., hi) come from pair d; d = -1: zeros */
   int d = act ? q : -1;
   int hi = act ? a.blockN[q] : S;
   int n = 0;
   float sx = 0.0f, sy = 0.0f, sxx = 0.0f, syy = 0.0f, sxy = 0.0f;
   double dx = 0.0, dy = 0.0, dxx = 0.0, dyy = 0.0, dxy = 0.0;

   constexpr int U = 8;
   for ( int i0 = 0; i0 < S; i0 += U )
   {
      int src[U];
#pragma unroll
      for ( int j = 0; j < U; ++j )
      {
         const int i = i0 + j;
         while ( d >= 0 && i >= hi && i < S )
         {
            d = a.prev[d];
            hi = d >= 0 ? a.blockN[d] : S;
         }
         src[j] = d;
      }
      float g1[U], g2[U];
      float2 w[U];
#pragma unroll
      for ( int j = 0; j < U; ++j )
      {
         const int i = i0 + j;
         const bool in = act && i < S;
         g1[j] = in ? r1[i] : 0.0f;
         g2[j] = in ? r2[i] : 0.0f;
         w[j] = (in && src[j] >= 0) ? xb[(size_t)src[j] * (size_t)S + i] : make_float2(0.0f, 0.0f);
      }
#pragma unroll
      for ( int j = 0; j < U; ++j )
      {
         const int i = i0 + j;
         if ( !act || i >= S ) continue;
         const bool nan = isnan(g1[j]) || isnan(g2[j]);
         const bool low = !nan && (g1[j] < me || g2[j] < me);
         if ( lab_out ) lab_out[i] = nan ? (int8_t)-9 : (low ? (int8_t)-6 : (int8_t)0);
         if ( nan || low ) continue;
         const float x = w[j].x;
         const float y = w[j].y;
         ++n;
         if ( !a.pearson_fp64 )
         {
            sx = sx + x;
            sy = sy + y;
            sxx = sxx + x * x;
            syy = syy + y * y;
            sxy = sxy + x * y;
         }
         else
         {
            const double xd = x, yd = y;
            dx = dx + xd;
            dy = dy + yd;
            dxx = dxx + xd * xd;
            dyy = dyy + yd * yd;
            dxy = dxy + xd * yd;
         }
      }
   }
   if ( !act ) return;

   float r = NAN;
   if ( n >= a.min_samples && !a.pearson_fp64 )
   {
      const float fn = (float)n;
      r = (fn * sxy - sx * sy) / sqrtf((fn * sxx - sx * sx) * (fn * syy - sy * sy));
   }
   else if ( n >= a.min_samples )
   {
      const double dn = (double)n;
      r = (float)((dn * dxy - dx * dy) / sqrt((dn * dxx - dx * dx) * (dn * dyy - dy * dy)));
   }
   a.K_out[q] = 1;
   float* co = a.corr_out + (size_t)q * MC;
   co[0] = r;
   for ( int k = 1; k < MC; ++k ) co[k] = NAN;
   if ( a.iters_out )
   {
      for ( int k = 0; k < MC; ++k ) a.iters_out[(size_t)q * MC + k] = 0;
   }
}

/* Zero a small int array (used instead of a memset node). */
__global__ void k_zero_i32(int32_t* p, int n)
{
   const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
   if ( i < n ) p[i] = 0;
}


extern "C" {
const void* kinc_kernel_count_valid() { return (const void*)&k_count_valid; }
const void* kinc_kernel_chunk_max() { return (const void*)&k_chunk_max; }
const void* kinc_kernel_prev_greater() { return (const void*)&k_prev_greater; }
const void* kinc_kernel_pearson_lanes() { return (const void*)&k_pearson_lanes; }
const void* kinc_kernel_bin_pairs() { return (const void*)&k_bin_pairs; }
const void* kinc_kernel_zero_i32() { return (const void*)&k_zero_i32; }
}

/* Device evaluation of the shared FP32 functions (diagnostic). */
__global__ void k_eval_math(int fn, const float* in, float* out, int64_t n)
{
   const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
   if ( i >= n ) return;
   const float x = in[i];
   float r;
   switch ( fn )
   {
   case 0: r = kinc_expf(x); break;
   case 1: r = kinc_logf(x); break;
   case 2: r = sqrtf(x); break;
   case 4: r = kinc_logf_ge1(x); break;
   case 5: r = kinc_expf_le0(x); break;
   default: r = 1.0f / x; break;
   }
   out[i] = r;
}

extern "C" int kinc_selftest_math(int32_t fn, const float* in, float* out, int64_t n)
{
   if ( fn < 0 || fn > 5 || n < 0 || (n > 0 && (!in || !out)) ) return -1;
   if ( n == 0 ) return 0;
   float* d_in = nullptr;
   float* d_out = nullptr;
   const size_t bytes = sizeof(float) * (size_t)n;
   if ( hipMalloc(&d_in, bytes) != hipSuccess ) return -3;
   if ( hipMalloc(&d_out, bytes) != hipSuccess )
   {
      (void)hipFree(d_in);
      return -3;
   }
   int rc = 0;
   if ( hipMemcpy(d_in, in, bytes, hipMemcpyHostToDevice) != hipSuccess ) rc = -2;
   if ( rc == 0 )
   {
      const unsigned grid = (unsigned)((n + 255) / 256);
      hipLaunchKernelGGL(k_eval_math, dim3(grid), dim3(256), 0, 0, fn, d_in, d_out, n);
      if ( hipGetLastError() != hipSuccess ) rc = -2;
   }
   if ( rc == 0 && hipMemcpy(out, d_out, bytes, hipMemcpyDeviceToHost) != hipSuccess ) rc = -2;
   (void)hipFree(d_in);
   (void)hipFree(d_out);
   return rc;
}

/*
 * VALU issue-rate probe (diagnostic).  Each wave runs `iters` rounds of 128
 * instructions of one kind OP on its first LANES lanes, spread over CHAINS
 * independent accumulators (CHAINS = 1: one dependent chain).  The issue
 * count per wave is exactly 128 * iters VALU instructions (the loop counter
 * is scalar), so the launch time gives the device's wave64 issue rate for
 * that instruction, exec mask and dependency pattern: the ceilings the pair
 * kernels' instruction mix is measured against.
 *   OP 0 v_add_f32   1 v_mul_f32   2 v_fma_f32   3 v_fmac_f32   4 v_max_f32
 *   5 v_rndne_f32    6 v_ldexp_f32 7 v_cvt_i32_f32 8 v_cndmask_b32
 *   9 v_fmaak_f32 (FMA with a literal)
 *   10 one kinc_expf_le0 (the pair kernels' E-step exponential), 11 one
 *   kinc_expf: compiled code, so the unit is one exponential of a known
 *   instruction count, and the rate is that of the exponentials' own mix of
 *   opcodes (acc = exp(-acc) stays in (0.36, 1], no subnormals)
 */
template <int OP>
__device__ __forceinline__ void probe_op(float& acc, float inc)
{
   if constexpr ( OP == 0 ) asm volatile("v_add_f32 %0, %0, %1" : "+v"(acc) : "v"(inc));
   if constexpr ( OP == 1 ) asm volatile("v_mul_f32 %0, %0, %1" : "+v"(acc) : "v"(inc));
   if constexpr ( OP == 2 ) asm volatile("v_fma_f32 %0, %0, %1, %1" : "+v"(acc) : "v"(inc));
   if constexpr ( OP == 3 ) asm volatile("v_fmac_f32 %0, %1, %1" : "+v"(acc) : "v"(inc));
   if constexpr ( OP == 4 ) asm volatile("v_max_f32 %0, %0, %1" : "+v"(acc) : "v"(inc));
   if constexpr ( OP == 5 ) asm volatile("v_rndne_f32 %0, %0" : "+v"(acc));
   if constexpr ( OP == 6 ) asm volatile("v_ldexp_f32 %0, %0, %1" : "+v"(acc) : "v"(inc));
   if constexpr ( OP == 7 ) asm volatile("v_cvt_i32_f32 %0, %0" : "+v"(acc));
   if constexpr ( OP == 8 ) asm volatile("v_cndmask_b32 %0, %0, %1, vcc" : "+v"(acc) : "v"(inc) : "vcc");
   if constexpr ( OP == 9 ) asm volatile("v_fmaak_f32 %0, %0, %1, 0x3f800000" : "+v"(acc) : "v"(inc));
   if constexpr ( OP == 10 ) acc = kinc_expf_le0(-acc);
   if constexpr ( OP == 11 ) acc = kinc_expf(-acc);
}

template <int CHAINS, int LANES, int OP>
__global__ __launch_bounds__(64) void k_valu_probe(float* out, int iters)
{
   const int lane = (int)threadIdx.x;
   float acc[CHAINS];
#pragma unroll
   for ( int c = 0; c < CHAINS; ++c ) acc[c] = (float)c;
   const float inc = 1.0e-7f * (float)(lane + 1);
   if ( lane < LANES )
   {
      for ( int it = 0; it < iters; ++it )
      {
#pragma unroll
         for ( int r = 0; r < 128 / CHAINS; ++r )
         {
#pragma unroll
            for ( int c = 0; c < CHAINS; ++c ) probe_op<OP>(acc[c], inc);
         }
      }
   }
   float s = 0.0f;
#pragma unroll
   for ( int c = 0; c < CHAINS; ++c ) s = s + acc[c];
   out[(size_t)blockIdx.x * 64 + lane] = s;
}

/* modes of kinc_selftest_valu_probe: (independent chains, active lanes, op) */
static const int kProbeModes[][3] = {
   {8, 64, 0}, {8, 16, 0}, {8, 4, 0}, {1, 64, 0}, {1, 16, 0}, {8, 32, 0},
   {8, 64, 1}, {8, 64, 2}, {8, 64, 3}, {8, 64, 4}, {8, 64, 5}, {8, 64, 6},
   {8, 64, 7}, {8, 64, 8}, {8, 64, 9}, {1, 64, 2}, {8, 16, 2}, {8, 64, 10}, {8, 64, 11}};
static const int kProbeCount = (int)(sizeof(kProbeModes) / sizeof(kProbeModes[0]));

template <int CH, int LN, int OP>
static const void* probe_fn() { return (const void*)&k_valu_probe<CH, LN, OP>; }

extern "C" int kinc_selftest_valu_probe(int32_t mode, int32_t iters, int32_t workgroups,
                                        int32_t lds_bytes, float* ms)
{
   if ( mode < 0 || mode >= kProbeCount || iters < 1 ||
        workgroups < 1 || lds_bytes < 0 || lds_bytes > 65536 || !ms )
   {
      return -1;
   }
   static const void* const fns[] = {
      probe_fn<8, 64, 0>(), probe_fn<8, 16, 0>(), probe_fn<8, 4, 0>(), probe_fn<1, 64, 0>(),
      probe_fn<1, 16, 0>(), probe_fn<8, 32, 0>(), probe_fn<8, 64, 1>(), probe_fn<8, 64, 2>(),
      probe_fn<8, 64, 3>(), probe_fn<8, 64, 4>(), probe_fn<8, 64, 5>(), probe_fn<8, 64, 6>(),
      probe_fn<8, 64, 7>(), probe_fn<8, 64, 8>(), probe_fn<8, 64, 9>(), probe_fn<1, 64, 2>(),
      probe_fn<8, 16, 2>(), probe_fn<8, 64, 10>(), probe_fn<8, 64, 11>()};
   const void* fn = fns[mode];
   float* d_out = nullptr;
   if ( hipMalloc(&d_out, sizeof(float) * 64 * (size_t)workgroups) != hipSuccess ) return -3;
   hipEvent_t e0 = nullptr, e1 = nullptr;
   int rc = 0;
   if ( hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ) rc = -2;
   for ( int rep = 0; rep < 2 && rc == 0; ++rep )
   {
      /* the first launch warms up; the second is timed */
      void* args[] = {&d_out, &iters};
      if ( hipEventRecord(e0, 0) != hipSuccess ) rc = -2;
      if ( rc == 0 && hipLaunchKernel(fn, dim3((unsigned)workgroups), dim3(64), args,
                                      (size_t)lds_bytes, 0) != hipSuccess ) rc = -2;
      if ( rc == 0 && hipEventRecord(e1, 0) != hipSuccess ) rc = -2;
      if ( rc == 0 && hipEventSynchronize(e1) != hipSuccess ) rc = -2;
      if ( rc == 0 && hipEventElapsedTime(ms, e0, e1) != hipSuccess ) rc = -2;
   }
   if ( e0 ) (void)hipEventDestroy(e0);
   if ( e1 ) (void)hipEventDestroy(e1);
   (void)hipFree(d_out);
   return rc;
}
