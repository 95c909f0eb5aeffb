/*
 * kinc_sim.cpp — host side of the C-ABI in include/kinc_sim.h.
 *
 * Replaces Similarity::OpenCL::initialize / Worker::execute
 * (src/core/similarity_opencl.cpp:34-70, src/core/similarity_opencl_worker.cpp:93-271):
 * the expression matrix is uploaded once and stays resident; a work block is
 * executed either in the plain form (one launch per ksize sub-batch, one
 * wave per pair, the register-slot count fixed by the sample count S) or in
 * the persistent form (valid-sample counts first, pairs binned by the slot
 * count their N needs, each bin drained by resident waves popping a device
 * queue).  Both forms run the same per-pair device code and give identical
 * results.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kinc_selftest.h"
#include "../../include/kinc_sim.h"
#include "sim_device.h"

#define KINC_FOR_EACH_INSTANCE(X)                                                   \
   X(5, 0) X(5, 1) X(5, 2) X(5, 3) X(5, 4) X(5, 5) X(5, 6) X(5, 7) X(5, 8) X(5, 10)   \
   X(5, 12) X(5, 14) X(5, 16) X(8, 0) X(8, 2) X(8, 4) X(8, 8) X(64, 0)
/* small instances also built for 5 waves per SIMD */
#define KINC_FOR_EACH_INSTANCE_W5(X) X(5, 1) X(5, 2) X(5, 3) X(5, 4) X(5, 5) X(5, 6) X(8, 2) X(8, 4)
/* the register-slot instances that spill at 4 waves per SIMD, also built for
 * 3 (168 VGPRs) */
#define KINC_FOR_EACH_INSTANCE_WIDE(X)                                              \
   X(5, 5) X(5, 6) X(5, 7) X(5, 8) X(5, 10) X(5, 12) X(5, 14) X(5, 16) X(8, 4) X(8, 8)

extern "C" {
#define KINC_DECL(K, P) const void* kinc_pair_kernel_k##K##_p##P();
KINC_FOR_EACH_INSTANCE(KINC_DECL)
#undef KINC_DECL
#define KINC_DECL(K, P) const void* kinc_pair_kernel_k##K##_p##P##_w5();
KINC_FOR_EACH_INSTANCE_W5(KINC_DECL)
#undef KINC_DECL
#define KINC_DECL(K, P) const void* kinc_pair_kernel_k##K##_p##P##_w3();
KINC_FOR_EACH_INSTANCE_WIDE(KINC_DECL)
#undef KINC_DECL
const void* kinc_kernel_count_valid();
const void* kinc_kernel_chunk_max();
const void* kinc_kernel_prev_greater();
const void* kinc_kernel_pearson_lanes();
const void* kinc_kernel_bin_pairs();
const void* kinc_kernel_zero_i32();
const void* kinc_pair_kernel_census_k5_p7();
}

namespace {

/* Register-slot counts that have a kernel instance (64*P samples per pair),
 * per cluster bound KMAX. */
const int kSlotList5[] = {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16};
const int kSlotList8[] = {2, 4, 8};
const int kMaxBins = KINC_MAX_BINS;

thread_local char g_create_error[256];

/* Slot classes used for a handle: every instance up to the one S needs,
 * or the memory-store instance (P = 0) alone above the register range
 * (1,024 samples for up to 5 clusters, 512 for up to 8).  Writes at most
 * kMaxBins classes to out and returns their number. */
int slot_classes(int S, int kmax, int* out)
{
   if ( kmax > 8 )
   {
      /* max_clusters 9..64: the run-time-K memory-store instance */
      out[0] = 0;
      return 1;
   }
   const int* list = kmax == 5 ? kSlotList5 : kSlotList8;
   const int n = kmax == 5 ? 12 : 3;
   if ( S > 64 * list[n - 1] )
   {
      out[0] = 0;
      return 1;
   }
   int c = 0;
   for ( int i = 0; i < n; ++i )
   {
      out[c++] = list[i];
      if ( 64 * list[i] >= S ) break;
   }
   return c;
}

/* The pair kernel for (P, KMAX); one instance serves both forms.  waves
 * picks the build for that many waves per SIMD where one exists (5: the small
 * instances; 3: the instances of 5 slots or more), else the 4-wave one. */
const void* kinc_select_pair_kernel(int P, int kmax, int waves = 4)
{
#define KINC_PICK(K, PP)                                                      \
   if ( kmax == K && P == PP ) return kinc_pair_kernel_k##K##_p##PP##_w5();
   if ( waves == 5 )
   {
      KINC_FOR_EACH_INSTANCE_W5(KINC_PICK)
   }
#undef KINC_PICK
#define KINC_PICK(K, PP)                                                      \
   if ( kmax == K && P == PP ) return kinc_pair_kernel_k##K##_p##PP##_w3();
   if ( waves == 3 )
   {
      KINC_FOR_EACH_INSTANCE_WIDE(KINC_PICK)
   }
#undef KINC_PICK
#define KINC_PICK(K, PP)                                                      \
   if ( kmax == K && P == PP ) return kinc_pair_kernel_k##K##_p##PP();
   KINC_FOR_EACH_INSTANCE(KINC_PICK)
#undef KINC_PICK
   return nullptr;
}

int next_pow2(int n)
{
   int p = 2;
   while ( p < n ) p <<= 1;
   return p;
}

size_t align16(size_t v)
{
   return (v + 15) & ~(size_t)15;
}

} // namespace

struct kinc_sim
{
   kinc_sim_params p;
   int device = 0;
   int G = 0;
   int S = 0;
   int kmax = 5;
   int num_cus = 0;
   hipStream_t stream = nullptr;
   float* d_emx = nullptr;
   double* d_logN = nullptr;
   int32_t* d_blockN = nullptr;
   int32_t* d_chunkMax = nullptr;
   float* d_xbuf = nullptr;        /* Pearson, register-slot layout: compact X per pair */
   int32_t* d_prev = nullptr;      /* with d_xbuf: previous-greater link per pair */
   int32_t* d_counters = nullptr;  /* [2*kMaxBins]: bin sizes, then pop heads */
   int32_t* d_lists = nullptr;     /* [kMaxBins][KINC_WORK_BLOCK_SIZE] */
   int classes[kMaxBins] = {};     /* slot count of each bin (0 = memory store) */
   int nbins = 0;
   BinBounds bins{};               /* the same, as the binning kernel takes it */
   char* d_scratch = nullptr;      /* memory-store slices */
   int scratch_slots = 0;
   int cap = KINC_WORK_BLOCK_SIZE; /* largest count of one call */
   int32_t* d_pairs = nullptr;     /* staging of host pair lists [cap][2] */
   int32_t* d_seg = nullptr;       /* staging of host segment starts [cap] */
   /* staging for the host-buffer entry point */
   int8_t* d_K = nullptr;
   int8_t* d_labels = nullptr;
   float* d_corr = nullptr;
   uint16_t* d_iters = nullptr;
   /* the persistent form drains its bins concurrently on these streams */
   static const int kAux = 4;
   hipStream_t aux[kAux] = {};
   int naux = kAux;                /* auxiliary streams the bins are spread over (1..kAux) */
   hipEvent_t binned = nullptr;
   hipEvent_t drained[kAux] = {};
   SimArgs base{};
   size_t lds_bytes = 0;           /* per-wave LDS of the current configuration */
   size_t lds_small = 0;           /* per-wave LDS without room for the lockstep 4+5-cluster tile */
   size_t lds_j45 = 0;             /* per-wave LDS with that room, or 0 when the fits cannot run so */
   char err[256] = {};
   /* kernel timing */
   bool profiling = false;
   bool capturing = false;         /* between kinc_sim_graph_begin and _end */
   /* waves per SIMD of the pair-kernel builds the register-slot bins run:
    * 4, or 5 for the small instances where the per-wave LDS leaves room for
    * 20 waves per CU, or 3 (168 VGPRs per wave) where it allows fewer than
    * 16 */
   int waves = 4;
   struct Timed { int P; hipEvent_t a, b; };
   std::vector<Timed> timed;
   std::vector<hipEvent_t> event_pool;
};

/* The pair kernel a launch of slot count P uses: the phase-census build
 * while a census is on and one exists for P, else the product build. */
static const void* pair_kernel(const kinc_sim* h, int P)
{
   if ( h->base.census && P == 7 && h->kmax == 5 && h->waves == 4 ) return kinc_pair_kernel_census_k5_p7();
   return kinc_select_pair_kernel(P, h->kmax, h->waves);
}

static int fail(kinc_sim* h, int code, const std::string& msg)
{
   char* dst = h ? h->err : g_create_error;
   std::snprintf(dst, 256, "%s", msg.c_str());
   return code;
}

/* a failed HIP call (formats into the handle's buffer; builds no string) */
static int fail_hip(kinc_sim* h, const char* call, hipError_t e)
{
   char* dst = h ? h->err : g_create_error;
   std::snprintf(dst, 256, "%s: %s", call, hipGetErrorString(e));
   return KINC_E_DEVICE;
}

/* fail() for a message that needs no formatting (builds no string) */
static int fail_c(kinc_sim* h, int code, const char* msg)
{
   char* dst = h ? h->err : g_create_error;
   std::snprintf(dst, 256, "%s", msg);
   return code;
}

#define HIP_TRY(h, call)                                                              \
   do                                                                                 \
   {                                                                                  \
      hipError_t e_ = (call);                                                         \
      if ( e_ != hipSuccess )                                                         \
      {                                                                               \
         return fail_hip((h), #call, e_);                                              \
      }                                                                               \
   } while ( 0 )

extern "C" void kinc_sim_default_params(kinc_sim_params* p)
{
   std::memset(p, 0, sizeof(*p));
   p->clus_method = KINC_CLUS_NONE;
   p->corr_method = KINC_CORR_PEARSON;
   p->min_expr = -INFINITY;
   p->min_samples = 30;
   p->min_clusters = 1;
   p->max_clusters = 5;
   p->criterion = KINC_CRIT_ICL;
   p->pre_out = 0;
   p->post_out = 0;
   p->min_corr = 0.5f;
   p->max_corr = 1.0f;
   p->ksize = 4096;
   p->form = KINC_FORM_PERSISTENT;
}

extern "C" int64_t kinc_pair_count(int32_t G)
{
   return (int64_t)G * (G - 1) / 2;
}

extern "C" void kinc_pair_index(int64_t index, int32_t* x, int32_t* y)
{
   int64_t xx = (int64_t)((1.0 + std::sqrt(1.0 + 8.0 * (double)index)) / 2.0);
   while ( xx * (xx - 1) / 2 > index ) --xx;
   while ( (xx + 1) * xx / 2 <= index ) ++xx;
   *x = (int32_t)xx;
   *y = (int32_t)(index - xx * (xx - 1) / 2);
}

static int validate(const kinc_sim_params* p, int G, int S, const char** why)
{
   if ( p->clus_method < 0 || p->clus_method > 2 ) { *why = "clus_method must be 0 (none), 1 (gmm) or 2 (kmeans)"; return KINC_E_INVALID; }
   if ( p->corr_method < 0 || p->corr_method > 1 ) { *why = "corr_method must be 0 (pearson) or 1 (spearman)"; return KINC_E_INVALID; }
   if ( p->criterion < 0 || p->criterion > 1 ) { *why = "criterion must be 0 (BIC) or 1 (ICL)"; return KINC_E_INVALID; }
   if ( p->min_samples < 1 ) { *why = "min_samples must be >= 1"; return KINC_E_INVALID; }
   if ( p->min_clusters < 1 || p->max_clusters < 1 || p->max_clusters > KINC_MAX_CLUSTER_SIZE )
   {
      *why = "cluster range must lie in [1, 64]";
      return KINC_E_INVALID;
   }
   if ( p->max_clusters < p->min_clusters )
   {
      /* similarity.cpp:181-187 */
      *why = "Minimum clusters must be less than or equal to maximum clusters.";
      return KINC_E_INVALID;
   }
   if ( p->ksize < 1 ) { *why = "ksize must be >= 1"; return KINC_E_INVALID; }
   if ( p->batch_capacity < 0 || p->batch_capacity > (1 << 24) )
   {
      *why = "batch_capacity must lie in [0, 2^24]";
      return KINC_E_INVALID;
   }
   if ( p->form != KINC_FORM_PLAIN && p->form != KINC_FORM_PERSISTENT ) { *why = "form must be 0 (plain) or 1 (persistent)"; return KINC_E_INVALID; }
   if ( p->pearson_fp64 != 0 && p->pearson_fp64 != 1 ) { *why = "pearson_fp64 must be 0 or 1"; return KINC_E_INVALID; }
   for ( int r : p->reserved )
   {
      if ( r != 0 ) { *why = "reserved fields must be zero"; return KINC_E_INVALID; }
   }
   if ( G < 2 || S < 1 ) { *why = "need at least 2 genes and 1 sample"; return KINC_E_INVALID; }
   if ( S > KINC_MAX_SAMPLES )
   {
      *why = "sample_size > 32768 is not supported";
      return KINC_E_UNSUPPORTED;
   }
   return KINC_OK;
}

static const size_t kLdsPerCu = (size_t)160 * 1024;

/* per-wave LDS of a handle with or without room for the lockstep 4- and
 * 5-cluster tile (0 with it when those fits are not offered) */
static size_t lds_for(const kinc_sim* h, bool j45)
{
   return j45 ? h->lds_j45 : h->lds_small;
}

/* the 5-wave builds of the small register-slot instances: where the LDS
 * leaves room for 20 waves per CU */
static bool w5_ok(const kinc_sim* h, bool j45)
{
   const size_t lds = lds_for(h, j45);
   return h->classes[0] != 0 && lds != 0 && lds * 20 <= kLdsPerCu;
}

/* the 3-wave builds (168 VGPRs) of the instances of 5 slots or more: where
 * the handle has such bins and the LDS allows fewer than 16 waves per CU */
static bool w3_ok(const kinc_sim* h, bool j45)
{
   const size_t lds = lds_for(h, j45);
   return h->classes[0] != 0 && lds != 0 && h->S > 4 * 64 && kLdsPerCu / lds < 16;
}

/* the wave counts a handle's register-slot bins can run with */
static bool waves_allowed(const kinc_sim* h, int waves, bool j45)
{
   return waves == 4 || (waves == 5 && w5_ok(h, j45)) || (waves == 3 && w3_ok(h, j45));
}

/* set a launch configuration (checked by the caller) */
static void apply_config(kinc_sim* h, int waves, int naux, bool j45)
{
   h->waves = waves;
   h->naux = naux;
   h->base.joint45 = j45 ? 1 : 0;
   h->lds_bytes = lds_for(h, j45);
}

/*
 * The parameters, bins and working-set carve-up of a handle for an S-sample
 * matrix: everything but the device buffers.  Allocates nothing and touches
 * no device (kinc_sim_plan_get runs it on a handle of its own).
 */
static void setup_layout(kinc_sim* h, const kinc_sim_params* params, int G, int S)
{
   h->p = *params;
   h->G = G;
   h->S = S;
   h->kmax = params->max_clusters <= 5 ? 5 : (params->max_clusters <= 8 ? 8 : 64);
   h->cap = params->batch_capacity > 0 ? params->batch_capacity : KINC_WORK_BLOCK_SIZE;

   /* bins: every register-slot count up to the one S needs, or the single
    * memory-store class; a bin takes pairs with 64*P >= N (the memory-store
    * bin takes every pair) */
   h->nbins = slot_classes(S, h->kmax, h->classes);
   for ( int b = 0; b < kMaxBins; ++b )
   {
      h->bins.P[b] = b < h->nbins ? (h->classes[b] > 0 ? h->classes[b] : (1 << 20)) : 0;
   }

   SimArgs& a = h->base;
   a = SimArgs{};
   a.G = G;
   a.S = S;
   a.clus = params->clus_method;
   a.corr = params->corr_method;
   a.min_expr = params->min_expr;
   a.min_samples = params->min_samples;
   a.min_k = params->min_clusters;
   a.max_k = params->max_clusters;
   a.crit = params->criterion;
   a.pre_out = params->pre_out ? 1 : 0;
   a.post_out = params->post_out ? 1 : 0;
   a.pearson_fp64 = params->pearson_fp64;
   a.np2 = next_pow2(S);
   a.tile_stride = KINC_TILE_STRIDE;
   a.mem_slots = (S + 63) / 64;
   const size_t tile_bytes = sizeof(float) * (size_t)(1 + 3 * h->kmax) * (size_t)a.tile_stride;
   const size_t sort_bytes = sizeof(float) * 2 * (size_t)a.np2;

   a.joint45 = 0;
   if ( h->classes[0] != 0 )
   {
      /* everything in LDS: labels, work labels, component parameters (nine
       * sets: the fits with 4 and 5 clusters run together), then the compact
       * X and after it the shared region that the sort buffers and K-means's
       * chain tile and log-likelihood terms take in turn.  A GMM fit holds X
       * in registers and lays its chain tile over X and the shared region,
       * which therefore follow each other: 1+3K rows for a single fit, 17
       * for the fits with 2 and 3 clusters in lockstep and 29 for those with
       * 4 and 5; the second lockstep fit's labels wait in the shared region
       * past its first row. */
      size_t off = 0;
      a.scratch = nullptr;
      a.lds_lab = (int32_t)off;
      off = align16(off + (size_t)S);
      a.lds_w = (int32_t)off;
      off = align16(off + (size_t)S);
      a.lds_prm = (int32_t)off;
      off = align16(off + sizeof(float) * 8 * 9);
      a.lds_x = (int32_t)off;
      const size_t x_bytes = align16(sizeof(float) * 2 * (size_t)S);
      off += x_bytes;
      a.lds_u = (int32_t)off;
      a.lds_tile = a.lds_u;
      const size_t row = sizeof(float) * (size_t)a.tile_stride;
      /* the sort buffers serve outlier removal and Spearman only */
      const bool sorts = a.corr == 1 || a.pre_out || a.post_out;
      size_t shared = std::max(row, sorts ? sort_bytes : (size_t)0);
      if ( a.clus == 1 )
      {
         /* the joint fits' second labels after the entropy chain's row; a
          * single fit's 1+3K rows and the lockstep fits' 17 over X and this
          * region */
         shared = std::max(shared, row + (size_t)S);
         size_t over_x = tile_bytes;
         if ( a.max_k >= 3 && a.min_k <= 2 ) over_x = std::max(over_x, 17 * row);
         if ( over_x > x_bytes ) shared = std::max(shared, over_x - x_bytes);
      }
      else if ( a.clus == 2 )
      {
         /* the mean chains' 2K rows; the log-likelihood terms pass through
          * the same region in chunks of a.ll_cap */
         shared = std::max(shared, 2 * (size_t)h->kmax * row);
      }
      a.ll_cap = (int32_t)((shared / sizeof(float)) & ~(size_t)3);
      h->lds_small = align16(off + shared);
      h->lds_j45 = 0;
      if ( a.clus == 1 && a.max_k >= 5 && a.min_k <= 4 )
      {
         /* the lockstep 4+5-cluster tile: 29 rows over X and the shared
          * region, offered while at least 8 waves per CU stay resident */
         const size_t need = 29 * row > x_bytes ? 29 * row - x_bytes : 0;
         const size_t lds = align16(off + std::max(shared, need));
         if ( lds * 8 <= kLdsPerCu ) h->lds_j45 = lds;
      }
      /* the lockstep fits where they can run; 5-wave builds where the LDS
       * then leaves room for 20 waves per CU, 3-wave ones where it allows
       * fewer than 16 (their 168 VGPRs beat the 4-wave builds' spills at
       * 12-15 waves per CU) */
      const bool j45 = h->lds_j45 != 0;
      const int waves = w5_ok(h, j45) ? 5 : (w3_ok(h, j45) ? 3 : 4);
      apply_config(h, waves, h->naux, j45);
   }
   else
   {
      /* memory store: per-pair arrays (X, labels, work labels, sort buffers)
       * in a scratch slice per workgroup, the parameters and the chain tile
       * in LDS */
      size_t off = 0;
      a.lds_x = (int32_t)off;
      off = align16(off + sizeof(float) * 2 * (size_t)S);
      a.lds_lab = (int32_t)off;
      off = align16(off + (size_t)S);
      a.lds_w = (int32_t)off;
      off = align16(off + (size_t)S);
      a.lds_u = (int32_t)off;
      off = align16(off + sort_bytes);
      a.ll_cap = 2 * a.np2;   /* the sort buffers: room for every term */
      a.scr_lg = (int32_t)off;
      /* per-point responsibilities and log p(x); the run-time-K fits keep
       * log p(x) only and recompute the rest */
      const size_t lg_rows = h->kmax > 8 ? 1 : (size_t)(h->kmax + 1);
      off = align16(off + sizeof(float) * lg_rows * (size_t)a.mem_slots * 64);
      a.scr_y = (int32_t)off;
      off = align16(off + (size_t)a.mem_slots * 64);
      a.scratch_stride = (int64_t)((off + 255) & ~(size_t)255);
      a.lds_prm = 0;
      if ( h->kmax > 8 )
      {
         a.lds_tile = (int32_t)align16(KINC_DYN_PRM_BYTES);
         h->lds_bytes = align16((size_t)a.lds_tile + sizeof(float) * KINC_DYN_TILE_ROWS * (size_t)a.tile_stride);
      }
      else
      {
         a.lds_tile = (int32_t)align16(sizeof(float) * 8 * 8);
         h->lds_bytes = align16((size_t)a.lds_tile + tile_bytes);
      }
      h->lds_small = h->lds_bytes;
      h->lds_j45 = 0;
      h->waves = 4;
   }
}

extern "C" int kinc_sim_plan_get(const kinc_sim_params* params, int32_t G, int32_t S,
                                 kinc_sim_plan* out)
{
   if ( !params || !out ) return fail(nullptr, KINC_E_INVALID, "null argument");
   const char* why = nullptr;
   const int rc = validate(params, G, S, &why);
   if ( rc != KINC_OK ) return fail_c(nullptr, rc, why);
   /* the layout a handle would get (setup_layout touches no device) */
   kinc_sim h;
   setup_layout(&h, params, G, S);
   std::memset(out, 0, sizeof(*out));
   out->bins = h.nbins;
   for ( int b = 0; b < h.nbins && b < (int)(sizeof(out->slots) / sizeof(out->slots[0])); ++b ) out->slots[b] = h.classes[b];
   out->lds_small = (int32_t)h.lds_small;
   out->lds_joint45 = (int32_t)h.lds_j45;
   out->pair_waves_allowed = (1 << 4) | ((w5_ok(&h, false) || w5_ok(&h, true)) ? 1 << 5 : 0) |
                             ((w3_ok(&h, false) || w3_ok(&h, true)) ? 1 << 3 : 0);
   out->tuning.pair_waves = h.waves;
   out->tuning.aux_streams = h.naux;
   out->tuning.joint45 = h.base.joint45;
   return KINC_OK;
}

extern "C" int kinc_sim_create(const kinc_sim_params* params, const float* emx,
                               int32_t G, int32_t S, int32_t device, kinc_sim** out)
{
   if ( !params || !emx || !out ) return fail(nullptr, KINC_E_INVALID, "null argument");
   *out = nullptr;
   const char* why = nullptr;
   int rc = validate(params, G, S, &why);
   if ( rc != KINC_OK ) return fail_c(nullptr, rc, why);

   kinc_sim* h = new kinc_sim();
   h->device = device;
   setup_layout(h, params, G, S);

   auto bail = [&](int code) {
      std::snprintf(g_create_error, sizeof(g_create_error), "%s", h->err);
      kinc_sim_destroy(h);
      return code;
   };
#define CREATE_TRY(call)                                                             \
   do                                                                                \
   {                                                                                 \
      hipError_t e_ = (call);                                                        \
      if ( e_ != hipSuccess )                                                        \
      {                                                                              \
         std::snprintf(h->err, sizeof(h->err), "%s: %s", #call, hipGetErrorString(e_)); \
         return bail(KINC_E_DEVICE);                                                 \
      }                                                                              \
   } while ( 0 )

   CREATE_TRY(hipSetDevice(device));
   hipDeviceProp_t prop;
   CREATE_TRY(hipGetDeviceProperties(&prop, device));
   h->num_cus = prop.multiProcessorCount;
   CREATE_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
   for ( int i = 0; i < kinc_sim::kAux; ++i )
   {
      CREATE_TRY(hipStreamCreateWithFlags(&h->aux[i], hipStreamNonBlocking));
      CREATE_TRY(hipEventCreateWithFlags(&h->drained[i], hipEventDisableTiming));
   }
   CREATE_TRY(hipEventCreateWithFlags(&h->binned, hipEventDisableTiming));

   const size_t emx_bytes = (size_t)G * (size_t)S * sizeof(float);
   CREATE_TRY(hipMalloc(&h->d_emx, emx_bytes));
   CREATE_TRY(hipMemcpy(h->d_emx, emx, emx_bytes, hipMemcpyHostToDevice));

   /* log(N) of the BIC term (pairwise_clustering.cpp:160), evaluated by the
    * host libm once per possible N */
   std::vector<double> logN((size_t)S + 1);
   for ( int n = 0; n <= S; ++n ) logN[(size_t)n] = std::log((double)n);
   CREATE_TRY(hipMalloc(&h->d_logN, sizeof(double) * logN.size()));
   CREATE_TRY(hipMemcpy(h->d_logN, logN.data(), sizeof(double) * logN.size(), hipMemcpyHostToDevice));

   const size_t B = (size_t)h->cap;
   CREATE_TRY(hipMalloc(&h->d_blockN, sizeof(int32_t) * B));
   CREATE_TRY(hipMalloc(&h->d_chunkMax, sizeof(int32_t) * (B / 64 + 1)));
   CREATE_TRY(hipMalloc(&h->d_counters, sizeof(int32_t) * 2 * kMaxBins));
   CREATE_TRY(hipMalloc(&h->d_lists, sizeof(int32_t) * B * kMaxBins));

   SimArgs& a = h->base;
   a.emx = h->d_emx;
   a.logN = h->d_logN;
   a.blockN = h->d_blockN;
   a.chunkMax = h->d_chunkMax;
   if ( params->corr_method == KINC_CORR_PEARSON &&
        (h->classes[0] != 0 || params->clus_method == KINC_CLUS_NONE) &&
        sizeof(float) * 2 * B * (size_t)S <= ((size_t)4 << 30) )
   {
      /* Pearson's stale reads copy earlier pairs' compact X from here
       * instead of fetching and compacting their rows again (256 MiB for
       * 32768 pairs of 1024 samples; at most 4 GiB) */
      CREATE_TRY(hipMalloc(&h->d_xbuf, sizeof(float) * 2 * B * (size_t)S));
      CREATE_TRY(hipMalloc(&h->d_prev, sizeof(int32_t) * B));
      a.xbuf = h->d_xbuf;
      a.prev = h->d_prev;
   }
   if ( h->classes[0] == 0 )
   {
      /* memory store: one scratch slice per resident workgroup */
      const void* fn = kinc_select_pair_kernel(0, h->kmax);
      int per_cu = 0;
      CREATE_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, h->lds_bytes));
      h->scratch_slots = std::max(1, per_cu) * h->num_cus;
      CREATE_TRY(hipMalloc(&h->d_scratch, (size_t)h->scratch_slots * (size_t)a.scratch_stride));
      a.scratch = h->d_scratch;
   }

   *out = h;
   return KINC_OK;
#undef CREATE_TRY
}

static int ensure_staging(kinc_sim* h)
{
   if ( h->d_K ) return KINC_OK;
   const size_t B = (size_t)h->cap;
   HIP_TRY(h, hipMalloc(&h->d_pairs, sizeof(int32_t) * 2 * B));
   HIP_TRY(h, hipMalloc(&h->d_seg, sizeof(int32_t) * B));
   HIP_TRY(h, hipMalloc(&h->d_K, B));
   HIP_TRY(h, hipMalloc(&h->d_labels, B * (size_t)h->S));
   HIP_TRY(h, hipMalloc(&h->d_corr, sizeof(float) * B * (size_t)h->p.max_clusters));
   HIP_TRY(h, hipMalloc(&h->d_iters, sizeof(uint16_t) * B * (size_t)h->p.max_clusters));
   return KINC_OK;
}

static int launch(kinc_sim* h, const void* fn, unsigned grid, unsigned block, void** args,
                  size_t lds, hipStream_t s)
{
   HIP_TRY(h, hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, s));
   return KINC_OK;
}

static int take_event(kinc_sim* h, hipEvent_t* ev)
{
   if ( !h->event_pool.empty() )
   {
      *ev = h->event_pool.back();
      h->event_pool.pop_back();
      return KINC_OK;
   }
   HIP_TRY(h, hipEventCreate(ev));
   return KINC_OK;
}

/* A pair-kernel launch, bracketed by events when profiling is on (P keys the
 * timing: the slot count, or -1 for the one-pair-per-lane Pearson kernel). */
static int launch_pairs(kinc_sim* h, int P, const void* fn, unsigned grid, void** args, hipStream_t s,
                        size_t lds = SIZE_MAX)
{
   if ( lds == SIZE_MAX ) lds = h->lds_bytes;
   if ( !h->profiling ) return launch(h, fn, grid, 64, args, lds, s);
   kinc_sim::Timed t;
   t.P = P;
   int rc = take_event(h, &t.a);
   if ( rc ) return rc;
   rc = take_event(h, &t.b);
   if ( rc ) return rc;
   HIP_TRY(h, hipEventRecord(t.a, s));
   rc = launch(h, fn, grid, 64, args, lds, s);
   if ( rc ) return rc;
   HIP_TRY(h, hipEventRecord(t.b, s));
   h->timed.push_back(t);
   return KINC_OK;
}

extern "C" int kinc_sim_set_profiling(kinc_sim* h, int32_t enabled)
{
   if ( !h ) return KINC_E_INVALID;
   if ( h->capturing && enabled ) return fail_c(h, KINC_E_INVALID, "no kernel timing while a graph is recorded");
   h->profiling = enabled != 0;
   return KINC_OK;
}

extern "C" int kinc_sim_profile_read(kinc_sim* h, int32_t* slots, int64_t* launches,
                                     double* total_ms, int32_t cap)
{
   if ( !h ) return KINC_E_INVALID;
   HIP_TRY(h, hipSetDevice(h->device));
   std::vector<int> ps;
   std::vector<int64_t> ns;
   std::vector<double> ms;
   for ( auto& t : h->timed )
   {
      HIP_TRY(h, hipEventSynchronize(t.b));
      float e = 0.0f;
      HIP_TRY(h, hipEventElapsedTime(&e, t.a, t.b));
      size_t j = 0;
      while ( j < ps.size() && ps[j] != t.P ) ++j;
      if ( j == ps.size() )
      {
         ps.push_back(t.P);
         ns.push_back(0);
         ms.push_back(0.0);
      }
      ns[j] += 1;
      ms[j] += e;
      h->event_pool.push_back(t.a);
      h->event_pool.push_back(t.b);
   }
   h->timed.clear();
   int n = 0;
   for ( size_t j = 0; j < ps.size() && n < cap; ++j, ++n )
   {
      if ( slots ) slots[n] = ps[j];
      if ( launches ) launches[n] = ns[j];
      if ( total_ms ) total_ms[n] = ms[j];
   }
   return n;
}

/* One call: `count` pairs given by a linear start (pairs == null) or by an
 * explicit device pair list with optional work-block segment starts. */
static int run(kinc_sim* h, int64_t start, const int32_t* pairs, const int32_t* seg_lo,
               int cnt, int8_t* K, int8_t* labels, float* corr, uint16_t* iters, hipStream_t s)
{
   SimArgs a = h->base;
   a.block_start = start;
   a.block_count = cnt;
   a.pairs = (const int2*)pairs;
   a.seg_lo = seg_lo;
   a.K_out = K;
   a.labels_out = labels;
   a.corr_out = corr;
   a.iters_out = iters;

   const bool binned = h->p.form == KINC_FORM_PERSISTENT && h->p.clus_method != KINC_CLUS_NONE;
   const bool need_counts = binned || (h->p.corr_method == KINC_CORR_PEARSON);
   if ( need_counts )
   {
      const float* emx = h->d_emx;
      int S = h->S;
      int64_t bs = start;
      const int2* pl = a.pairs;
      float me = h->p.min_expr;
      int32_t* outN = h->d_blockN;
      int G = h->G;
      float* xb = h->d_xbuf;
      void* args[] = {&emx, &S, &G, &bs, &pl, (void*)&cnt, &me, &outN, &xb};
      int rc = launch(h, kinc_kernel_count_valid(), (unsigned)((cnt + 3) / 4), 256, args, 0, s);
      if ( rc ) return rc;
   }
   if ( h->p.corr_method == KINC_CORR_PEARSON )
   {
      const int32_t* Np = h->d_blockN;
      int32_t* cm = h->d_chunkMax;
      void* args[] = {&Np, (void*)&cnt, &cm};
      const int chunks = (cnt + 63) / 64;
      int rc = launch(h, kinc_kernel_chunk_max(), (unsigned)((chunks + 3) / 4), 256, args, 0, s);
      if ( rc ) return rc;
      if ( h->d_prev )
      {
         const int32_t* sl = seg_lo;
         int32_t* pv = h->d_prev;
         void* pargs[] = {&Np, &cm, &sl, (void*)&cnt, &pv};
         rc = launch(h, kinc_kernel_prev_greater(), (unsigned)((cnt + 3) / 4), 256, pargs, 0, s);
         if ( rc ) return rc;
      }
   }

   if ( h->p.form == KINC_FORM_PLAIN )
   {
      /* the slot count S needs (or the memory store) for every pair */
      const int P = h->classes[h->nbins - 1];
      const void* fn = pair_kernel(h, P);
      if ( !fn ) return fail_c(h, KINC_E_UNSUPPORTED, "no kernel instance for this shape");
      a.list = nullptr;
      for ( int first = 0; first < cnt; first += h->p.ksize )
      {
         a.first = first;
         a.launch_count = std::min(h->p.ksize, cnt - first);
         unsigned grid = (unsigned)a.launch_count;
         if ( P == 0 ) grid = (unsigned)std::min(a.launch_count, h->scratch_slots);
         void* args[] = {&a};
         int rc = launch_pairs(h, P, fn, grid, args, s);
         if ( rc ) return rc;
      }
      return KINC_OK;
   }

   if ( h->p.clus_method == KINC_CLUS_NONE && h->p.corr_method == KINC_CORR_PEARSON && h->d_xbuf )
   {
      /* no fits, Pearson: one pair per lane over the whole block */
      a.list = nullptr;
      a.first = 0;
      a.launch_count = cnt;
      void* args[] = {&a};
      return launch_pairs(h, -1, kinc_kernel_pearson_lanes(), (unsigned)((cnt + 63) / 64), args, s, 0);
   }
   if ( h->p.clus_method == KINC_CLUS_NONE )
   {
      /* no fits: the slot count does not matter and every pair costs about
       * the same, so resident waves stride over the whole block in one launch */
      const int P = h->classes[0];
      const void* fn = pair_kernel(h, P);
      if ( !fn ) return fail_c(h, KINC_E_UNSUPPORTED, "no kernel instance for this shape");
      int per_cu = 0;
      HIP_TRY(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, h->lds_bytes));
      if ( per_cu < 1 ) per_cu = 1;
      long long waves = (long long)per_cu * h->num_cus;
      if ( P == 0 ) waves = std::min<long long>(waves, h->scratch_slots);
      a.list = nullptr;
      a.first = 0;
      a.launch_count = cnt;
      void* args[] = {&a};
      return launch_pairs(h, P, fn, (unsigned)std::min<long long>(waves, (long long)cnt), args, s);
   }

   /* persistent: bin by slot count, then drain each bin with resident waves */
   {
      int32_t* ctr = h->d_counters;
      int n = 2 * kMaxBins;
      void* args[] = {&ctr, &n};
      int rc = launch(h, kinc_kernel_zero_i32(), 1, 64, args, 0, s);
      if ( rc ) return rc;
   }
   {
      const int32_t* Np = h->d_blockN;
      BinBounds bb = h->bins;
      int nb = h->nbins;
      int32_t* ctr = h->d_counters;
      int32_t* lists = h->d_lists;
      void* args[] = {&Np, (void*)&cnt, &bb, &nb, &ctr, &lists};
      int rc = launch(h, kinc_kernel_bin_pairs(), (unsigned)((cnt + 255) / 256), 256, args, 0, s);
      if ( rc ) return rc;
   }
   /* Bins run concurrently on the handle's auxiliary streams: a bin with few
    * pairs would otherwise leave most of the device idle for one pair's
    * latency (milliseconds).  They wait for the binning on s, and s waits
    * for all of them. */
   const int naux = std::max(1, std::min(h->naux, (int)kinc_sim::kAux));
   HIP_TRY(h, hipEventRecord(h->binned, s));
   for ( int i = 0; i < naux; ++i ) HIP_TRY(h, hipStreamWaitEvent(h->aux[i], h->binned, 0));
   for ( int b = h->nbins - 1; b >= 0; --b )
   {
      const int P = h->classes[b];
      hipStream_t sb = h->aux[(h->nbins - 1 - b) % naux];
      const void* fn = pair_kernel(h, P);
      if ( !fn ) return fail_c(h, KINC_E_UNSUPPORTED, "no kernel instance for this shape");
      int per_cu = 0;
      HIP_TRY(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, h->lds_bytes));
      if ( per_cu < 1 ) per_cu = 1;
      long long waves = (long long)per_cu * h->num_cus;
      if ( P == 0 ) waves = std::min<long long>(waves, h->scratch_slots);
      long long grid = std::min<long long>(waves, (long long)cnt);
      a.list = h->d_lists + (size_t)b * (size_t)cnt;
      a.list_len = h->d_counters + b;
      a.list_head = h->d_counters + kMaxBins + b;
      void* args[] = {&a};
      int rc = launch_pairs(h, P, fn, (unsigned)grid, args, sb);
      if ( rc ) return rc;
   }
   for ( int i = 0; i < naux; ++i )
   {
      HIP_TRY(h, hipEventRecord(h->drained[i], h->aux[i]));
      HIP_TRY(h, hipStreamWaitEvent(s, h->drained[i], 0));
   }
   return KINC_OK;
}

extern "C" int kinc_sim_execute_device(kinc_sim* h, int64_t start, int64_t count,
                                       int8_t* K, int8_t* labels, float* corr,
                                       uint16_t* iters, void* stream)
{
   if ( !h ) return KINC_E_INVALID;
   if ( count < 0 || count > KINC_WORK_BLOCK_SIZE || count > h->cap )
   {
      return fail(h, KINC_E_INVALID, "count must lie in [0, min(32768, batch_capacity)]");
   }
   if ( start < 0 || start + count > kinc_pair_count(h->G) ) return fail(h, KINC_E_INVALID, "pair range outside the matrix");
   if ( count == 0 ) return KINC_OK;
   if ( !K || !corr ) return fail(h, KINC_E_INVALID, "K and corr outputs are required");
   if ( h->capturing && stream && (hipStream_t)stream != h->stream )
   {
      return fail_c(h, KINC_E_INVALID, "while a graph is recorded, work goes on the handle's own stream");
   }
   HIP_TRY(h, hipSetDevice(h->device));
   hipStream_t s = stream ? (hipStream_t)stream : h->stream;
   return run(h, start, nullptr, nullptr, (int)count, K, labels, corr, iters, s);
}

extern "C" int kinc_sim_execute_pairs_device(kinc_sim* h, const int32_t* pairs, const int32_t* seg_lo,
                                             int64_t count, int8_t* K, int8_t* labels, float* corr,
                                             uint16_t* iters, void* stream)
{
   if ( !h ) return KINC_E_INVALID;
   if ( count < 0 || count > h->cap ) return fail(h, KINC_E_INVALID, "count must lie in [0, batch_capacity]");
   if ( count == 0 ) return KINC_OK;
   if ( !pairs || !K || !corr ) return fail(h, KINC_E_INVALID, "pairs, K and corr are required");
   if ( h->capturing && stream && (hipStream_t)stream != h->stream )
   {
      return fail_c(h, KINC_E_INVALID, "while a graph is recorded, work goes on the handle's own stream");
   }
   HIP_TRY(h, hipSetDevice(h->device));
   hipStream_t s = stream ? (hipStream_t)stream : h->stream;
   return run(h, 0, pairs, seg_lo, (int)count, K, labels, corr, iters, s);
}

extern "C" int kinc_sim_execute_pairs(kinc_sim* h, const int32_t* pairs, const int32_t* seg_lo,
                                      int64_t count, int8_t* K, int8_t* labels, float* corr,
                                      uint16_t* iters)
{
   if ( h && h->capturing ) return fail_c(h, KINC_E_INVALID, "host-buffer calls cannot be recorded into a graph");
   if ( !h ) return KINC_E_INVALID;
   if ( count < 0 || count > h->cap ) return fail(h, KINC_E_INVALID, "count must lie in [0, batch_capacity]");
   if ( count == 0 ) return KINC_OK;
   if ( !pairs || !K || !corr ) return fail(h, KINC_E_INVALID, "pairs, K and corr are required");
   for ( int64_t q = 0; q < count; ++q )
   {
      const int32_t x = pairs[2 * q], y = pairs[2 * q + 1];
      if ( !(0 <= y && y < x && x < h->G) ) return fail(h, KINC_E_INVALID, "pair " + std::to_string(q) + " is not x > y >= 0 within the matrix");
      if ( seg_lo && (seg_lo[q] < 0 || seg_lo[q] > q) ) return fail(h, KINC_E_INVALID, "seg_lo[q] must lie in [0, q]");
   }
   int rc = ensure_staging(h);
   if ( rc ) return rc;
   HIP_TRY(h, hipSetDevice(h->device));
   HIP_TRY(h, hipMemcpyAsync(h->d_pairs, pairs, sizeof(int32_t) * 2 * (size_t)count, hipMemcpyHostToDevice, h->stream));
   if ( seg_lo ) HIP_TRY(h, hipMemcpyAsync(h->d_seg, seg_lo, sizeof(int32_t) * (size_t)count, hipMemcpyHostToDevice, h->stream));
   rc = run(h, 0, h->d_pairs, seg_lo ? h->d_seg : nullptr, (int)count, h->d_K, labels ? h->d_labels : nullptr,
            h->d_corr, iters ? h->d_iters : nullptr, h->stream);
   if ( rc ) return rc;
   const size_t n = (size_t)count;
   const size_t MC = (size_t)h->p.max_clusters;
   HIP_TRY(h, hipMemcpyAsync(K, h->d_K, n, hipMemcpyDeviceToHost, h->stream));
   HIP_TRY(h, hipMemcpyAsync(corr, h->d_corr, sizeof(float) * n * MC, hipMemcpyDeviceToHost, h->stream));
   if ( labels ) HIP_TRY(h, hipMemcpyAsync(labels, h->d_labels, n * (size_t)h->S, hipMemcpyDeviceToHost, h->stream));
   if ( iters ) HIP_TRY(h, hipMemcpyAsync(iters, h->d_iters, sizeof(uint16_t) * n * MC, hipMemcpyDeviceToHost, h->stream));
   HIP_TRY(h, hipStreamSynchronize(h->stream));
   return KINC_OK;
}

extern "C" int kinc_sim_execute(kinc_sim* h, int64_t start, int64_t count,
                                int8_t* K, int8_t* labels, float* corr, uint16_t* iters)
{
   if ( h && h->capturing ) return fail_c(h, KINC_E_INVALID, "host-buffer calls cannot be recorded into a graph");
   if ( !h ) return KINC_E_INVALID;
   if ( count == 0 ) return KINC_OK;
   if ( !K || !corr ) return fail(h, KINC_E_INVALID, "K and corr outputs are required");
   int rc = ensure_staging(h);
   if ( rc ) return rc;
   rc = kinc_sim_execute_device(h, start, count, h->d_K, labels ? h->d_labels : nullptr, h->d_corr,
                                iters ? h->d_iters : nullptr, nullptr);
   if ( rc ) return rc;
   const size_t n = (size_t)count;
   const size_t MC = (size_t)h->p.max_clusters;
   HIP_TRY(h, hipMemcpyAsync(K, h->d_K, n, hipMemcpyDeviceToHost, h->stream));
   HIP_TRY(h, hipMemcpyAsync(corr, h->d_corr, sizeof(float) * n * MC, hipMemcpyDeviceToHost, h->stream));
   if ( labels ) HIP_TRY(h, hipMemcpyAsync(labels, h->d_labels, n * (size_t)h->S, hipMemcpyDeviceToHost, h->stream));
   if ( iters ) HIP_TRY(h, hipMemcpyAsync(iters, h->d_iters, sizeof(uint16_t) * n * MC, hipMemcpyDeviceToHost, h->stream));
   HIP_TRY(h, hipStreamSynchronize(h->stream));
   return KINC_OK;
}

extern "C" int kinc_sim_synchronize(kinc_sim* h)
{
   if ( !h ) return KINC_E_INVALID;
   HIP_TRY(h, hipSetDevice(h->device));
   for ( int i = 0; i < kinc_sim::kAux; ++i ) HIP_TRY(h, hipStreamSynchronize(h->aux[i]));
   HIP_TRY(h, hipStreamSynchronize(h->stream));
   return KINC_OK;
}

extern "C" void kinc_sim_destroy(kinc_sim* h)
{
   if ( !h ) return;
   (void)hipSetDevice(h->device);
   if ( h->stream ) (void)hipStreamSynchronize(h->stream);
   (void)hipFree(h->d_emx);
   (void)hipFree(h->d_logN);
   (void)hipFree(h->d_blockN);
   (void)hipFree(h->d_chunkMax);
   (void)hipFree(h->d_xbuf);
   (void)hipFree(h->d_prev);
   (void)hipFree(h->d_counters);
   (void)hipFree(h->d_lists);
   (void)hipFree(h->d_scratch);
   (void)hipFree(h->d_pairs);
   (void)hipFree(h->d_seg);
   (void)hipFree(h->d_K);
   (void)hipFree(h->d_labels);
   (void)hipFree(h->d_corr);
   (void)hipFree(h->d_iters);
   for ( auto& t : h->timed )
   {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
   }
   for ( auto e : h->event_pool ) (void)hipEventDestroy(e);
   for ( int i = 0; i < kinc_sim::kAux; ++i )
   {
      if ( h->aux[i] ) (void)hipStreamDestroy(h->aux[i]);
      if ( h->drained[i] ) (void)hipEventDestroy(h->drained[i]);
   }
   if ( h->binned ) (void)hipEventDestroy(h->binned);
   if ( h->stream ) (void)hipStreamDestroy(h->stream);
   delete h;
}

extern "C" const char* kinc_sim_last_error(const kinc_sim* h)
{
   return h ? h->err : g_create_error;
}

/* ------------------------------------------------------------- autotuner */

/*
 * The launch configurations a handle can run with and that give the same
 * results: the waves per SIMD the register-slot bins' builds are made for
 * (4; 5 for the small instances where the LDS leaves room for 20 waves per
 * CU; 3 where it allows fewer than 16), how many auxiliary streams the
 * persistent form's bins are spread over, and whether the GMM fits with 4
 * and 5 clusters run in lockstep (where the LDS layout holds their tile).
 * kinc_sim_autotune_group times each on a sample of pairs, run by all the
 * handles of a group at once, and keeps the fastest; kinc_sim_autotune is
 * the group of one.  Timed alone, one handle of a two-handle run preferred
 * two auxiliary streams where the pair runs 5% faster with four, so the
 * handles that will share the device are timed together.
 */

/* Time one configuration on a group of handles: every handle runs the
 * sample on its own stream, all starting together, and the time is that of
 * the whole group (the worker handles of one run share the device, so a
 * configuration is judged as they will use it). */
static int time_config(kinc_sim* const* hs, int n, int64_t start, int count, int waves, int naux,
                       bool j45, double* ms)
{
   kinc_sim* h = hs[0];
   for ( int i = 0; i < n; ++i ) apply_config(hs[i], waves, naux, j45);
   std::vector<hipEvent_t> ev((size_t)n + 1, nullptr);
   int rc = KINC_OK;
   for ( size_t i = 0; i < ev.size() && rc == KINC_OK; ++i )
   {
      if ( hipEventCreate(&ev[i]) != hipSuccess ) rc = fail_c(h, KINC_E_DEVICE, "hipEventCreate failed");
   }
   double best = 1e30;
   for ( int rep = 0; rep < 3 && rc == KINC_OK; ++rep )
   {
      /* the first run warms up (code objects, caches); the best of the others counts */
      if ( hipEventRecord(ev[0], h->stream) != hipSuccess ) rc = fail_c(h, KINC_E_DEVICE, "hipEventRecord failed");
      for ( int i = 1; i < n && rc == KINC_OK; ++i )
      {
         if ( hipStreamWaitEvent(hs[i]->stream, ev[0], 0) != hipSuccess ) rc = fail_c(h, KINC_E_DEVICE, "hipStreamWaitEvent failed");
      }
      for ( int i = 0; i < n && rc == KINC_OK; ++i )
      {
         rc = run(hs[i], start, nullptr, nullptr, count, hs[i]->d_K, nullptr, hs[i]->d_corr, nullptr, hs[i]->stream);
         if ( rc && i > 0 ) std::snprintf(h->err, sizeof(h->err), "%s", hs[i]->err);
      }
      for ( int i = 1; i < n && rc == KINC_OK; ++i )
      {
         if ( hipEventRecord(ev[i], hs[i]->stream) != hipSuccess ||
              hipStreamWaitEvent(h->stream, ev[i], 0) != hipSuccess )
         {
            rc = fail_c(h, KINC_E_DEVICE, "joining the group's streams failed");
         }
      }
      if ( rc == KINC_OK && hipEventRecord(ev[n], h->stream) != hipSuccess ) rc = fail_c(h, KINC_E_DEVICE, "hipEventRecord failed");
      if ( rc == KINC_OK && hipEventSynchronize(ev[n]) != hipSuccess ) rc = fail_c(h, KINC_E_DEVICE, "hipEventSynchronize failed");
      float t = 0.0f;
      if ( rc == KINC_OK && hipEventElapsedTime(&t, ev[0], ev[n]) != hipSuccess ) rc = fail_c(h, KINC_E_DEVICE, "hipEventElapsedTime failed");
      if ( rc == KINC_OK && rep > 0 ) best = std::min(best, (double)t);
   }
   for ( hipEvent_t e : ev )
   {
      if ( e ) (void)hipEventDestroy(e);
   }
   *ms = best;
   return rc;
}

/* handles one configuration means the same on: device, matrix shape and
 * every parameter (kinc_sim_params holds 4-byte fields only) */
static bool same_setup(const kinc_sim* a, const kinc_sim* b)
{
   return a->device == b->device && a->G == b->G && a->S == b->S &&
          std::memcmp(&a->p, &b->p, sizeof(a->p)) == 0;
}

extern "C" int kinc_sim_autotune_group(kinc_sim* const* hs, int32_t n, int64_t start, int64_t count,
                                       kinc_sim_tuning* out)
{
   if ( !hs || n < 1 || n > KINC_AUTOTUNE_MAX_HANDLES ) return fail(nullptr, KINC_E_INVALID, "autotune: 1..16 handles");
   for ( int i = 0; i < n; ++i )
   {
      if ( !hs[i] ) return fail(nullptr, KINC_E_INVALID, "autotune: null handle");
   }
   kinc_sim* h = hs[0];
   for ( int i = 0; i < n; ++i )
   {
      for ( int j = 0; j < i; ++j )
      {
         if ( hs[j] == hs[i] ) return fail_c(h, KINC_E_INVALID, "autotune: a handle appears twice");
      }
      if ( !same_setup(h, hs[i]) ) return fail_c(h, KINC_E_INVALID, "autotune: the handles differ in device, matrix shape or parameters");
      if ( hs[i]->capturing ) return fail_c(h, KINC_E_INVALID, "autotune: a handle is recording a graph");
   }
   if ( count < 1 || count > KINC_WORK_BLOCK_SIZE || count > h->cap )
   {
      return fail_c(h, KINC_E_INVALID, "autotune sample must hold 1..min(32768, batch_capacity) pairs");
   }
   if ( start < 0 || start + count > kinc_pair_count(h->G) ) return fail_c(h, KINC_E_INVALID, "pair range outside the matrix");
   for ( int i = 0; i < n; ++i )
   {
      const int rc = ensure_staging(hs[i]);
      if ( rc )
      {
         if ( i > 0 ) std::snprintf(h->err, sizeof(h->err), "%s", hs[i]->err);
         return rc;
      }
   }
   HIP_TRY(h, hipSetDevice(h->device));
   /* each handle's configuration, restored if the tuning fails */
   struct Saved { int waves, naux; bool j45; };
   std::vector<Saved> saved((size_t)n);
   for ( int i = 0; i < n; ++i ) saved[(size_t)i] = {hs[i]->waves, hs[i]->naux, hs[i]->base.joint45 != 0};
   /* candidates; the other builds and the lockstep fits only where the LDS
    * leaves room for them */
   const int wavess[3] = {3, 4, 5};
   const int nauxs[3] = {1, 2, 4};
   int best_waves = saved[0].waves, best_naux = saved[0].naux;
   bool best_j45 = saved[0].j45;
   double best_ms = 1e30;
   int tried = 0;
   int rc = KINC_OK;
   std::vector<int8_t> K0((size_t)count), K1((size_t)count);
   std::vector<float> C0((size_t)count * (size_t)h->p.max_clusters), C1(C0.size());
   for ( int c = 0; c < 18 && rc == KINC_OK; ++c )
   {
      const int waves = wavess[c / 6];
      const bool j45 = ((c / 3) & 1) != 0;
      const int naux = nauxs[c % 3];
      if ( (j45 && h->lds_j45 == 0) || !waves_allowed(h, waves, j45) ) continue;
      double ms = 0.0;
      rc = time_config(hs, n, start, (int)count, waves, naux, j45, &ms);
      if ( rc ) break;
      /* every configuration and every handle must give the same bits */
      for ( int i = 0; i < n && rc == KINC_OK; ++i )
      {
         const bool first = tried == 0 && i == 0;
         std::vector<int8_t>& K = first ? K0 : K1;
         std::vector<float>& C = first ? C0 : C1;
         if ( hipMemcpy(K.data(), hs[i]->d_K, (size_t)count, hipMemcpyDeviceToHost) != hipSuccess ||
              hipMemcpy(C.data(), hs[i]->d_corr, sizeof(float) * C.size(), hipMemcpyDeviceToHost) != hipSuccess )
         {
            rc = fail_c(h, KINC_E_DEVICE, "autotune: copying the sample's results failed");
         }
         else if ( !first && (std::memcmp(K0.data(), K1.data(), K0.size()) != 0 ||
                              std::memcmp(C0.data(), C1.data(), sizeof(float) * C0.size()) != 0) )
         {
            rc = fail_c(h, KINC_E_DEVICE, "autotune: launch configurations or handles gave different results");
         }
      }
      if ( rc ) break;
      ++tried;
      if ( ms < best_ms )
      {
         best_ms = ms;
         best_waves = waves;
         best_naux = naux;
         best_j45 = j45;
      }
   }
   if ( rc )
   {
      for ( int i = 0; i < n; ++i )
      {
         apply_config(hs[i], saved[(size_t)i].waves, saved[(size_t)i].naux, saved[(size_t)i].j45);
      }
      return rc;
   }
   for ( int i = 0; i < n; ++i ) apply_config(hs[i], best_waves, best_naux, best_j45);
   if ( out )
   {
      out->pair_waves = h->waves;
      out->aux_streams = h->naux;
      out->joint45 = h->base.joint45;
      out->candidates = tried;
      out->sample_ms = best_ms;
   }
   return KINC_OK;
}

extern "C" int kinc_sim_autotune(kinc_sim* h, int64_t start, int64_t count, kinc_sim_tuning* out)
{
   if ( !h ) return KINC_E_INVALID;
   return kinc_sim_autotune_group(&h, 1, start, count, out);
}

extern "C" int kinc_selftest_census(kinc_sim* h, uint64_t* d_counters)
{
   if ( !h ) return KINC_E_INVALID;
   if ( h->capturing ) return fail_c(h, KINC_E_INVALID, "the census cannot change while a graph is captured");
   if ( d_counters )
   {
      bool has7 = false;
      for ( int b = 0; b < h->nbins; ++b ) has7 = has7 || h->classes[b] == 7;
      if ( h->kmax != 5 || !has7 ) return fail_c(h, KINC_E_UNSUPPORTED, "the census build covers the 7-slot, 5-cluster instance only");
   }
   h->base.census = (unsigned long long*)d_counters;
   return KINC_OK;
}

extern "C" int kinc_sim_get_tuning(const kinc_sim* h, kinc_sim_tuning* out)
{
   if ( !h || !out ) return KINC_E_INVALID;
   out->pair_waves = h->waves;
   out->aux_streams = h->naux;
   out->joint45 = h->base.joint45;
   out->candidates = 0;
   out->sample_ms = 0.0;
   return KINC_OK;
}

extern "C" int kinc_sim_set_tuning(kinc_sim* h, const kinc_sim_tuning* t)
{
   if ( !h || !t ) return KINC_E_INVALID;
   if ( t->aux_streams < 1 || t->aux_streams > kinc_sim::kAux ) return fail_c(h, KINC_E_INVALID, "aux_streams must lie in [1, 4]");
   if ( t->pair_waves < 3 || t->pair_waves > 5 ) return fail_c(h, KINC_E_INVALID, "pair_waves must be 3, 4 or 5");
   if ( t->joint45 && h->lds_j45 == 0 ) return fail_c(h, KINC_E_UNSUPPORTED, "the LDS layout holds no tile for the lockstep 4- and 5-cluster fits");
   if ( !waves_allowed(h, t->pair_waves, t->joint45 != 0) )
   {
      return fail_c(h, KINC_E_UNSUPPORTED, "the per-wave LDS makes no use of builds for that many waves per SIMD");
   }
   apply_config(h, t->pair_waves, t->aux_streams, t->joint45 != 0);
   return KINC_OK;
}

/* ------------------------------------------------------------ HIP graphs */

struct kinc_sim_graph
{
   hipGraph_t graph = nullptr;
   hipGraphExec_t exec = nullptr;
   int device = 0;
};

extern "C" int kinc_sim_graph_begin(kinc_sim* h)
{
   if ( !h ) return KINC_E_INVALID;
   if ( h->capturing ) return fail_c(h, KINC_E_INVALID, "a graph is already being recorded");
   if ( h->profiling ) return fail_c(h, KINC_E_INVALID, "no graph recording while kernel timing is on");
   HIP_TRY(h, hipSetDevice(h->device));
   HIP_TRY(h, hipStreamSynchronize(h->stream));
   HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
   h->capturing = true;
   return KINC_OK;
}

extern "C" int kinc_sim_graph_end(kinc_sim* h, kinc_sim_graph** out)
{
   if ( !h || !out ) return KINC_E_INVALID;
   if ( !h->capturing ) return fail_c(h, KINC_E_INVALID, "no graph is being recorded");
   *out = nullptr;
   h->capturing = false;
   kinc_sim_graph* g = new kinc_sim_graph();
   g->device = h->device;
   hipError_t e = hipStreamEndCapture(h->stream, &g->graph);
   if ( e == hipSuccess ) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
   if ( e != hipSuccess )
   {
      kinc_sim_graph_destroy(g);
      return fail_hip(h, "recording the graph", e);
   }
   *out = g;
   return KINC_OK;
}

extern "C" int kinc_sim_graph_launch(kinc_sim_graph* g, void* stream)
{
   if ( !g || !g->exec ) return KINC_E_INVALID;
   if ( hipSetDevice(g->device) != hipSuccess ) return KINC_E_DEVICE;
   return hipGraphLaunch(g->exec, (hipStream_t)stream) == hipSuccess ? KINC_OK : KINC_E_DEVICE;
}

extern "C" void kinc_sim_graph_destroy(kinc_sim_graph* g)
{
   if ( !g ) return;
   (void)hipSetDevice(g->device);
   if ( g->exec ) (void)hipGraphExecDestroy(g->exec);
   if ( g->graph ) (void)hipGraphDestroy(g->graph);
   delete g;
}
