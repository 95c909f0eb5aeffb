/*
 * kinc_sim.h — C-ABI of the MI355X pairwise Similarity engine.
 *
 * This is the drop-in boundary for KINC's Similarity analytic.  It replaces
 * the per-block backend interface of the reference:
 *   - Similarity::Serial::execute(const Block*) -> unique_ptr<Block>
 *       (src/core/similarity_serial.h:12, src/core/similarity_serial.cpp:33-101)
 *   - Similarity::OpenCL::Worker::execute(const Block*) -> unique_ptr<Block>
 *       (src/core/similarity_opencl_worker.h:12, .cpp:93-271)
 *   - Similarity::OpenCL::initialize(Context*) which uploads the expression
 *       matrix once (src/core/similarity_opencl.cpp:34-70)
 * A work block [start, start+count) of linear pair indices goes in; per pair
 * the selected cluster count K, the per-sample labels and the per-cluster
 * correlations come out, exactly the fields of Similarity::Pair
 * (src/core/similarity.h:19-24).  Output buffers are caller-owned with fixed
 * strides instead of a heap-allocated ResultBlock.  No C++ exceptions cross
 * this boundary: every entry point returns 0 on success or a negative code,
 * and kinc_sim_last_error() gives the message.
 *
 * One handle is bound to one device and one HIP stream and keeps the
 * expression matrix resident in device memory.  A handle is not thread-safe:
 * use one handle per host thread.
 */
#ifndef KINC_SIM_H
#define KINC_SIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KINC_OK              0
#define KINC_E_INVALID      -1   /* bad argument or parameter combination */
#define KINC_E_DEVICE       -2   /* HIP runtime error */
#define KINC_E_NOMEM        -3   /* device or host allocation failed */
#define KINC_E_UNSUPPORTED  -4   /* valid in the reference, not built here */

/* --clusmethod none|gmm|kmeans (src/core/similarity.h:39-44) */
#define KINC_CLUS_NONE    0
#define KINC_CLUS_GMM     1
#define KINC_CLUS_KMEANS  2
/* --corrmethod pearson|spearman (src/core/similarity.h:46-50) */
#define KINC_CORR_PEARSON   0
#define KINC_CORR_SPEARMAN  1
/* --crit BIC|ICL (src/core/pairwise_clustering.h:9-13) */
#define KINC_CRIT_BIC  0
#define KINC_CRIT_ICL  1
/* execution form of the device path */
#define KINC_FORM_PLAIN       0  /* one launch per ksize sub-batch, one wave per pair */
#define KINC_FORM_PERSISTENT  1  /* sample-count-binned queue drained by resident waves */

/* Largest cluster count a pair may have (Pairwise::Index::MAX_CLUSTER_SIZE,
 * src/core/pairwise_index.h:36). */
#define KINC_MAX_CLUSTER_SIZE 64
/* Pairs per work block (Similarity::size/makeWork, src/core/similarity.cpp:17-39). */
#define KINC_WORK_BLOCK_SIZE 32768

/*
 * Analytic parameters.  Field meanings and defaults follow the reference's
 * Similarity members (src/core/similarity.h:55-71) and CLI flags
 * (src/core/similarity_input.cpp:96-254).
 */
typedef struct kinc_sim_params
{
   int32_t clus_method;    /* --clusmethod, default none */
   int32_t corr_method;    /* --corrmethod, default pearson */
   float   min_expr;       /* --minexpr, default -inf */
   int32_t min_samples;    /* --minsamp, default 30 */
   int32_t min_clusters;   /* --minclus, default 1 */
   int32_t max_clusters;   /* --maxclus, default 5 */
   int32_t criterion;      /* --crit, default ICL */
   int32_t pre_out;        /* --preout, default 0 */
   int32_t post_out;       /* --postout, default 0 */
   float   min_corr;       /* --mincorr, default 0.5 (used by the CCM/CMX writer) */
   float   max_corr;       /* --maxcorr, default 1.0 (used by the CCM/CMX writer) */
   int32_t ksize;          /* --ksize, pairs per plain-form launch, default 4096 */
   int32_t form;           /* KINC_FORM_*, default persistent */
   int32_t batch_capacity; /* largest count of one execute call; 0 = 32768 */
   int32_t pearson_fp64;   /* 0 (default): Pearson moments summed as the reference
                              sums them (plain FP32, pairwise_pearson.cpp:26-48);
                              1: the same sums in FP64 over exact products and the
                              quotient in FP64, rounded to float: an extension for
                              callers who want correlations within half an ulp of
                              the exact value (K, labels and Spearman unchanged) */
   int32_t reserved[3];    /* must be zero */
} kinc_sim_params;

typedef struct kinc_sim kinc_sim;

/* Fill *p with the reference defaults. */
void kinc_sim_default_params(kinc_sim_params* p);

/*
 * Create a handle on `device`, upload the expression matrix emx[G][S]
 * (row-major float32, NaN = missing; the EMX payload layout of
 * src/core/expressionmatrix.cpp:302-312) and keep it resident.
 */
int kinc_sim_create(const kinc_sim_params* params, const float* emx,
                    int32_t gene_size, int32_t sample_size, int32_t device,
                    kinc_sim** out);

/*
 * Execute one work block of `count` pairs starting at linear pair index
 * `start` (pair k = x(x-1)/2 + y with x > y, src/core/pairwise_index.cpp:45-56).
 * Host output buffers, written for every pair of the block:
 *   K      [count]                   selected cluster count (0 = too few samples)
 *   labels [count][sample_size]      per-sample labels (-9 NaN, -6 below minexpr,
 *                                    -7/-8 pre/post outlier, >= 0 cluster);
 *                                    may be NULL
 *   corr   [count][max_clusters]     correlation of cluster k for k < K, NaN after
 *   iters  [count][max_clusters]     optional (may be NULL): for the fit with k+1
 *                                    clusters, the EM or Lloyd iteration count in
 *                                    bits 0-11, bit 12 set if EM converged, bit 13
 *                                    set if the fit failed; 0 if not run
 * A block is the unit of the reference's stale-buffer semantics: the X
 * workspace is shared by the pairs of one call in index order
 * (src/core/similarity_serial.cpp:42), so results depend on the block split
 * exactly as they do in the reference.  count <= KINC_WORK_BLOCK_SIZE (and
 * <= batch_capacity).
 */
int kinc_sim_execute(kinc_sim* h, int64_t start, int64_t count,
                     int8_t* K, int8_t* labels, float* corr, uint16_t* iters);

/*
 * Same as kinc_sim_execute but the outputs are device pointers on the
 * handle's device and the call is asynchronous on `stream` (a hipStream_t,
 * or NULL for the handle's own stream): the work is ordered after what is
 * already queued on `stream`, and what is queued on `stream` afterwards
 * runs after it.  labels and iters may be NULL.  Calls on one handle share
 * its per-block scratch, so they must be issued on one stream (or otherwise
 * ordered).
 */
int kinc_sim_execute_device(kinc_sim* h, int64_t start, int64_t count,
                            int8_t* K, int8_t* labels, float* corr,
                            uint16_t* iters, void* stream);

/*
 * Execute an explicit list of pairs: pairs[2*q], pairs[2*q+1] are the genes
 * (x, y) of pair q, x > y.  seg_lo[q] (optional, NULL = all 0) is the index of
 * the first pair of q's work block within this call: the pairs of one work
 * block share the reference's X workspace (stale-read Pearson semantics, see
 * kinc_sim_execute), blocks do not.  Several requests' work blocks can so be
 * batched into one call with results identical to running each block alone.
 * count <= batch_capacity.  Host buffers in, host buffers out.
 */
int kinc_sim_execute_pairs(kinc_sim* h, const int32_t* pairs, const int32_t* seg_lo,
                           int64_t count, int8_t* K, int8_t* labels, float* corr,
                           uint16_t* iters);

/* Same with device buffers (pairs as int32 x,y interleaved), asynchronous on
 * `stream` like kinc_sim_execute_device.  The device lists are not checked
 * on the host: a pair outside x > y >= 0, x < G reads gene 0 twice and its
 * results are meaningless, and a seg_lo[q] outside [0, q] is taken as the
 * nearest end of that range; neither reads outside the handle's buffers. */
int kinc_sim_execute_pairs_device(kinc_sim* h, const int32_t* pairs, const int32_t* seg_lo,
                                  int64_t count, int8_t* K, int8_t* labels, float* corr,
                                  uint16_t* iters, void* stream);

/* Wait for all work queued by the handle. */
int kinc_sim_synchronize(kinc_sim* h);

/* Release the handle and its device memory. */
void kinc_sim_destroy(kinc_sim* h);

/* Message of the last failed call on this handle (or of the last failed
 * kinc_sim_create when h is NULL). */
const char* kinc_sim_last_error(const kinc_sim* h);

/*
 * Kernel timing (instrumentation; the reference has no counterpart).  While
 * enabled, every pair-kernel launch of the handle is bracketed by HIP events
 * on the stream it runs on.  kinc_sim_profile_read waits for the handle's
 * work, then reports, per register-slot class P (64*P samples per wave),
 * the number of launches and their summed duration in milliseconds, and
 * clears the record.  Returns the number of classes written (<= cap).
 */
int kinc_sim_set_profiling(kinc_sim* h, int32_t enabled);
int kinc_sim_profile_read(kinc_sim* h, int32_t* slots, int64_t* launches,
                          double* total_ms, int32_t cap);

/*
 * Launch configuration of a handle (instrumentation and tuning; the
 * reference's only performance knob is --ksize, src/core/similarity_input.cpp:
 * 241-251).  Every configuration gives the same results.
 *   pair_waves   waves per SIMD the register-slot bins' kernel builds are
 *                made for: 4; 5 (the instances of at most 6 slots, only when
 *                the per-wave LDS leaves room for 20 waves per CU); 3 (the
 *                instances of 5 slots or more, only when the per-wave LDS
 *                allows fewer than 16 waves per CU: 168 VGPRs per wave
 *                instead of 128, at the same residency where the LDS allows
 *                at most 12).  A handle starts with 5 where it fits, else 3
 *                where it is allowed, else 4
 *   aux_streams  streams the persistent form's bins run on concurrently (1-4)
 *   joint45      1: the GMM fits with 4 and 5 clusters run in lockstep on the
 *                register-slot instances whose VGPR budget holds them (at
 *                most 7 slots at 4 waves per SIMD, 10 at 3), sharing
 *                their chain-sum passes; their 29-row tile lies over the
 *                compact X and the shared LDS region, enlarging that region
 *                where the two are smaller (offered while 8 waves per CU
 *                stay resident); 0: one after the other.  A handle starts
 *                with 1 where it is offered.
 * kinc_sim_autotune runs the sample of pairs [start, start+count) (at most
 * one work block) with every candidate configuration, three times each,
 * checks that all of them give identical K and correlations, keeps the
 * fastest on the handle and reports it (candidates tried, best time of the
 * sample in ms).  Results of later calls are unchanged by it.
 * kinc_sim_autotune_group does the same for n handles (1..16) created with
 * the same device, matrix and parameters, such as the worker handles of one
 * run: each candidate runs the sample on every handle at once, each on its
 * own stream, is timed from the common start to the last finish, must give
 * the same bits on every handle, and the fastest is set on all of them
 * (kinc_sim_autotune is the group of one).  Errors are reported on hs[0].
 */
typedef struct kinc_sim_tuning
{
   int32_t pair_waves;
   int32_t aux_streams;
   int32_t joint45;
   int32_t candidates;
   double sample_ms;
} kinc_sim_tuning;

#define KINC_AUTOTUNE_MAX_HANDLES 16

int kinc_sim_autotune(kinc_sim* h, int64_t start, int64_t count, kinc_sim_tuning* out);
int kinc_sim_autotune_group(kinc_sim* const* hs, int32_t n, int64_t start, int64_t count,
                            kinc_sim_tuning* out);
int kinc_sim_get_tuning(const kinc_sim* h, kinc_sim_tuning* out);
int kinc_sim_set_tuning(kinc_sim* h, const kinc_sim_tuning* t);

/*
 * The launch plan a handle would get for these parameters and this matrix
 * shape, computed on the host without a device (no counterpart in the
 * reference, whose OpenCL worker sizes nothing per shape; for callers that
 * size work before creating handles, and for tests of the layout rules):
 * the register-slot classes of the bins (0 = the memory-store instance), the
 * per-wave LDS bytes without and with room for the lockstep 4- and 5-cluster
 * tile (0: not offered), a bit mask of the pair_waves values kinc_sim_set_
 * tuning accepts for some joint45 setting (bit w), and the configuration
 * the handle starts with.  Returns 0 or a negative code as kinc_sim_create
 * would for the same arguments (before touching a device).
 */
typedef struct kinc_sim_plan
{
   int32_t bins;
   int32_t slots[12];
   int32_t lds_small;
   int32_t lds_joint45;
   int32_t pair_waves_allowed;
   kinc_sim_tuning tuning;
} kinc_sim_plan;

int kinc_sim_plan_get(const kinc_sim_params* params, int32_t G, int32_t S, kinc_sim_plan* out);

/*
 * HIP graphs of work blocks (no counterpart in the reference, whose OpenCL
 * worker enqueues every kernel of every sub-batch itself).  Between
 * kinc_sim_graph_begin and kinc_sim_graph_end, kinc_sim_execute_device and
 * kinc_sim_execute_pairs_device calls on the handle's own stream (stream
 * NULL) are recorded instead of run; kinc_sim_graph_end instantiates them as
 * one graph, which kinc_sim_graph_launch replays on `stream` with the same
 * pair ranges and output buffers (the device queues are cleared by kernel
 * nodes, so every replay gives the results of the first).  Host-buffer
 * calls and kernel timing are refused while recording.  A graph outlives
 * neither its handle nor its buffers.
 */
typedef struct kinc_sim_graph kinc_sim_graph;
int kinc_sim_graph_begin(kinc_sim* h);
int kinc_sim_graph_end(kinc_sim* h, kinc_sim_graph** graph);
int kinc_sim_graph_launch(kinc_sim_graph* graph, void* stream);
void kinc_sim_graph_destroy(kinc_sim_graph* graph);

/* Number of linear pairs for G genes: G(G-1)/2 (src/core/similarity.cpp:19). */
int64_t kinc_pair_count(int32_t gene_size);

/* Map a linear pair index to (x, y), x > y >= 0 (src/core/pairwise_index.cpp:30-56). */
void kinc_pair_index(int64_t index, int32_t* x, int32_t* y);

#ifdef __cplusplus
}
#endif

#endif
