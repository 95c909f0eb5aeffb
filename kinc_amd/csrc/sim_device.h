/*
 * sim_device.h — argument block shared by the host launcher and the HIP
 * kernels of the pairwise Similarity path.
 */
#ifndef KINC_SIM_DEVICE_H
#define KINC_SIM_DEVICE_H

#include <stdint.h>

/* 2*log(2*pi) in double: the D*log(2.0f*M_PI) term of the GMM normaliser
 * (src/core/pairwise_gmm.cpp:46, D = 2). */
#define KINC_TWO_LOG_2PI 3.6757541328186907

/* Points held per lane in registers: a pair with N valid samples runs on the
 * register-slot instance with 64*P >= N; above 64*KINC_MAX_P samples the
 * memory-store instance (P = 0) runs. */
#define KINC_MAX_P 16
#define KINC_MAX_REG_SAMPLES (64 * KINC_MAX_P)
/* Largest sample count (Spearman pads to the next power of two, 32768). */
#define KINC_MAX_SAMPLES 32768
/* Floats per chain row of the LDS term tile: 64 points and 4 of padding, so
 * the 16-byte reads of one row per lane spread over the banks.  A compile-
 * time constant lets every tile address fold into the LDS instruction's
 * offset field. */
#define KINC_TILE_STRIDE 68

/* The run-time-K instance (max_clusters 9..64, KMAX = 64, memory store):
 * LDS bytes of its parameter area (two sets of 64 components x 8 floats and
 * 64 per-component maxima) and rows of its chain tile. */
#define KINC_DYN_PRM_BYTES (4 * (128 * 8 + 64))
#define KINC_DYN_TILE_ROWS 64

/* Bins of the persistent form: at most this many register-slot classes. */
#define KINC_MAX_BINS 12

/* Slot count of each bin (a pair goes to the first bin with 64*P >= N; the
 * memory-store bin's entry is 2^20), passed to the binning kernel by value. */
struct BinBounds
{
   int32_t P[KINC_MAX_BINS];
};

/* Pair-kernel argument block (passed by value). */
struct SimArgs
{
   const float* emx;         /* [G][S] resident expression matrix */
   int32_t G;
   int32_t S;
   int64_t block_start;      /* linear index of block-relative pair 0 */
   int32_t block_count;      /* pairs in the block */
   int32_t first;            /* plain form: first block-relative pair of the launch */
   int32_t launch_count;     /* plain form: pairs in the launch */
   const int32_t* list;      /* persistent form: block-relative pair ids of one bin */
   int32_t* list_head;       /* persistent form: atomic pop counter of the bin */
   const int32_t* list_len;  /* persistent form: number of ids in the bin */

   int32_t clus;
   int32_t corr;
   float min_expr;
   int32_t min_samples;
   int32_t min_k;
   int32_t max_k;
   int32_t crit;
   int32_t pre_out;
   int32_t post_out;
   int32_t pearson_fp64;     /* Pearson sums and quotient in double */
   int32_t joint45;          /* GMM fits with 4 and 5 clusters in lockstep (register-
                                slot instances of at most 7 slots): the host sets it
                                when X and the shared LDS region after it hold their
                                29-row chain tile */

   const int2* pairs;        /* explicit (x, y) of each block-relative pair, or null
                                for linear indices from block_start */
   const int32_t* seg_lo;    /* first block-relative pair of each pair's stale-read
                                segment (the work block it belongs to), or null (0) */
   const double* logN;       /* log((double)n) for n in [0, S] */
   const int32_t* blockN;    /* valid-sample count of every pair of the block */
   const int32_t* chunkMax;  /* max of blockN over 64-pair chunks */
   const float* xbuf;        /* [block_count][S] float pairs: each pair's compact X,
                                written by the count kernel, or null */
   const int32_t* prev;      /* [block_count] latest earlier pair of the segment with
                                more valid samples (-1: none), with xbuf */

   int8_t* K_out;            /* [block_count] */
   int8_t* labels_out;       /* [block_count][S] or null */
   float* corr_out;          /* [block_count][max_k] */
   uint16_t* iters_out;      /* [block_count][max_k] or null */

   /* Working-set carve-up.  lds_prm and lds_tile are offsets in the dynamic
    * shared segment.  lds_x, lds_lab, lds_w and lds_u are offsets in the
    * shared segment too when scratch is null (register-slot instances), and
    * offsets in the workgroup's scratch slice otherwise (memory-store
    * instance, any sample count). */
   int32_t lds_x;            /* float2[S]  compact X, then the stale-read view */
   int32_t lds_lab;          /* int8[S]    full-index labels */
   int32_t lds_w;            /* int8[S]    work labels of the current fit */
   int32_t lds_prm;          /* float[9][8] component parameters (register-slot
                                instances; [8][8] on the memory-store instance) */
   int32_t lds_tile;         /* chain-term tile */
   int32_t lds_u;            /* sort buffers / sequence scratch */
   char* scratch;            /* memory-store instance: slices of scratch_stride bytes */
   int64_t scratch_stride;
   int32_t scr_lg;           /* float[(KMAX+1)][mem_slots][64] responsibilities, log p(x) */
   int32_t scr_y;            /* int8[mem_slots][64] K-means labels */
   int32_t mem_slots;        /* ceil(S / 64) */
   int32_t np2;              /* next power of two >= S (min 2) */
   int32_t tile_stride;      /* KINC_TILE_STRIDE */
   int32_t ll_cap;           /* floats at lds_u the K-means log-likelihood terms
                                may fill before they are added (a multiple of 4,
                                at least 64) */
   unsigned long long* census;  /* phase-census build only: its counters */
};

#endif
