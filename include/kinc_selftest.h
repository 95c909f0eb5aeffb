/*
 * kinc_selftest.h — diagnostic entry point of the MI355X Similarity library
 * (no counterpart in the reference).
 *
 * The pair kernels and the CPU code share one FP32 exp/log implementation
 * and rely on correctly rounded sqrt and division on the device; this entry
 * point evaluates those functions on the device so that tests can check the
 * device results bit for bit against the host.
 */
#ifndef KINC_SELFTEST_H
#define KINC_SELFTEST_H

#include <stdint.h>

#include "kinc_sim.h"

/*
 * Phase census of the pair kernel: counters of an instrumented build of the
 * 7-slot, 5-cluster instance (the register-slot kernel of pairs with 385 to
 * 448 valid samples), for profiling only.  Indices 0..15 are shader-clock
 * cycles (s_memtime deltas) that waves spent in each phase, summed over
 * waves; 16..31 are event counts.
 */
#define KINC_CENSUS_FETCH        0   /* pair ids, fetch, compaction, pre-clustering outliers */
#define KINC_CENSUS_FIT_INIT     1   /* LCG draws and initial parameters of a GMM fit */
#define KINC_CENSUS_ESTEP        2   /* E step (log densities, log-sum-exp) */
#define KINC_CENSUS_GMAX         3   /* per-component maxima of the log responsibilities */
#define KINC_CENSUS_A_TERMS      4   /* pass A: per-point terms into the LDS tile */
#define KINC_CENSUS_A_CHAIN      5   /* pass A: sequential chain sums over the tile */
#define KINC_CENSUS_MSTEP        6   /* convergence test, Gamma_k, pi_k, mu_k */
#define KINC_CENSUS_B_TERMS      7   /* pass B: per-point covariance terms into the tile */
#define KINC_CENSUS_B_CHAIN      8   /* pass B: chain sums */
#define KINC_CENSUS_COV          9   /* covariance inverse and determinant */
#define KINC_CENSUS_LABELS      10   /* final E step of a converged fit, labels */
#define KINC_CENSUS_ENTROPY     11   /* ICL entropy terms and their chain sum */
#define KINC_CENSUS_SELECT      12   /* BIC/ICL, label copy-back, post-clustering outliers */
#define KINC_CENSUS_CORR        13   /* per-cluster correlation and result stores */
#define KINC_CENSUS_OTHER_FIT   14   /* K-means fits */
#define KINC_CENSUS_WAVE        15   /* whole wave lifetime (kernel entry to exit) */
#define KINC_CENSUS_PAIRS       16   /* pairs processed */
#define KINC_CENSUS_FITS        17   /* GMM fits started */
#define KINC_CENSUS_ITERS       18   /* EM iterations (E steps inside the loop) */
#define KINC_CENSUS_ESTEP_SLOTS 19   /* register slots (64 points) processed by E steps */
#define KINC_CENSUS_A_SLOTS     20   /* slots processed by pass A */
#define KINC_CENSUS_B_SLOTS     21   /* slots processed by pass B */
#define KINC_CENSUS_CHAIN_ADDS  22   /* wave instructions of the chain sums (one add each) */
#define KINC_CENSUS_CHAIN_LANES 23   /* the same, weighted by the lanes active in each */
#define KINC_CENSUS_CHAIN_CALLS 24   /* chain-sum calls (one tile row set each) */
#define KINC_CENSUS_EXPS        25   /* exp evaluations over a whole wave (E step, passes A, M step) */
#define KINC_CENSUS_LOGS        26   /* log evaluations over a whole wave */
#define KINC_CENSUS_QUEUE       27   /* cycles in queue pops (atomic on the bin counter) */
#define KINC_CENSUS_COUNTERS    32

#ifdef __cplusplus
extern "C" {
#endif

/* out[i] = f(in[i]) on the current device for f = 0 exp, 1 log, 2 sqrt,
 * 3 reciprocal (1/x), 4 log for finite x >= 1 or NaN, 5 exp for x <= 0 or
 * NaN (the log-sum-exp forms the E step uses).  Host buffers.  Returns 0 on
 * success, < 0 on error. */
int kinc_selftest_math(int32_t fn, const float* in, float* out, int64_t n);

/*
 * VALU issue-rate probe: launches `workgroups` one-wave workgroups (with
 * lds_bytes of dynamic LDS each, which caps the waves resident per CU), each
 * issuing exactly 128 * iters instructions of one kind on the active lanes
 * of `mode` (independent chains / active lanes / instruction):
 *    0: 8 / 64 / v_add_f32      1: 8 / 16 / v_add_f32    2: 8 / 4 / v_add_f32
 *    3: 1 / 64 / v_add_f32      4: 1 / 16 / v_add_f32    5: 8 / 32 / v_add_f32
 *    6: 8 / 64 / v_mul_f32      7: 8 / 64 / v_fma_f32    8: 8 / 64 / v_fmac_f32
 *    9: 8 / 64 / v_max_f32     10: 8 / 64 / v_rndne_f32 11: 8 / 64 / v_ldexp_f32
 *   12: 8 / 64 / v_cvt_i32_f32 13: 8 / 64 / v_cndmask_b32
 *   14: 8 / 64 / v_fmaak_f32   15: 1 / 64 / v_fma_f32   16: 8 / 16 / v_fma_f32
 * and writes the duration of the timed launch (after one warm-up launch) to
 * *ms.  Returns 0 on success, < 0 on error.
 */
int kinc_selftest_valu_probe(int32_t mode, int32_t iters, int32_t workgroups,
                             int32_t lds_bytes, float* ms);

/*
 * Switch the handle's pair launches to the census build for pairs of the
 * 7-slot, 5-cluster instance (d_counters: KINC_CENSUS_COUNTERS device
 * uint64 counters that the kernels add to; the caller zeroes them), or back
 * to the product build (d_counters == NULL).  Results are the same with and
 * without the census; only the timing differs.  KINC_E_UNSUPPORTED if the
 * handle has no such instance (max_clusters above 5, or no 7-slot bin).
 */
int kinc_selftest_census(kinc_sim* h, uint64_t* d_counters);

#ifdef __cplusplus
}
#endif

#endif
