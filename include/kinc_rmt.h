/*
 * kinc_rmt.h — the RMT threshold analytic on one MI355X: the random-matrix
 * test that picks a correlation threshold for the network (RMT::process,
 * src/core/rmt.cpp:33-109).
 *
 * The correlation matrix (a CMX file's records) stays resident in HBM as one
 * dense G x G float plane per cluster index.  For each threshold t, from
 * threshold_start down by threshold_step (accumulated in float, as the
 * reference does):
 *   - the rows/columns (gene, cluster) whose largest |r| reaches t are kept
 *     (RMT::computeMaximums/computePruneMatrix, rmt.cpp:160-236);
 *   - the pruned matrix (|r| >= t between genes of the same cluster index,
 *     ones on the diagonal) is the direct sum of the connected components of
 *     that graph; each component that changed since the previous threshold
 *     is gathered on the device in double and solved by rocSOLVER (dsyevd),
 *     one- and two-row components in closed form; together their
 *     eigenvalues are the ones GSL's symmv returns for the whole matrix
 *     (rmt.cpp:243-275), ordered by increasing |value| and rounded to float;
 *   - the chi-square of their nearest-neighbour spacings against a Poisson
 *     distribution (rmt.cpp:282-439: unique values, Akima unfolding at paces
 *     min_pace..max_pace, a histogram of bins bins on [0, 3)) decides: the
 *     last t with chi < chi_square1 is the answer once a later chi reaches
 *     chi_square2.
 * The reference's GSL aborts where the spline knots (every pace-th unique
 * eigenvalue in |value| order) are not strictly increasing, which happens
 * when negative eigenvalues interleave the positive ones; here that ends the
 * run with KINC_E_INVALID and a message instead.
 */
#ifndef KINC_RMT_H
#define KINC_RMT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* RMT members and their command-line defaults (src/core/rmt.h:31-40,
 * rmt_input.cpp:76-136). */
typedef struct kinc_rmt_params
{
   float threshold_start;    /* --tstart, 0.99 */
   float threshold_step;     /* --tstep, 0.001 */
   float threshold_stop;     /* --tstop, 0.5 */
   float chi_square1;        /* 99.607: a chi below it is non-random */
   float chi_square2;        /* 200: a later chi at or above it ends the sweep */
   int32_t min_eigenvalues;  /* 50 unique eigenvalues at least */
   int32_t min_pace;         /* --minpace, 10 */
   int32_t max_pace;         /* --maxpace, 40 */
   int32_t bins;             /* --bins, 60 */
   int32_t reserved[3];
} kinc_rmt_params;

void kinc_rmt_default_params(kinc_rmt_params* p);

typedef struct kinc_rmt kinc_rmt;

/*
 * Upload a correlation matrix given as CMX records: record i is the
 * correlation corr[i] of genes x[i] > y[i] in cluster cluster[i]
 * (0 <= cluster < max_clusters), records in strictly increasing
 * (x(x-1)/2 + y, cluster) order as a CMX file stores them
 * (CorrelationMatrix::dumpRawData, correlationmatrix.cpp:144-177, fills the
 * same dense matrix; absent entries are 0).
 */
int kinc_rmt_create(const kinc_rmt_params* params, int32_t gene_size, int32_t max_clusters,
                    const int32_t* x, const int32_t* y, const int8_t* cluster, const float* corr,
                    int64_t n_records, int32_t device, kinc_rmt** out);

/* Eigenvalues of the matrix pruned at threshold, |value|-ascending, as
 * float; *size = the pruned matrix's size.  With eigens == NULL only *size is
 * set; otherwise cap must be >= *size. */
int kinc_rmt_eigenvalues(kinc_rmt* h, float threshold, float* eigens, int64_t cap, int64_t* size);

/* RMT::computeChiSquare of an eigenvalue list (host computation): the mean
 * chi-square over the paces, -1 with fewer than min_eigenvalues unique
 * values, -2 where the spline knots are not strictly increasing. */
float kinc_rmt_chi_square(const kinc_rmt_params* params, const float* eigens, int64_t n);

/*
 * RMT::process: the threshold sweep.  Row s of the log (s < cap) gets the
 * threshold, the pruned size and the chi-square (-1: not computed) of step s;
 * *steps = number of thresholds evaluated.  Returns KINC_OK with the final
 * threshold, or KINC_E_INVALID when the threshold fell below threshold_stop
 * first ("Could not find non-random threshold above stopping threshold.") or
 * the spline knots failed; the log holds the steps done.
 */
int kinc_rmt_run(kinc_rmt* h, float* log_threshold, int32_t* log_size, float* log_chi, int64_t cap,
                 int64_t* steps, float* final_threshold);

/* The analytic end to end: read a CMX file, sweep, write the reference's log
 * (one "threshold<TAB>size<TAB>chi" line per step, then the final threshold,
 * numbers as QTextStream prints them, "%g"). */
int kinc_rmt_run_file(const kinc_rmt_params* params, const char* cmx_path, const char* log_path,
                      int32_t device, float* final_threshold);

void kinc_rmt_destroy(kinc_rmt* h);

/* Message of the handle's last error (or of the last failed create/run_file
 * of this thread when h is NULL). */
const char* kinc_rmt_last_error(const kinc_rmt* h);

#ifdef __cplusplus
}
#endif

#endif
