/*
 * kinc_server.h — query serving on one MI355X with the expression matrix
 * resident (no counterpart in the reference, which only runs whole-matrix
 * batch jobs).
 *
 * A query is a set of genes; its answer is every pair (query gene, any gene)
 * whose per-cluster correlation r passes !NaN and min_abs_corr <= |r| <=
 * max_corr, with the cluster selected by the Similarity analytic exactly as
 * kinc_sim_execute computes it.  The pairs of a query are listed query gene
 * by query gene (in the order given, duplicates dropped), partners in
 * increasing gene order, a pair between two query genes once (under the
 * earlier one), each as (max, min); the list is cut into work blocks of
 * 32,768 pairs, the unit of the reference's stale-buffer Pearson semantics.
 *
 * Continuous batching: a scheduler thread admits queued queries in arrival
 * order, each one whose device result buffer fits in the HBM budget (a query
 * that does not fit holds back the later ones only after it has waited 5 s),
 * then packs whole work blocks of the admitted queries into batches, each
 * one persistent-kernel launch: fewest remaining blocks first, one work
 * block per batch while more than one query has work or waits (the
 * scheduling quantum), up to batch_pairs pairs while one query runs alone.
 * A query that arrives while others run joins the next batch, so it waits
 * at most for the two batches in flight (one on each of the server's two
 * engine handles).  Results are the same whether a query runs alone or
 * mixed with others.
 */
#ifndef KINC_SERVER_H
#define KINC_SERVER_H

#include <stdint.h>

#include "kinc_sim.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kinc_server kinc_server;

/* One answer row: pair (x, y), x > y, its cluster (0-based, renumbered over
 * the clusters that pass, as Similarity::process numbers CMX records) and r. */
typedef struct kinc_hit
{
   int32_t x;
   int32_t y;
   int32_t cluster;
   float corr;
} kinc_hit;

/* Start a server: params as kinc_sim_create (its batch_capacity is replaced
 * by batch_pairs), hbm_budget_bytes caps the device result buffers of the
 * admitted queries (1 + 4*max_clusters bytes per pair). */
int kinc_server_create(const kinc_sim_params* params, const float* emx, int32_t gene_size,
                       int32_t sample_size, int32_t device, int64_t hbm_budget_bytes,
                       int32_t batch_pairs, kinc_server** out);

/* Queue a query; returns its id in *id.  A query whose result buffer alone
 * exceeds the HBM budget is refused with KINC_E_NOMEM. */
int kinc_server_submit(kinc_server* s, const int32_t* genes, int32_t n_genes,
                       float min_abs_corr, int64_t* id);

/* Wait up to timeout_s seconds; returns 1 if the query is done, 0 if not,
 * < 0 on error (an unknown id, or the query failed). */
int kinc_server_wait(kinc_server* s, int64_t id, double timeout_s);

/*
 * Answer of a finished query.  With hits == NULL only *n_hits is set.
 * Otherwise up to cap rows are copied and, if all fitted, the query is
 * released.  latency_s = submit to completion; n_pairs = pairs evaluated.
 */
int kinc_server_result(kinc_server* s, int64_t id, kinc_hit* hits, int64_t cap,
                       int64_t* n_hits, double* latency_s, int64_t* n_pairs);

/* Timing of a finished query (before kinc_server_result releases it):
 * submit to the launch of its first block, and from there to completion. */
int kinc_server_timing(kinc_server* s, int64_t id, double* start_wait_s, double* run_s);

/* Totals since start: pairs evaluated, batches launched, seconds with a
 * batch on the device. */
int kinc_server_stats(kinc_server* s, int64_t* pairs_done, int64_t* batches, double* busy_s);

/* Stop the scheduler (queued queries are dropped) and free everything. */
void kinc_server_destroy(kinc_server* s);

const char* kinc_server_last_error(const kinc_server* s);

#ifdef __cplusplus
}
#endif

#endif
