/*
 * kinc_formats.h — EMX input and CCM/CMX output of the Similarity analytic.
 *
 * Payload layouts are KINC's (the data section ACE exposes at seek(0)):
 *   EMX  int32 geneSize, int32 sampleSize, float32[G][S] gene-major, NaN =
 *        missing (src/core/expressionmatrix.cpp:21-37,302-312)
 *   pairwise base header, 30 bytes: int32 geneSize, int32 maxClusterSize,
 *        int32 dataSize, int64 pairSize, int64 clusterSize, int16 offset,
 *        then `offset` bytes of sub-header, then clusterSize records of a
 *        9-byte item header (int32 x, int32 y, int8 cluster) and dataSize
 *        bytes, sorted strictly by (x(x-1)/2 + y)*64 + cluster
 *        (src/core/pairwise_matrix.h:32-40, .cpp:35-57,122-151)
 *   CCM  sub-header int32 sampleSize; dataSize = (S+1)/2; sample i in the
 *        low nibble of byte i/2 for even i, the high nibble for odd i
 *        (src/core/ccmatrix.cpp:108-127,223-243)
 *   CMX  sub-header int8 correlationSize; dataSize = 4*correlationSize
 *        (src/core/correlationmatrix.cpp:108-127,231-243)
 *
 * The ACE file envelope is not available, so files carry this library's own
 * little-endian envelope around that payload:
 *     0  char[8]  "KINCAMD1"
 *     8  uint32   type: 1 EMX, 2 CCM, 3 CMX
 *    12  uint32   0
 *    16  uint64   payload offset (64)
 *    24  uint64   payload size
 *    32  uint64   metadata offset
 *    40  uint64   metadata size
 *    48  16 bytes zero
 *    64  payload, then the metadata as UTF-8 JSON (an object with the
 *        reference's keys: "genes", "samples", "transform", "correlations")
 */
#ifndef KINC_FORMATS_H
#define KINC_FORMATS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KINC_FILE_EMX 1
#define KINC_FILE_CCM 2
#define KINC_FILE_CMX 3

/* Write an expression matrix.  metadata_json may be NULL (then "{}"). */
int kinc_emx_write(const char* path, const float* data, int32_t gene_size,
                   int32_t sample_size, const char* metadata_json);

/* Read the EMX header. */
int kinc_emx_info(const char* path, int32_t* gene_size, int32_t* sample_size);

/* Read the expression values into data[G][S]. */
int kinc_emx_read(const char* path, float* data);

/* Copy up to cap bytes of a file's metadata JSON into buf; returns the full
 * metadata size (or < 0 on error). */
int64_t kinc_file_metadata(const char* path, char* buf, int64_t cap);

/* Element-wise transforms of ExpressionMatrix::Transform
 * (src/core/expressionmatrix.cpp:8-14: "none", "natural logarithm",
 * "logarithm base 2", "logarithm base 10"). */
#define KINC_TRANSFORM_NONE 0
#define KINC_TRANSFORM_NLOG 1
#define KINC_TRANSFORM_LOG2 2
#define KINC_TRANSFORM_LOG10 3

/*
 * ImportExpressionMatrix::process (src/core/importexpressionmatrix.cpp:20-170):
 * whitespace-delimited text to EMX.  With sample_size == 0 the first line
 * holds the sample names, otherwise samples are named 0..sample_size-1 and
 * every line is data.  A data line is a gene name and sample_size values; a
 * value equal to nan_token (NULL or "" matches nothing) is NaN, any other is
 * read as a double, transformed (log, log2, log10 in double) and stored as
 * float.  A line with another number of fields, or a value that does not
 * parse, fails with the reference's message.  threads <= 0 uses every host
 * thread (at most 16).  gene_size / sample_size_out may be NULL.
 */
int kinc_emx_import_text(const char* text_path, const char* emx_path, const char* nan_token,
                         int32_t sample_size, int32_t transform, int32_t threads,
                         int32_t* gene_size, int32_t* sample_size_out);

/*
 * ExportExpressionMatrix::process (src/core/exportexpressionmatrix.cpp:20-99):
 * EMX to text.  First line: every sample name followed by a tab; then per
 * gene its name and, per sample, a tab and either nan_token (for NaN) or the
 * value with the stored transform undone, printed with 12 significant
 * digits (QTextStream's SmartNotation at precision 12, i.e. "%.12g").
 * Names come from the file's "genes"/"samples" metadata, else 0, 1, 2, ...
 */
int kinc_emx_export_text(const char* emx_path, const char* text_path, const char* nan_token);

/*
 * CCM + CMX writer: Similarity::initialize / process / finish
 * (src/core/similarity.cpp:66-127,164-197; pairwise_matrix.cpp:51-57).
 * process() takes the results of one block in pair order; blocks must come
 * in increasing pair order (Matrix::write refuses a non-increasing indent).
 */
typedef struct kinc_similarity_writer kinc_similarity_writer;

int kinc_similarity_writer_open(const char* ccm_path, const char* cmx_path,
                                int32_t gene_size, int32_t sample_size,
                                int32_t max_clusters, int32_t corr_method,
                                float min_corr, float max_corr,
                                const char* metadata_json,
                                kinc_similarity_writer** out);

/* K[count], labels[count][sample_size], corr[count][corr_stride]. */
int kinc_similarity_writer_process(kinc_similarity_writer* w, int64_t start, int64_t count,
                                   const int8_t* K, const int8_t* labels,
                                   const float* corr, int32_t corr_stride);

/* Rewrite the headers with the final counts and close both files. */
int kinc_similarity_writer_close(kinc_similarity_writer* w);

/*
 * ExportCorrelationMatrix::process (src/core/exportcorrelationmatrix.cpp:17-117):
 * one text line per CMX cluster, in file order:
 *   x, y, cluster, cluster count, and for pairs with more than one cluster
 *   the counts of CCM nibbles 1 (in the cluster), 9 (missing), 8 (post-
 *   clustering outlier), 7 (pre-clustering outlier), 6 (below the minimum
 *   expression) (else zeros), the correlation ("%g"), and the CCM mask as
 *   one character '0' + nibble per sample (all '0' for a one-cluster pair);
 * fields separated by tabs.
 */
int kinc_export_cmx_text(const char* ccm_path, const char* cmx_path, const char* out_path);

/*
 * Extract::process (src/core/extract.cpp:14-275): the network of the
 * clusters with min_corr <= |r| <= max_corr (a cluster outside is skipped),
 * as a tab-separated edge list with the header
 *   Source Target sc Interaction Cluster Num_Clusters Cluster_Samples
 *   Missing_Samples Cluster_Outliers Pair_Outliers Too_Low Samples
 * (gene names from the CMX metadata, interaction "co", the counts and mask
 * as kinc_export_cmx_text, except that a one-cluster pair's mask comes from
 * the expression matrix: '9' where either gene is NaN, else '1'), and, when
 * graphml_path is not NULL, the same edges as GraphML with every gene a node.
 */
int kinc_extract_network(const char* emx_path, const char* ccm_path, const char* cmx_path,
                         const char* out_path, const char* graphml_path, float min_corr,
                         float max_corr);

/* Message of the last failed format call of this thread. */
const char* kinc_formats_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
