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
 * issuing exactly 128 * iters v_add_f32 on the active lanes of `mode`:
 *   0: 8 independent chains, 64 lanes   1: 8 chains, 16 lanes
 *   2: 8 chains, 4 lanes                3: 1 dependent chain, 64 lanes
 *   4: 1 dependent chain, 16 lanes      5: 8 chains, 32 lanes
 * and writes the duration of the timed launch (after one warm-up launch) to
 * *ms.  Returns 0 on success, < 0 on error.
 */
int kinc_selftest_valu_probe(int32_t mode, int32_t iters, int32_t workgroups,
                             int32_t lds_bytes, float* ms);

#ifdef __cplusplus
}
#endif

#endif
