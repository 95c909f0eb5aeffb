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

#ifdef __cplusplus
}
#endif

#endif
