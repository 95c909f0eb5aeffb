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

#ifdef __cplusplus
}
#endif

#endif
