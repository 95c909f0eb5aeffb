/*
 * sim_pairs_inst.hip — one pair-kernel instance per translation unit.
 *
 * Compiled once per (KINC_INST_P, KINC_INST_KMAX, KINC_PAIR_WAVES) by
 * kinc_amd/build.py so that the instances build in parallel; each exports a
 * getter named kinc_pair_kernel_k<KMAX>_p<P> (4 waves per SIMD) or
 * kinc_pair_kernel_k<KMAX>_p<P>_w<W> (W = 3 or 5) returning the kernel's
 * host handle;
 * with KINC_CENSUS, the instrumented build kinc_pair_kernel_census_k<KMAX>_p<P>.
 */
#include "sim_pair_impl.h"

#ifndef KINC_INST_P
#error "KINC_INST_P must be defined"
#endif
#ifndef KINC_INST_KMAX
#error "KINC_INST_KMAX must be defined"
#endif

#define KINC_CAT2(a, b, c, d) a##b##c##d
#define KINC_CAT(a, b, c, d) KINC_CAT2(a, b, c, d)

#if defined(KINC_CENSUS)
/* the phase-census build of an instance (profiling only) */
extern "C" const void* KINC_CAT(kinc_pair_kernel_census_k, KINC_INST_KMAX, _p, KINC_INST_P)()
{
   return (const void*)&k_pairs_census<KINC_INST_P, KINC_INST_KMAX, KINC_PAIR_WAVES>;
}
#elif KINC_PAIR_WAVES == 4
extern "C" const void* KINC_CAT(kinc_pair_kernel_k, KINC_INST_KMAX, _p, KINC_INST_P)()
#else
/* a build of an instance for another number of waves per SIMD: 5 (small
 * instances, where the LDS leaves room for 20 waves per CU) or 3 (where the
 * LDS allows at most 12 waves per CU, so the kernel may use 168 VGPRs at the
 * same residency) */
#define KINC_CAT3(a, b, c, d, e, f) a##b##c##d##e##f
#define KINC_CATW(a, b, c, d, e, f) KINC_CAT3(a, b, c, d, e, f)
extern "C" const void* KINC_CATW(kinc_pair_kernel_k, KINC_INST_KMAX, _p, KINC_INST_P, _w, KINC_PAIR_WAVES)()
#endif
#if !defined(KINC_CENSUS)
{
   return (const void*)&k_pairs<KINC_INST_P, KINC_INST_KMAX, KINC_PAIR_WAVES>;
}
#endif
