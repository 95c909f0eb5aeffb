/*
 * sim_pairs_inst.hip — one pair-kernel instance per translation unit.
 *
 * Compiled once per (KINC_INST_P, KINC_INST_KMAX) by kinc_amd/build.py so
 * that the instances build in parallel; each exports a getter named
 * kinc_pair_kernel_k<KMAX>_p<P> returning the kernel's host handle.
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

extern "C" const void* KINC_CAT(kinc_pair_kernel_k, KINC_INST_KMAX, _p, KINC_INST_P)()
{
   return (const void*)&k_pairs<KINC_INST_P, KINC_INST_KMAX>;
}
