/*
 * sim_pairs_k5.hip — pair-kernel instances for up to 5 clusters per pair.
 */
#include "sim_pair_impl.h"

#define KINC_FOR_EACH_P(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(10) X(12) X(14) X(16)

extern "C" const void* kinc_select_pair_kernel_k5(int P, int persistent)
{
#define KINC_CASE(PP)                                                                \
   if ( P == PP )                                                                   \
   {                                                                                \
      return persistent ? (const void*)&k_pairs_persistent<PP, 5>                  \
                        : (const void*)&k_pairs_plain<PP, 5>;                      \
   }
   KINC_FOR_EACH_P(KINC_CASE)
#undef KINC_CASE
   return nullptr;
}
