/* FP32 variant of the oracle: the reference's float arithmetic, in its order.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#define _POSIX_C_SOURCE 199309L
#define R float
#define FN(x) x##_f32
#define ORC_WITH_TIMER
#include "orc_impl.h"
