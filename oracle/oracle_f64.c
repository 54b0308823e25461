/* FP64 variant of the oracle: the same algorithm carried out in double.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#define _POSIX_C_SOURCE 199309L
#define R double
#define FN(x) x##_f64
#include "orc_impl.h"
