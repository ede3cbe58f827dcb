/* ora_internal.h -- private layout shared by the oracle's translation units.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef ORA_INTERNAL_H
#define ORA_INTERNAL_H
#include "oracle.h"

/* approx/approx.c:10-24 */
struct ora_approx {
        size_t nrhs, nvars;
        ora_matrix_t * matrix;
        double * rhs, * weight, * linear, * lower, * upper;
        uint32_t * beta;
        double * v, * inv_v;
        struct ora_approx * permuted;
};

#endif
