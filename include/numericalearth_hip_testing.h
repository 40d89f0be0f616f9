/*
 * numericalearth_hip_testing.h — entry points of libnumericalearth_hip that exist for tests
 * of the library's own numerics and have no counterpart in the reference. They are exported by
 * the same library; a caller of the coupled step never needs them.
 */
#ifndef NUMERICALEARTH_HIP_TESTING_H
#define NUMERICALEARTH_HIP_TESTING_H

#include "numericalearth_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The flux solve's FP64 natural logarithm applied elementwise to x[n] (DEVICE arrays):
 * y[i] = log(x[i]) within 2 ulp for positive, finite, normal x (other arguments are not
 * defined). It exposes the library's own logarithm (argument reduction to [√½, √2) and the
 * atanh series, in place of the math library's) so its accuracy can be checked against a
 * correctly rounded one. */
int ne_log_f64(ne_ctx* ctx, const double* x, double* y, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NUMERICALEARTH_HIP_TESTING_H */
