/*
 * numericalearth_hip_testing.h — entry points of libnumericalearth_hip that exist for tests
 * of the library's own numerics and for measurements of its kernels, and have no counterpart in
 * the reference. They are exported by the same library; a caller of the coupled step never needs
 * them.
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

/* The flux solve's FP64 arctangent for arguments x ≥ 1 applied elementwise to x[n] (DEVICE
 * arrays): y[i] = atan(x[i]) within 2 ulp for x ≥ 1 (+∞ included; other arguments are not
 * defined). It exposes the library's own arctangent (one quotient of magnitude ≤ tan(π/8) and a
 * degree-10 polynomial in its square, in place of the math library's) so its accuracy can be
 * checked against a correctly rounded one. */
int ne_atan_ge1_f64(ne_ctx* ctx, const double* x, double* y, int64_t n, void* stream);

/* The flux solve's FP64 cube root for arguments x ≥ 1 applied elementwise to x[n] (DEVICE
 * arrays): y[i] = cbrt(x[i]) within 1 ulp for 1 ≤ x < 2¹²⁸ (other arguments are not defined). It
 * exposes the library's own cube root (an FP32 first value and one FP64 Halley step, in place of
 * the math library's) so its accuracy can be checked against a correctly rounded one. */
int ne_cbrt_ge1_f64(ne_ctx* ctx, const double* x, double* y, int64_t n, void* stream);

/* The flux solve's FP64 square root applied elementwise to x[n] (DEVICE arrays): y[i] = √x[i]
 * within 1 ulp for positive, finite, normal x (other arguments are not defined). It exposes the
 * library's own square root (the hardware's reciprocal square root refined by Goldschmidt steps,
 * in place of the math library's correctly rounded one) so its accuracy can be checked. */
int ne_sqrt_f64(ne_ctx* ctx, const double* x, double* y, int64_t n, void* stream);

/* The flux solve's FP64 exponential applied elementwise to x[n] (DEVICE arrays): y[i] = e^x[i]
 * within 2 ulp for |x| ≤ 700 (other arguments are not defined). It exposes the library's own
 * exponential (reduction by ln 2 and a degree-13 Taylor polynomial, in place of the math
 * library's) so its accuracy can be checked. */
int ne_exp_f64(ne_ctx* ctx, const double* x, double* y, int64_t n, void* stream);

/* A kernel of known HBM traffic, to read hardware memory counters against: out_j[i] =
 * Σ_k in_k[i] + j for i < n, coalesced and i fastest, so one launch reads n_in·n·8 bytes and
 * writes n_out·n·8 bytes. `in` and `out` are HOST arrays of n_in ≤ NE_STREAM_PROBE_MAX_IN and
 * n_out ≤ NE_STREAM_PROBE_MAX_OUT DEVICE pointers to n doubles each (distinct, not
 * overlapping). */
#define NE_STREAM_PROBE_MAX_IN 8
#define NE_STREAM_PROBE_MAX_OUT 4
int ne_stream_probe_f64(ne_ctx* ctx, const double* const* in, int32_t n_in, double* const* out, int32_t n_out,
                        int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NUMERICALEARTH_HIP_TESTING_H */
