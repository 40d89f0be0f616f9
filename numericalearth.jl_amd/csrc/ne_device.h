// ne_device.h — device-side building blocks shared by the HIP kernels.
//
// Formulas follow the reference expression they restate (cited per function), and the library
// is compiled with -ffp-contract=off. The general FP64 flux instance keeps the reference's
// evaluation order with the math library's elementary functions, apart from the saturation
// vapour pressure (one exponential of a sum, with ne_log, for a power times an exponential,
// within a few ulp). The instance for the default
// Edson constants ("Fast", and "Consts" with the default surface constants as literals) departs
// from it at the ulp level: its own logarithm ne_log, arctangent ne_atan_ge1, cube root
// ne_cbrt_ge1, square root ne_sqrt and exponential ne_exp (each within 2 ulp), A·exp(−b·log R★) for A/R★^b, x·√x for (1 + x)^{3/2}, shared subexpressions of ψ_m and
// ψ_s, and a degree-7 Taylor polynomial of ψ at |ζ| ≤ 5·10⁻⁴; with the default surface constants
// also logarithms of quotients as differences of logarithms, ζ as products with 1/L, and its
// remaining quotients as products with a reciprocal refined by two Newton steps (quotient).
// Those differences stay far inside the 1e-12 comparison with
// the oracle, but near a stop threshold they can move a column's fixed point by one iteration
// (its fluxes then differ by that iteration's drift, below the stop tolerance). The WENO
// reconstructions use fused multiply-adds and a reciprocal-based division (stated there).
#pragma once
#include <string.h>

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "numericalearth_hip.h"

#define NE_CELSIUS_TO_KELVIN 273.15  // src/EarthSystemModels/components.jl:10
#define NE_WENO_EPS 1e-8

namespace ne {

// ---------------------------------------------------------------------------
// Views
// ---------------------------------------------------------------------------

template <typename T>
struct View3 {
    T* __restrict__ p;
    int64_t sx, sy, sz;
    __host__ __device__ T& operator()(int i, int j, int k) const {
        return p[(int64_t)i * sx + (int64_t)j * sy + (int64_t)k * sz];
    }
    __host__ __device__ T& at2(int i, int j) const { return p[(int64_t)i * sx + (int64_t)j * sy]; }
};

template <typename T>
__host__ __device__ inline View3<const T> cview(const View3<T>& v) {
    View3<const T> r;
    r.p = v.p;
    r.sx = v.sx;
    r.sy = v.sy;
    r.sz = v.sz;
    return r;
}

template <typename T>
__host__ inline View3<T> view(const ne_view& v) {
    View3<T> r;
    r.p = static_cast<T*>(v.data);
    r.sx = v.sx;
    r.sy = v.sy;
    r.sz = v.sz;
    return r;
}

// Julia's min/max on floats: NaN-propagating, -0.0 < +0.0. Written as selects without early
// returns: branches here become divergent blocks that assign a literal operand, for which the
// compiler holds 64-bit constants in (spilled) vector registers inside the flux iteration.
template <typename T>
__device__ __forceinline__ T jmin(T a, T b) {
    T r = a < b ? a : b;                                  // b when b < a
    r = a == b ? (__builtin_signbit(a) ? a : b) : r;      // ±0: the negative one
    return (a != a || b != b) ? a + b : r;               // NaN
}
template <typename T>
__device__ __forceinline__ T jmax(T a, T b) {
    T r = a > b ? a : b;
    r = a == b ? (__builtin_signbit(a) ? b : a) : r;
    return (a != a || b != b) ? a + b : r;
}

__host__ __device__ __forceinline__ int wrapi(int i, int n) {
    int r = i % n;
    return r < 0 ? r + n : r;
}

// wrapi for the indices grids and source fields use, −n ≤ i < 2n: one conditional add or
// subtract in place of a remainder by a run-time n (about 30 instructions); any other i takes
// the remainder, in a branch of its own.
__host__ __device__ __forceinline__ int wrap_near(int i, int n) {
    const int r = i < 0 ? i + n : (i >= n ? i - n : i);
    if (r >= 0 && r < n) return r;
    return wrapi(i, n);
}

// Grid geometry as seen by kernels (device pointers, passed by value).
struct GridDev {
    int Nx, Ny, Nz;
    int tx, ty;  // topology
    int advection;
    const double* __restrict__ dz;
    const double* __restrict__ zc;
    const double* __restrict__ dy_xface;
    const double* __restrict__ dx_yface;
    const double* __restrict__ Az;
    const double* __restrict__ lambda_c;
    const double* __restrict__ phi_c;
    const int32_t* __restrict__ kmin;
    const double* __restrict__ zf;
    int ring;   // 1: the per-phase surface entry points also write the ring i = -1, Nx, j = -1, Ny

    // Column index after periodic wrapping; -1 outside a bounded direction.
    __device__ __forceinline__ int64_t column(int i, int j) const {
        if (tx == NE_PERIODIC) i = wrap_near(i, Nx);
        else if (i < 0 || i >= Nx) return -1;
        if (ty == NE_PERIODIC) j = wrap_near(j, Ny);
        else if (j < 0 || j >= Ny) return -1;
        return (int64_t)i + (int64_t)j * Nx;
    }
    // Complement of inactive_node at (Center, Center, Center).
    __device__ __forceinline__ bool active(int i, int j, int k) const {
        if (k < 0 || k >= Nz) return false;
        int64_t c = column(i, j);
        return c >= 0 && k >= kmin[c];
    }
    __device__ __forceinline__ int wx(int i) const { return tx == NE_PERIODIC ? wrap_near(i, Nx) : i; }
    __device__ __forceinline__ int wy(int j) const { return ty == NE_PERIODIC ? wrap_near(j, Ny) : j; }
};

__host__ inline GridDev grid_dev(const ne_grid& g) {
    GridDev d;
    d.Nx = g.Nx; d.Ny = g.Ny; d.Nz = g.Nz;
    d.tx = g.topology_x; d.ty = g.topology_y;
    d.advection = g.advection;
    d.dz = g.dz; d.zc = g.z_center; d.dy_xface = g.dy_xface; d.dx_yface = g.dx_yface;
    d.Az = g.Az; d.lambda_c = g.lambda_c; d.phi_c = g.phi_c; d.kmin = g.kmin; d.zf = g.z_face;
    d.ring = (g.halo >= 1 && !(g.Nx == 1 && g.Ny == 1)) ? 1 : 0;
    return d;
}

// ---------------------------------------------------------------------------
// Thermodynamics (Thermodynamics.jl forms; src/Atmospheres/thermodynamic_parameters.jl)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double Rv_over_Rd(const ne_params& p) { return p.dry_air_molar_mass / p.water_molar_mass; }
__device__ __forceinline__ double R_v(const ne_params& p) { return p.gas_constant / p.water_molar_mass; }
__device__ __forceinline__ double R_d(const ne_params& p) { return p.gas_constant / p.dry_air_molar_mass; }
__device__ __forceinline__ double cp_d(const ne_params& p) { return R_d(p) / p.dry_air_adiabatic_exponent; }

__device__ __forceinline__ double gas_constant_air(const ne_params& p, double q) {
    double eps = Rv_over_Rd(p);
    return R_d(p) * ((1.0 + (eps - 1.0) * q) - eps * (0.0 + 0.0));
}
__device__ __forceinline__ double cp_m(const ne_params& p, double q) {
    double cpd = cp_d(p);
    return cpd + (p.water_vapor_heat_capacity - cpd) * q;
}
__device__ __forceinline__ double air_density(const ne_params& p, double T, double pa, double q) {
    return pa / (gas_constant_air(p, q) * T);
}
__device__ __forceinline__ double latent_heat_vapor(const ne_params& p, double T) {
    double dcp = p.water_vapor_heat_capacity - p.liquid_water_heat_capacity;
    return p.reference_vaporization_enthalpy + dcp * (T - p.reference_temperature);
}
__device__ __forceinline__ double virtual_temperature(const ne_params& p, double T, double q) {
    return gas_constant_air(p, q) / R_d(p) * T;
}
// a/b for b > 0 normal: hardware reciprocal, two Newton steps, one product.
__device__ __forceinline__ double quotient(double a, double b) {
#pragma clang fp contract(fast)
    double r = __builtin_amdgcn_rcp(b);
    double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    return a * r;
}

// Natural logarithm of a positive, finite, normal double in about 30 vector instructions (the
// library's log takes about 100): x = 2^e·m with m in [√½, √2), log x = e·ln 2 + 2·atanh(s),
// s = (m − 1)/(m + 1), |s| ≤ 0.1716 — the argument reduction of Cody & Waite (Software Manual
// for the Elementary Functions, 1980) — with atanh summed as its Taylor series to s¹⁹, whose
// remainder is below 2⁻⁵⁵ of the result. m − 1 is exact; m + 1 and the quotient (a reciprocal
// refined by two Newton steps, 6 instructions in place of a division's 11) round, and e·ln 2 is
// added in one fused operation with ln 2 carried in two parts, so the result is within 2 ulp. Zero, negative, infinite and NaN arguments are not handled: the flux solve
// passes none.
__device__ __forceinline__ double ne_log(double x) {
    double m = __builtin_amdgcn_frexp_mant(x);   // [0.5, 1)
    int e = __builtin_amdgcn_frexp_exp(x);
    if (m < 0.7071067811865476) {
        m += m;
        e -= 1;
    }
    const double s = quotient(m - 1.0, m + 1.0);
    const double z = s * s;
    double q = 2.0 / 19.0;
    q = fma(q, z, 2.0 / 17.0);
    q = fma(q, z, 2.0 / 15.0);
    q = fma(q, z, 2.0 / 13.0);
    q = fma(q, z, 2.0 / 11.0);
    q = fma(q, z, 2.0 / 9.0);
    q = fma(q, z, 2.0 / 7.0);
    q = fma(q, z, 2.0 / 5.0);
    q = fma(q, z, 2.0 / 3.0);
    const double lm = fma(s * z, q, s + s);      // log m = 2s + 2s³/3 + … + 2s¹⁹/19
    const double de = (double)e;
    return fma(de, 0.6931471805599453, fma(de, 2.3190468138462996e-17, lm));
}

// saturation_vapor_pressure, liquid (interface_states.jl:63): pₜᵣ (T/Tₜᵣ)^(Δc/Rᵥ) e^b, evaluated
// as pₜᵣ e^((Δc/Rᵥ)·log(T/Tₜᵣ) + b) — one logarithm (ne_log) and one exponential in place of the
// library's pow (≈ 215 instructions) and a second exponential; both exponents are of size
// below 3, so the result agrees to a few ulp. T is a positive, finite temperature.
__device__ __forceinline__ double saturation_vapor_pressure(const ne_params& p, double T) {
    double dcp = p.water_vapor_heat_capacity - p.liquid_water_heat_capacity;
    double Rv = R_v(p);
    double Ttr = p.triple_point_temperature;
    const double a = dcp / Rv * ne_log(T / Ttr);
    const double b = (p.reference_vaporization_enthalpy - dcp * p.reference_temperature) / Rv * (1.0 / Ttr - 1.0 / T);
    return p.triple_point_pressure * exp(a + b);
}
// surface_specific_humidity(::ImpureSaturationSpecificHumidity) (interface_states.jl:55-74)
__device__ __forceinline__ double surface_specific_humidity(const ne_params& p, double pa, double Ts) {
    if (p.specific_humidity == NE_HUMIDITY_FIXED) return p.fixed_specific_humidity;
    double pv = p.water_mole_fraction * saturation_vapor_pressure(p, Ts);
    double edv = 1.0 / Rv_over_Rd(p);
    pv = jmin(pv, 0.999 * pa);
    return edv * pv / (pa - (1.0 - edv) * pv);
}

// ---------------------------------------------------------------------------
// Stability functions (similarity_theory_turbulent_fluxes.jl:501-532, :586-618).
// Only the branch selected by sign(ζ) is evaluated: the reference computes both and
// selects with ifelse, which gives the same value.
// ---------------------------------------------------------------------------

__device__ __forceinline__ double psi_momentum(const ne_params& p, double z) {
    if (p.stability_functions == NE_STABILITY_NONE) return 0.0;
    const double* c = p.edson_momentum;
    if (z < 0.0) {
        double Am = c[5], Bm = c[6], Cm = c[7], Dm = c[8], Em = c[9], Fm = c[10];
        double zm = jmin(0.0, z);
        double f1 = sqrt(sqrt(1.0 - Am * zm));
        double psi_m1 = Bm * log((1.0 + f1) / Bm) + log((1.0 + f1 * f1) / Bm) - Bm * atan(f1) + Cm;
        double f2 = cbrt(1.0 - Dm * zm);
        double psi_m2 = Em / 2.0 * log((1.0 + f2 + f2 * f2) / Em) - sqrt(Em) * atan((1.0 + 2.0 * f2) / sqrt(Em)) + Fm;
        double f = zm * zm / (1.0 + zm * zm);
        return (1.0 - f) * psi_m1 + f * psi_m2;
    } else {
        double zmax = c[0], Ap = c[1], Bp = c[2], Cp = c[3], Dp = c[4];
        double zp = jmax(0.0, z);
        double dz = jmin(zmax, Ap * zp);
        return -Bp * zp - Cp * (zp - Dp) * exp(-dz) - Cp * Dp;
    }
}

__device__ __forceinline__ double psi_scalar(const ne_params& p, double z) {
    if (p.stability_functions == NE_STABILITY_NONE) return 0.0;
    const double* c = p.edson_scalar;
    if (z < 0.0) {
        double Am = c[6], Bm = c[7], Cm = c[8], Dm = c[9], Em = c[10], Fm = c[11];
        double zm = jmin(0.0, z);
        double f1 = sqrt(1.0 - Am * zm);
        double psi_m1 = Bm * log((1.0 + f1) / Bm) + Cm;
        double f2 = cbrt(1.0 - Dm * zm);
        double psi_m2 = Em / 2.0 * log((1.0 + f2 + f2 * f2) / Em) - sqrt(Em) * atan((1.0 + 2.0 * f2) / sqrt(Em)) + Fm;
        double f = zm * zm / (1.0 + zm * zm);
        return (1.0 - f) * psi_m1 + f * psi_m2;
    } else {
        double zmax = c[0], Ap = c[1], Bp = c[2], Cp = c[3], Dp = c[4], Ep = c[5];
        double zp = jmax(0.0, z);
        double dz = jmin(zmax, Ap * zp);
        return -pow(1.0 + Bp * zp, Cp) - Bp * (zp - Dp) * exp(-dz) - Ep;
    }
}

// ---------------------------------------------------------------------------
// Stability functions with the reference's default Edson constants, arranged so that the
// momentum and scalar functions at the same ζ share their common subexpressions:
// √(1−15ζ) (its square root is the momentum's fourth root), log((1+√(1−15ζ))/2), the
// blend weight ζ²/(1+ζ²), and the stable-branch exp(−min(ζmax, 0.35ζ)). Near ζ = 0
// (|ζ| ≤ NE_PSI_SERIES_RADIUS, which is where ψ(ℓ/L) at the roughness length almost
// always falls) the functions are Taylor polynomials (ne_psi_series.h), accurate to
// < 1e-17 there. Same values as psi_momentum / psi_scalar up to rounding.
// ---------------------------------------------------------------------------

#include "ne_psi_series.h"
#include "ne_atan_poly.h"
#define NE_PSI_SERIES_RADIUS 5e-4

// Elementary functions by precision: FP64 uses the math library (the Fast flux instance its
// own ne_log above, ne_atan_ge1 and ne_cbrt_ge1 below); the FP32 pre-iteration of the mixed form
// uses the hardware log₂/exp₂-based approximations.

// Arctangent of x ≥ 1 in about 35 vector instructions with one division (the library's atan
// takes about 80 with two). The stability functions' arctangents have such arguments:
// (1 − 15ζ)^¼ and (1 + 2(1 − cζ)^⅓)/√3 with ζ < 0. atan x = π/2 − atan(1/x), and for
// x < 1 + √2 also = π/4 − atan((1 − x)/(1 + x)); the branch taken keeps the quotient u within
// |u| ≤ tan(π/8), where atan u = u + u·w·P(w), w = u² (ne_atan_poly.h: 0.07 ulp). 1 − x is
// exact for 1 ≤ x ≤ 1 + √2, the quotient rounds once, and π/4 and π/2 are added in two parts,
// so the result is within 2 ulp (test_elementary_gpu.py: against numpy's arctan). x = +∞
// gives π/2; arguments below 1 are outside its range.
__device__ __forceinline__ double ne_atan_ge1(double x) {
    const bool near = x < 2.4142135623730951;
    const double u = (near ? 1.0 - x : 1.0) / (near ? 1.0 + x : x);
    const double w = u * u;
    const double a = fma(u * w, ne_atan_poly(w), u);
    return near ? 0.78539816339744828 + (3.0616169978683830e-17 - a)
                : 1.5707963267948966 + (6.1232339957367660e-17 - a);
}

// Cube root of x ≥ 1 in 14 vector instructions (the library's cbrt takes about 31). The
// stability functions' cube roots have such arguments: 1 − cζ with ζ < 0. A first value from the
// hardware's FP32 log₂ and exp₂, y₀ = 2^(log₂x / 3) (relative error near 10⁻⁷), then one step of
// Halley's iteration for y³ = x in FP64, y = y₀ − y₀(y₀³ − x)/(2y₀³ + x), which cubes the
// relative error: the correction is of size 10⁻⁷·y₀, so its reciprocal needs 10⁻⁹ only — the
// hardware's approximation (about 2⁻²⁶ alone, which left 32 ulp) with one Newton step — and
// y₀³ − x is one fused operation; the result is within 1 ulp
// (test_elementary_gpu.py: against numpy's cbrt). Arguments below 1 or from 2¹²⁸ (beyond FP32)
// are outside its range.
__device__ __forceinline__ double ne_cbrt_ge1(double x) {
    const float l2 = __builtin_amdgcn_logf((float)x);
    const double y0 = (double)__builtin_amdgcn_exp2f(l2 * (1.0f / 3.0f));
    const double y2 = y0 * y0;
    const double t = fma(y2, y0, -x);
    const double d = fma(y2 + y2, y0, x);
    double r = __builtin_amdgcn_rcp(d);
    r = fma(r, fma(-d, r, 1.0), r);
    return fma(-(y0 * t), r, y0);
}

// Square root of a positive, finite, normal double in 12 vector instructions (the library's
// correctly rounded sqrt takes about 21): the hardware's reciprocal square root (about 2⁻²⁶),
// two Goldschmidt steps on g ≈ √x and h ≈ 1/(2√x), then g + h(x − g²) with the residual x − g²
// formed in one fused operation; within 1 ulp (test_elementary_gpu.py: against numpy's sqrt).
// Zero, negative, infinite and NaN arguments are not handled: the flux solve passes none.
__device__ __forceinline__ double ne_sqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y, h = 0.5 * y;
    double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    return fma(fma(-g, g, x), h, g);
}

// e^x for |x| ≤ 700 in about 19 vector instructions (the library's exp takes about 41): x = k ln 2 + r
// with k the nearest integer to x/ln 2 and |r| ≤ ln(2)/2, r formed by two fused operations with
// ln 2 in two parts; e^r by its Taylor polynomial to r¹³ (remainder below 0.06 ulp), then scaled by
// 2^k. Within 2 ulp (test_elementary_gpu.py: against numpy's exp); the flux solve's arguments lie
// within ±60.
__device__ __forceinline__ double ne_exp(double x) {
    const double k = rint(x * 1.4426950408889634);
    double r = fma(-k, 0.6931471805599453, x);
    r = fma(-k, 2.3190468138462996e-17, r);
    double q = 1.6059043836821613e-10;
    q = fma(q, r, 2.08767569878681e-09);
    q = fma(q, r, 2.505210838544172e-08);
    q = fma(q, r, 2.755731922398589e-07);
    q = fma(q, r, 2.7557319223985893e-06);
    q = fma(q, r, 2.48015873015873e-05);
    q = fma(q, r, 0.0001984126984126984);
    q = fma(q, r, 0.001388888888888889);
    q = fma(q, r, 0.008333333333333333);
    q = fma(q, r, 0.041666666666666664);
    q = fma(q, r, 0.16666666666666666);
    q = fma(q, r, 0.5);
    q = fma(q, r, 1.0);
    q = fma(q, r, 1.0);
    return __builtin_amdgcn_ldexp(q, (int)k);
}

// FP32 functions of the mixed form's flux iteration: the hardware's approximations (log₂, exp₂,
// square root and reciprocal, each within about 1 FP32 ulp for normal arguments) without the
// library's handling of denormal and special arguments, which takes 10–17 instructions per call
// where these take 1–3. The iteration passes positive, normal arguments (a cube root of zero
// gives zero); its iterate is accurate to its stop criterion, a few FP32 ulp of u★ + θ★ + q★.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fcbrt(float x) { return __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * (1.0f / 3.0f)); }

__device__ __forceinline__ double tlog(double x) { return ne_log(x); }
__device__ __forceinline__ float tlog(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
__device__ __forceinline__ double texp(double x) { return ne_exp(x); }
__device__ __forceinline__ float texp(float x) { return __expf(x); }
__device__ __forceinline__ double tatan_ge1(double x) { return ne_atan_ge1(x); }
__device__ __forceinline__ float tatan_ge1(float x) { return atanf(x); }
__device__ __forceinline__ double tcbrt_ge1(double x) { return ne_cbrt_ge1(x); }
__device__ __forceinline__ float tcbrt_ge1(float x) { return fcbrt(x); }
__device__ __forceinline__ double tsqrt(double x) { return ne_sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename T>
__device__ __forceinline__ T tdiv(T a, T b) { return quotient(a, b); }
template <>
__device__ __forceinline__ float tdiv<float>(float a, float b) { return a * frcp(b); }

template <typename T>
__device__ __forceinline__ T psi_m_default(T z) {
    if (fabs(z) <= (T)NE_PSI_SERIES_RADIUS) return z < (T)0 ? psi_series_momentum_unstable(z) : psi_series_momentum_stable(z);
    const T pi = (T)3.14159265358979323846, s3 = (T)1.7320508075688772935;
    if (z < (T)0) {
        T xs = tsqrt((T)1 - (T)15 * z), xm = tsqrt(xs);
        T psi1 = (T)2 * tlog(((T)1 + xm) * (T)0.5) + tlog(((T)1 + xs) * (T)0.5) - (T)2 * tatan_ge1(xm) + pi / (T)2;
        T f2 = tcbrt_ge1((T)1 - (T)10.15 * z);
        T psi2 = (T)1.5 * tlog(((T)1 + f2 + f2 * f2) * (T)(1.0 / 3.0)) - s3 * tatan_ge1(((T)1 + (T)2 * f2) * ((T)1 / s3)) + pi / s3;
        T f = tdiv(z * z, (T)1 + z * z);
        return ((T)1 - f) * psi1 + f * psi2;
    }
    T e = texp(-fmin((T)50, (T)0.35 * z));
    return (T)-0.7 * z - (T)0.75 * (z - (T)(5.0 / 0.35)) * e - (T)(0.75 * (5.0 / 0.35));
}

template <typename T>
__device__ __forceinline__ T psi_s_default(T z) {
    if (fabs(z) <= (T)NE_PSI_SERIES_RADIUS) return z < (T)0 ? psi_series_scalar_unstable(z) : psi_series_scalar_stable(z);
    const T pi = (T)3.14159265358979323846, s3 = (T)1.7320508075688772935;
    if (z < (T)0) {
        T xs = tsqrt((T)1 - (T)15 * z);
        T psi1 = (T)2 * tlog(((T)1 + xs) * (T)0.5);
        T f2 = tcbrt_ge1((T)1 - (T)34.15 * z);
        T psi2 = (T)1.5 * tlog(((T)1 + f2 + f2 * f2) * (T)(1.0 / 3.0)) - s3 * tatan_ge1(((T)1 + (T)2 * f2) * ((T)1 / s3)) + pi / s3;
        T f = tdiv(z * z, (T)1 + z * z);
        return ((T)1 - f) * psi1 + f * psi2;
    }
    T e = texp(-fmin((T)50, (T)0.35 * z));
    T x = (T)1 + (T)(2.0 / 3.0) * z;
    return -(x * tsqrt(x)) - (T)(2.0 / 3.0) * (z - (T)14.28) * e - (T)8.525;
}

// ψ_m(ζ) and ψ_s(ζ) together (the height term h/L of the two profiles).
template <typename T>
__device__ __forceinline__ void psi_pair_default(T z, T& pm, T& ps) {
    if (fabs(z) <= (T)NE_PSI_SERIES_RADIUS) {
        bool u = z < (T)0;
        pm = u ? psi_series_momentum_unstable(z) : psi_series_momentum_stable(z);
        ps = u ? psi_series_scalar_unstable(z) : psi_series_scalar_stable(z);
        return;
    }
    const T pi = (T)3.14159265358979323846, s3 = (T)1.7320508075688772935;
    if (z < (T)0) {
        T xs = tsqrt((T)1 - (T)15 * z), xm = tsqrt(xs);
        T l2 = tlog(((T)1 + xs) * (T)0.5);
        T psi1m = (T)2 * tlog(((T)1 + xm) * (T)0.5) + l2 - (T)2 * tatan_ge1(xm) + pi / (T)2;
        T psi1s = (T)2 * l2;
        T fm = tcbrt_ge1((T)1 - (T)10.15 * z), fs = tcbrt_ge1((T)1 - (T)34.15 * z);
        T psi2m = (T)1.5 * tlog(((T)1 + fm + fm * fm) * (T)(1.0 / 3.0)) - s3 * tatan_ge1(((T)1 + (T)2 * fm) * ((T)1 / s3)) + pi / s3;
        T psi2s = (T)1.5 * tlog(((T)1 + fs + fs * fs) * (T)(1.0 / 3.0)) - s3 * tatan_ge1(((T)1 + (T)2 * fs) * ((T)1 / s3)) + pi / s3;
        T f = tdiv(z * z, (T)1 + z * z);
        pm = ((T)1 - f) * psi1m + f * psi2m;
        ps = ((T)1 - f) * psi1s + f * psi2s;
        return;
    }
    T e = texp(-fmin((T)50, (T)0.35 * z));
    pm = (T)-0.7 * z - (T)0.75 * (z - (T)(5.0 / 0.35)) * e - (T)(0.75 * (5.0 / 0.35));
    T x = (T)1 + (T)(2.0 / 3.0) * z;
    ps = -(x * tsqrt(x)) - (T)(2.0 / 3.0) * (z - (T)14.28) * e - (T)8.525;
}

// True when the flux formulation is SimilarityTheoryFluxes with BulkTemperature and the
// reference's default Edson constants with the logarithmic similarity profile (the case
// the *_default functions above implement).
__host__ inline bool edson_default_constants(const ne_params& p) {
    if (p.turbulent_fluxes != NE_FLUXES_SIMILARITY_THEORY || p.interface_temperature != NE_TEMPERATURE_BULK)
        return false;
    const double pi = 3.14159265358979323846;
    const double em[11] = {50.0, 0.35, 0.7, 0.75, 5.0 / 0.35, 15.0, 2.0, pi / 2.0, 10.15, 3.0, pi / sqrt(3.0)};
    const double es[12] = {50.0, 0.35, 2.0 / 3.0, 3.0 / 2.0, 14.28, 8.525, 15.0, 2.0, 0.0, 34.15, 3.0, pi / sqrt(3.0)};
    if (p.stability_functions != NE_STABILITY_EDSON || p.similarity_form != NE_SIMILARITY_LOG) return false;
    for (int n = 0; n < 11; n++)
        if (p.edson_momentum[n] != em[n]) return false;
    for (int n = 0; n < 12; n++)
        if (p.edson_scalar[n] != es[n]) return false;
    return true;
}

// The reference's default surface-layer constants: von Kármán constant and gustiness
// (similarity_theory_turbulent_fluxes.jl:174-199, :45-98), roughness lengths
// (roughness_lengths.jl:93-146, :197-246), stop criterion, humidity and velocity difference
// (component_interfaces.jl:353-358), surface-layer and boundary-layer heights and the
// thermodynamic constants (Atmospheres/thermodynamic_parameters.jl:30-258). ne_params_default
// sets them through this function; a flux-solve instance with them compiled in (so that the
// fixed-point loop keeps no register for them) runs when a model's parameters equal them.
__host__ __device__ inline void set_default_surface_constants(ne_params& q) {
    q.von_karman_constant = 0.4;
    q.convective_gustiness = 1;
    q.mesoscale_gustiness = 0;
    q.gustiness_parameter = 1.2;
    q.minimum_gustiness = 0.01;
    q.mesoscale_velocity = 0.0;
    q.momentum_roughness = NE_ROUGHNESS_EDSON;
    q.temperature_roughness = NE_ROUGHNESS_EDSON;
    q.water_vapor_roughness = NE_ROUGHNESS_EDSON;
    q.roughness_gravitational_acceleration = 9.80665;
    q.air_kinematic_viscosity = 1.5e-5;
    q.charnock_parameter = 0.02;
    q.smooth_wall_parameter = 0.11;
    q.maximum_momentum_roughness = 1.0;
    q.reynolds_scaling_A = 5.85e-5;
    q.reynolds_scaling_b = 0.72;
    q.maximum_scalar_roughness = 1.6e-4;
    q.zero_plane_displacement = 0.0;
    q.stop_criteria = NE_STOP_CONVERGENCE;
    q.maxiter = 100;
    q.tolerance = 1e-8;
    q.specific_humidity = NE_HUMIDITY_IMPURE_SATURATION;
    q.velocity_difference = NE_VELOCITY_RELATIVE;
    q.water_mole_fraction = 0.98;
    q.surface_layer_height = 10.0;
    q.boundary_layer_height = 512.0;
    q.gravitational_acceleration = 9.80665;
    q.gas_constant = 8.3144598;
    q.dry_air_molar_mass = 0.02897;
    q.water_molar_mass = 0.018015;
    q.dry_air_adiabatic_exponent = 2.0 / 7.0;
    q.water_vapor_heat_capacity = 1859.0;
    q.liquid_water_heat_capacity = 4181.0;
    q.reference_vaporization_enthalpy = 2500800.0;
    q.reference_temperature = 273.16;
    q.triple_point_temperature = 273.16;
    q.triple_point_pressure = 611.657;
}

// True when every constant set_default_surface_constants sets already holds its default
// (compared bit for bit, so −0.0 for a zero default counts as different).
__host__ inline bool default_surface_constants(const ne_params& p) {
    ne_params q;
    memcpy(&q, &p, sizeof q);
    set_default_surface_constants(q);
    return memcmp(&q, &p, sizeof q) == 0;
}

// Flux-solve instance of a parameter set: 0 the general path, 1 the reference's default
// Edson constants (shared ψ), 2 those and the default surface constants compiled in.
__host__ inline int flux_path(const ne_params& p) {
    if (!edson_default_constants(p)) return 0;
    return default_surface_constants(p) ? 2 : 1;
}

// similarity_profile (similarity_theory_turbulent_fluxes.jl:242-253)
__device__ __forceinline__ double similarity_profile(const ne_params& p, bool scalar, double h, double l, double L) {
    double zeta = h / L;
    double psih = scalar ? psi_scalar(p, zeta) : psi_momentum(p, zeta);
    if (p.similarity_form == NE_SIMILARITY_COARE_LOG) return log(h / l) - psih;
    double psil = scalar ? psi_scalar(p, l / L) : psi_momentum(p, l / L);
    return log(h / l) - psih + psil;
}

// roughness_length(::MomentumRoughnessLength) (roughness_lengths.jl:197-210)
__device__ __forceinline__ double momentum_roughness(const ne_params& p, double ustar) {
    if (p.momentum_roughness == NE_ROUGHNESS_CONSTANT) return p.momentum_roughness_length;
    double nu = p.air_kinematic_viscosity;
    double g = p.roughness_gravitational_acceleration;
    double Cg = p.charnock_parameter;
    double Cnu = p.smooth_wall_parameter;
    double lW = Cg * (ustar * ustar) / g;
    double lR = Cnu == 0.0 ? 0.0 : Cnu * nu / ustar;
    return jmin(lW + lR, p.maximum_momentum_roughness);
}

// roughness_length(::ScalarRoughnessLength) (roughness_lengths.jl:231-246)
__device__ __forceinline__ double scalar_roughness(const ne_params& p, int kind, double constant, double lu, double ustar) {
    if (kind == NE_ROUGHNESS_CONSTANT) return constant;
    double Rstar = lu * ustar / p.air_kinematic_viscosity;
    double ls = Rstar == 0.0 ? 0.0 : p.reynolds_scaling_A / pow(Rstar, p.reynolds_scaling_b);
    return jmin(ls, p.maximum_scalar_roughness);
}

// vsgs² (similarity_theory_turbulent_fluxes.jl:88-98)
__device__ __forceinline__ double gustiness_squared(const ne_params& p, double ustar, double bstar, double hbl) {
    double v2 = 0.0;
    if (p.convective_gustiness) {
        double Jb = -ustar * bstar;
        double Ug = jmax(p.minimum_gustiness, p.gustiness_parameter * cbrt(jmax(0.0, Jb) * hbl));
        v2 = Ug * Ug;
    }
    if (p.mesoscale_gustiness) v2 = v2 + p.mesoscale_velocity * p.mesoscale_velocity;
    return v2;
}

// Result of one column of the air-sea flux solve.
struct ColumnFluxes {
    double Qv, Qt, Jv, rtx, rty, us, ts, qs, Ts_C;
    int iterations;
    bool converged;
    bool deferred;   // Defer: the FP32 iteration did not finish the column (nothing else is set)
};

// ---------------------------------------------------------------------------
// Coefficient-based fluxes and the skin temperature (the formulations other than
// SimilarityTheoryFluxes + BulkTemperature). FP64, one iteration as the reference
// writes it: the skin temperature, then qₛ, Δθ, Δq from it, then the scales.
// ---------------------------------------------------------------------------

// x^n, n > 3, as Julia's Float64^Integer: power by squaring in double-double with
// error-free products (fma), rounded once.
__device__ __forceinline__ double pow_int(double x, int n) {
    double yh = 1.0, yl = 0.0, xh = x, xl = 0.0;
    while (n > 0) {
        if (n & 1) {
            double ph = yh * xh;
            double pl = fma(yh, xh, -ph) + (yh * xl + yl * xh);
            yh = ph + pl;
            yl = pl - (yh - ph);
        }
        n >>= 1;
        if (n > 0) {
            double sh = xh * xh;
            double sl = fma(xh, xh, -sh) + 2.0 * xh * xl;
            xh = sh + sl;
            xl = sl - (xh - sh);
        }
    }
    return yh + yl;
}

// latent_heat_sublim = ℒₛ₀ + (cₚᵥ − cₚᵢ)(T − T₀) (taken at interface_states.jl:544)
__device__ __forceinline__ double latent_heat_sublim(const ne_params& p, double T) {
    double dcp = p.water_vapor_heat_capacity - p.water_ice_heat_capacity;
    return p.reference_sublimation_enthalpy + dcp * (T - p.reference_temperature);
}

// PolynomialNeutralDragCoefficient (coefficient_based_turbulent_fluxes.jl:48-54)
__device__ __forceinline__ double neutral_drag(const ne_params& p, double U) {
    U = jmax(U, p.minimum_wind_speed);
    if (!(U < p.high_wind_speed_threshold)) return p.high_wind_drag_coefficient;
    const double* c = p.drag_polynomial;
    return (c[0] / U + c[1] + c[2] * U - c[3] * pow_int(U, 6)) / 1000.0;
}

// SplitStabilityFunction(LinearStable, Paulson) of large_yeager_stability_functions
// (similarity_theory_turbulent_fluxes.jl:688-710, :742-752, :772-778); only the branch
// selected by sign(ζ) is evaluated.
__device__ __forceinline__ double psi_large_yeager(const ne_params& p, bool scalar, double zeta) {
    if (zeta > 0.0) {
        const double* ls = scalar ? p.linear_stable_scalar : p.linear_stable_momentum;
        return -ls[0] * jmin(zeta, ls[1]);
    }
    double a = scalar ? p.paulson_scalar : p.paulson_momentum[0];
    double z = sqrt(sqrt(1.0 - a * jmin(0.0, zeta)));
    if (scalar) return 2.0 * log((1.0 + z * z) / 2.0);
    return 2.0 * log((1.0 + z) / 2.0) + log((1.0 + z * z) / 2.0) + -2.0 * atan(z) + p.paulson_momentum[1];
}

// iterate_interface_fluxes(::CoefficientBasedFluxes) (coefficient_based_turbulent_fluxes.jl:346-371)
// with evaluate_coefficients (:269-340).
__device__ __forceinline__ void coefficient_scales(const ne_params& p, double du, double dv, double Ts, double qsurf,
                                                   double dtheta, double dq, double dh, double& us, double& ts,
                                                   double& qs) {
    const bool ly = p.transfer_coefficients == NE_TRANSFER_LARGE_YEAGER;
    double dU = jmax(sqrt(du * du + dv * dv), ly ? p.minimum_wind_speed : 0.0);
    double Cd, Ch, Cq;
    if (ly) {
        const double kappa = p.large_yeager_von_karman_constant;
        const double g = p.gravitational_acceleration;
        double Tv = virtual_temperature(p, Ts, qsurf);
        double delta = Rv_over_Rd(p) - 1.0;
        double bstar = g / Tv * (ts * (1.0 + delta * qsurf) + delta * Tv * qs);
        double L = bstar == 0.0 ? (double)INFINITY : us * us / (kappa * bstar);
        double zeta = dh / L;
        double psim = psi_large_yeager(p, false, zeta);
        double psih = psi_large_yeager(p, true, zeta);
        double lh = log(dh / p.large_yeager_reference_height);
        double Cdp = us == 0.0 ? neutral_drag(p, dU) : us * us / (dU * dU);
        double UN10 = jmax(dU / (1.0 + sqrt(Cdp) / kappa * (lh - psim)), p.minimum_wind_speed);
        double CdN = neutral_drag(p, UN10);
        double sCdN = sqrt(CdN);
        double ChN = sCdN / 1000.0 *
                     (zeta > 0.0 ? p.stable_heat_transfer_coefficient : p.unstable_heat_transfer_coefficient);
        double CqN = sCdN / 1000.0 * p.moisture_transfer_coefficient;
        double xim = sCdN / kappa * (lh - psim);
        Cd = CdN / ((1.0 + xim) * (1.0 + xim));
        double xih = sCdN / kappa * (lh - psih);
        double ratio = sqrt(Cd) / sCdN;
        Ch = ChN * ratio / (1.0 + ChN * xih);
        Cq = CqN * ratio / (1.0 + CqN * xih);
    } else {
        Cd = p.transfer_kind[0] == NE_COEFFICIENT_POLYNOMIAL ? neutral_drag(p, dU) : p.transfer_coefficient[0];
        Ch = p.transfer_kind[1] == NE_COEFFICIENT_POLYNOMIAL ? neutral_drag(p, dU) : p.transfer_coefficient[1];
        Cq = p.transfer_kind[2] == NE_COEFFICIENT_POLYNOMIAL ? neutral_drag(p, dU) : p.transfer_coefficient[2];
    }
    double sCd = sqrt(Cd);
    us = sCd * dU;
    ts = Cd == 0.0 ? 0.0 : Ch / sCd * dtheta;
    qs = Cd == 0.0 ? 0.0 : Cq / sCd * dq;
}

// flux_balance_temperature(::SkinTemperature{<:DiffusiveFlux}) (interface_states.jl:425-448)
// inside compute_interface_temperature(::SkinTemperature) (:536-577), constant diffusivity.
// Tm: the previous iterate's interface temperature; Ti: the top-cell temperature (K).
__device__ __forceinline__ double skin_temperature(const ne_params& p, double rhoa, double ca, double Li,
                                                   double thetaa, double Ti, double Tm, double us, double ts,
                                                   double qs, double sigma, double alpha, double eps, double SW,
                                                   double LW) {
    double Iup = sigma * eps * pow_int(Tm, 4);
    double Qd = -(1.0 - alpha) * SW - eps * LW;
    double QT = -rhoa * ca * us * ts;
    double Qv = -rhoa * Li * us * qs;
    double kappa = p.skin_diffusivity, delta = p.skin_depth;
    double lambda = 1.0 / (p.ocean_reference_density * p.ocean_heat_capacity);
    double JT = (Qv + Iup + Qd) * lambda;
    double dT = thetaa - Tm;
    double OmT = QT * lambda;
    double D = kappa * dT - OmT * delta;
    double Tstar = D == 0.0 ? Tm : (Ti * kappa * dT - (JT * dT + OmT * thetaa) * delta) / D;
    double x = Tstar - Ti, m = p.skin_max_temperature_change;
    return Ti + (x > m ? m : (x < -m ? -m : x));
}

// The fixed point of compute_interface_state (compute_interface_state.jl:31-58) for these
// formulations. Returns the scales and the interface temperature (K) through the references.
__device__ __noinline__ int flux_iterate_general(const ne_params& p, double ua, double va, double Ta, double pa,
                                                 double qa, double To, double su, double sv, double SW, double LW,
                                                 double& us, double& ts, double& qs, double& sT, bool& reached) {
    const double g = p.gravitational_acceleration;
    const double za = p.surface_layer_height;
    const bool converge = p.stop_criteria == NE_STOP_CONVERGENCE;
    const bool skin = p.interface_temperature == NE_TEMPERATURE_SKIN;
    const bool coeff = p.turbulent_fluxes == NE_FLUXES_COEFFICIENT_BASED;
    const double thetaa = Ta + g * za / cp_m(p, qa);  // surface_atmosphere_temperature (interface_states.jl:308-317)
    double rhoa = 0.0, Li = 0.0, ca = 0.0;
    double sigma = 0.0, alpha = 0.0, eps = 0.0;
    if (skin) {
        rhoa = air_density(p, Ta, pa, qa);
        ca = cp_m(p, qa);
        Li = latent_heat_sublim(p, Ta);
        if (p.radiation) { sigma = p.stefan_boltzmann_constant; alpha = p.ocean_albedo; eps = p.ocean_emissivity; }
        else { SW = 0.0; LW = 0.0; }
    }
    double du, dv;
    if (p.velocity_difference == NE_VELOCITY_RELATIVE) { du = ua - su; dv = va - sv; }
    else { du = ua; dv = va; }
    const double kappa = p.von_karman_constant;
    const double delta = Rv_over_Rd(p) - 1.0;
    int it = 0;
    for (;;) {
        double Ts = skin ? skin_temperature(p, rhoa, ca, Li, thetaa, To, sT, us, ts, qs, sigma, alpha, eps, SW, LW)
                         : sT;
        double qsurf = surface_specific_humidity(p, pa, Ts);
        double dq = qa - qsurf;
        double dtheta = thetaa - Ts;
        double nus, nts, nqs;
        if (coeff) {
            nus = us; nts = ts; nqs = qs;
            coefficient_scales(p, du, dv, Ts, qsurf, dtheta, dq, za, nus, nts, nqs);
        } else {
            // iterate_interface_fluxes(::SimilarityTheoryFluxes) (similarity:315-385)
            double Tv = virtual_temperature(p, Ts, qsurf);
            double bstar = g / Tv * (ts * (1.0 + delta * qsurf) + delta * Tv * qs);
            double U = sqrt(du * du + dv * dv + gustiness_squared(p, us, bstar, p.boundary_layer_height));
            double lu = momentum_roughness(p, us);
            double lq = scalar_roughness(p, p.water_vapor_roughness, p.water_vapor_roughness_length, lu, us);
            double lt = scalar_roughness(p, p.temperature_roughness, p.temperature_roughness_length, lu, us);
            double dh = jmax(za - p.zero_plane_displacement, 2.0 * lu);
            double L = bstar == 0.0 ? (double)INFINITY : us * us / (kappa * bstar);
            nus = kappa / similarity_profile(p, false, dh, lu, L) * U;
            nts = kappa / similarity_profile(p, true, dh, lt, L) * dtheta;
            nqs = kappa / similarity_profile(p, true, dh, lq, L) * dq;
        }
        it += 1;
        bool go;
        if (converge) {
            double drift = fabs(nus - us) + fabs(nts - ts) + fabs(nqs - qs);
            reached = it >= p.maxiter;
            go = !((drift < p.tolerance) | reached);
        } else {
            go = it < p.maxiter;
        }
        us = nus; ts = nts; qs = nqs; sT = Ts;
        if (!go) break;
    }
    return it;
}

// One column of _compute_atmosphere_ocean_interface_state! (atmosphere_ocean_fluxes.jl:80-197)
// with compute_interface_state (compute_interface_state.jl:31-58) and
// iterate_interface_fluxes(::SimilarityTheoryFluxes) (similarity:315-385).
// Quantities that the reference re-evaluates each iteration from unchanged inputs
// (qₛ, θᵃ, Δθ, Δq, 𝒯ₛ) are evaluated once: the values are identical.
// Fast = true (only for the default Edson constants, edson_default_constants) evaluates the
// two profiles with the shared/series stability functions above and R★^b as exp(b log R★).
// Mixed = true (the FP32 form) runs the same fixed point in FP32 with the hardware log/exp
// until its drift falls below NE_MIXED_DRIFT (a few times the FP32 rounding noise of
// u★ + θ★ + q★) and takes that iterate; only a column whose FP32 iteration does not get there
// within NE_MIXED_MAXITER iterations (or leaves the finite range) continues in FP64 from it
// with the reference's stop criterion. Defer = true (for kernels whose parameters take the FP32
// iteration) returns such a column as `deferred` instead of continuing: its FP64 iteration needs
// twice the registers of the FP32 one, so the kernel that defers runs at 6 wavefronts per SIMD
// instead of 4 and a second kernel finishes the few deferred columns with Defer = false, which
// gives them the same values. Finishing in FP64 as well (≈ 3 more FP64 iterations)
// changed neither the 200-step SST error against the FP64 model (9.8e-6 of the largest SST)
// nor the flux errors (2.4e-5 of their scale) at 1440×560×10, and cost 13 % of the step.
#ifndef NE_MIXED_DRIFT
#define NE_MIXED_DRIFT 2e-7f
#endif
#ifndef NE_MIXED_MAXITER
#define NE_MIXED_MAXITER 40
#endif

template <bool Fast, bool Mixed = false, bool Consts = false, bool Defer = false>
__device__ __forceinline__ ColumnFluxes flux_column(const ne_params& p_in, double ua, double va, double Ta,
                                                    double pa, double qa, double To_C, double uo, double vo,
                                                    bool inactive, double SW = 0.0, double LW = 0.0) {
    // Consts: the default surface constants as literals (the host has checked that p holds them)
    ne_params pc;
    if constexpr (Consts) {
        pc = p_in;
        set_default_surface_constants(pc);
    }
    const ne_params& p = Consts ? pc : p_in;
    const double g = p.gravitational_acceleration;
    const double za = p.surface_layer_height;
    const double hbl = p.boundary_layer_height;
    const double To = To_C + NE_CELSIUS_TO_KELVIN;

    double us = 1e-4, ts = 1e-4, qs = 1e-4;
    double su = uo, sv = vo, sT = To;
    int it = 0;
    const bool converge = p.stop_criteria == NE_STOP_CONVERGENCE;
    bool reached = false;

    const bool general = !Fast && (p.turbulent_fluxes != NE_FLUXES_SIMILARITY_THEORY ||
                                   p.interface_temperature != NE_TEMPERATURE_BULK);
    if (general && !inactive && (converge || p.maxiter > 0)) {
        it = flux_iterate_general(p, ua, va, Ta, pa, qa, To, su, sv, SW, LW, us, ts, qs, sT, reached);
    } else if (!inactive && (converge || p.maxiter > 0)) {
        // Iteration invariants: BulkTemperature keeps Tₛ = Tₒ, hence qₛ is fixed.
        const double Ts = To;
        const double qsurf = surface_specific_humidity(p, pa, Ts);
        const double dq = qa - qsurf;
        const double thetaa = Ta + g * za / cp_m(p, qa);
        const double dtheta = thetaa - Ts;
        const double Tv = virtual_temperature(p, Ts, qsurf);
        const double delta = Rv_over_Rd(p) - 1.0;
        const double onedq = 1.0 + delta * qsurf;
        const double dTv = delta * Tv;
        const double goTv = g / Tv;
        double du, dv;
        if (p.velocity_difference == NE_VELOCITY_RELATIVE) { du = ua - su; dv = va - sv; }
        else { du = ua; dv = va; }
        const double du2dv2 = du * du + dv * dv;
        const double kappa = p.von_karman_constant;
        const bool shared_scalar = p.temperature_roughness == p.water_vapor_roughness &&
                                   (p.temperature_roughness == NE_ROUGHNESS_EDSON ||
                                    p.temperature_roughness_length == p.water_vapor_roughness_length);
        const int maxit = p.maxiter;
        const double tol = p.tolerance;
        // LogForm (the default constants compiled in): the logarithms of the profiles' quotients
        // as differences of logarithms, log(h/ℓ) = log h − log ℓ, with log R★ = log ℓᵤ + log u★ −
        // log ν and log ℓq = min(log A − b·log R★, log ℓmax), so an iteration takes two
        // logarithms (ℓᵤ, u★) and no quotients where it took three of each, and the three
        // stability arguments as products with 1/L. The terms are of size ≤ 12 and the results
        // agree with the quotient forms to a few ulp of those terms. u★ > 0 (the iteration's first
        // guess, and κU over a positive profile with U ≥ the minimum gustiness).
        constexpr bool LogForm = Fast && Consts;
        const double zd = za - p.zero_plane_displacement;
        const double ln_nu = LogForm ? ne_log(p.air_kinematic_viscosity) : 0.0;
        const double ln_A = LogForm ? ne_log(p.reynolds_scaling_A) : 0.0;
        const double ln_lqmax = LogForm ? ne_log(p.maximum_scalar_roughness) : 0.0;
        const double ln_zd = LogForm ? ne_log(zd) : 0.0;
        const double ln2 = 0.6931471805599453;

        bool done32 = false;   // the FP32 fixed point converged: no FP64 iteration
        if (Mixed && Fast && converge && shared_scalar && p.momentum_roughness == NE_ROUGHNESS_EDSON &&
            p.water_vapor_roughness == NE_ROUGHNESS_EDSON) {
            const float fgoTv = (float)goTv, fonedq = (float)onedq, fdTv = (float)dTv;
            const float fdu2dv2 = (float)du2dv2, fdtheta = (float)dtheta, fdq = (float)dq;
            const float fkappa = (float)kappa, fhbl = (float)hbl;
            const float fCg_g = (float)(p.charnock_parameter / p.roughness_gravitational_acceleration);
            const float fCnu_nu = (float)(p.smooth_wall_parameter * p.air_kinematic_viscosity);
            const float finv_nu = (float)(1.0 / p.air_kinematic_viscosity);
            const float flmax = (float)p.maximum_momentum_roughness, fsmax = (float)p.maximum_scalar_roughness;
            const float fA = (float)p.reynolds_scaling_A, fb = (float)p.reynolds_scaling_b;
            const float fzd = (float)(za - p.zero_plane_displacement);
            const float fbeta = (float)p.gustiness_parameter, fgmin = (float)p.minimum_gustiness;
            const float fmeso2 = p.mesoscale_gustiness ? (float)(p.mesoscale_velocity * p.mesoscale_velocity) : 0.0f;
            const int max32 = maxit - 1 < NE_MIXED_MAXITER ? maxit - 1 : NE_MIXED_MAXITER;
            float fus = 1e-4f, fts = 1e-4f, fqs = 1e-4f;
            int k32 = 0;
            while (k32 < max32) {
                float bstar = fgoTv * (fts * fonedq + fdTv * fqs);
                float Usg2 = fmeso2;
                if (p.convective_gustiness) {
                    float Ug = fmaxf(fgmin, fbeta * fcbrt(fmaxf(0.0f, -fus * bstar) * fhbl));
                    Usg2 = Usg2 + Ug * Ug;
                }
                float U = tsqrt(fdu2dv2 + Usg2);
                float lu = fminf(fCg_g * fus * fus + tdiv(fCnu_nu, fus), flmax);
                float Rs = lu * fus * finv_nu;
                float lq = fminf(Rs == 0.0f ? 0.0f : fA * texp(-fb * tlog(Rs)), fsmax);
                float dh = fmaxf(fzd, 2.0f * lu);
                float L = bstar == 0.0f ? (float)INFINITY : tdiv(fus * fus, fkappa * bstar);
                float pmh, psh;
                psi_pair_default<float>(tdiv(dh, L), pmh, psh);
                float chiu = tdiv(fkappa, tlog(tdiv(dh, lu)) - pmh + psi_m_default<float>(tdiv(lu, L)));
                float chiq = tdiv(fkappa, tlog(tdiv(dh, lq)) - psh + psi_s_default<float>(tdiv(lq, L)));
                float nus = chiu * U, nts = chiq * fdtheta, nqs = chiq * fdq;
                float drift = fabsf(nus - fus) + fabsf(nts - fts) + fabsf(nqs - fqs);
                fus = nus; fts = nts; fqs = nqs;
                k32++;
                if (drift < NE_MIXED_DRIFT) {
                    done32 = true;
                    break;
                }
            }
            if (isfinite(fus) && isfinite(fts) && isfinite(fqs) && fus > 0.0f) {
                us = fus; ts = fts; qs = fqs;
                it = k32;
            } else {
                done32 = false;   // the FP64 iteration starts from the reference's first guess
            }
        }

        if constexpr (Defer && Mixed) {
            if (!done32) {
                ColumnFluxes d = {};
                d.deferred = true;
                return d;
            }
        }
        while (!done32) {
            // (iteration 0 always iterates; ConvergenceStopCriteria / FixedIterations tests below)
            double bstar = goTv * (ts * onedq + dTv * qs);
            double Usg2;
            if constexpr (LogForm) {
                // the default convective gustiness with ne_cbrt_ge1's first value and Halley step,
                // valid for any x in FP32's normal range: below 10⁻³⁰, β∛x is far under the
                // minimum gustiness the maximum keeps, so x is taken as 0 there
                const double x = jmax(0.0, -us * bstar) * hbl;
                const double Ug = jmax(p.minimum_gustiness, p.gustiness_parameter * (x < 1e-30 ? 0.0 : ne_cbrt_ge1(x)));
                Usg2 = Ug * Ug;
            } else {
                Usg2 = gustiness_squared(p, us, bstar, hbl);
            }
            double U = LogForm ? ne_sqrt(du2dv2 + Usg2) : sqrt(du2dv2 + Usg2);
            // LogForm: 1/u★ once (hardware reciprocal and two Newton steps), for ℓᵤ's smooth-wall
            // term and for 1/L; the Charnock term with Cg/g folded
            const double rus = LogForm ? quotient(1.0, us) : 0.0;
            double lu = LogForm ? jmin((p.charnock_parameter / p.roughness_gravitational_acceleration) * (us * us) +
                                           (p.smooth_wall_parameter * p.air_kinematic_viscosity) * rus,
                                       p.maximum_momentum_roughness)
                                : momentum_roughness(p, us);
            double lq, lt, lnlu = 0.0, lnlq = 0.0;
            if constexpr (LogForm) {
                lnlu = ne_log(lu);
                const double lnRs = lnlu + ne_log(us) - ln_nu;
                lq = jmin(p.reynolds_scaling_A * ne_exp(-p.reynolds_scaling_b * lnRs), p.maximum_scalar_roughness);
                lnlq = jmin(ln_A - p.reynolds_scaling_b * lnRs, ln_lqmax);
            } else if (Fast && p.water_vapor_roughness == NE_ROUGHNESS_EDSON) {
                double Rstar = lu * us / p.air_kinematic_viscosity;
                double ls = Rstar == 0.0 ? 0.0 : p.reynolds_scaling_A * exp(-p.reynolds_scaling_b * ne_log(Rstar));
                lq = jmin(ls, p.maximum_scalar_roughness);
            } else {
                lq = scalar_roughness(p, p.water_vapor_roughness, p.water_vapor_roughness_length, lu, us);
            }
            lt = shared_scalar ? lq
                               : scalar_roughness(p, p.temperature_roughness, p.temperature_roughness_length, lu, us);
            double dh = jmax(za - p.zero_plane_displacement, 2.0 * lu);
            double chiu, chiq, chit;
            if constexpr (LogForm) {
                const double invL = (kappa * bstar) * (rus * rus);
                const double lndh = 2.0 * lu > zd ? ln2 + lnlu : ln_zd;
                double pmh, psh;
                psi_pair_default(dh * invL, pmh, psh);
                chiu = quotient(kappa, (lndh - lnlu) - pmh + psi_m_default(lu * invL));
                chiq = quotient(kappa, (lndh - lnlq) - psh + psi_s_default(lq * invL));
                chit = chiq;
            } else if (Fast) {
                double L = bstar == 0.0 ? (double)INFINITY : us * us / (kappa * bstar);
                double pmh, psh;
                psi_pair_default(dh / L, pmh, psh);
                chiu = kappa / (ne_log(dh / lu) - pmh + psi_m_default(lu / L));
                chiq = kappa / (ne_log(dh / lq) - psh + psi_s_default(lq / L));
                chit = shared_scalar ? chiq : kappa / (ne_log(dh / lt) - psh + psi_s_default(lt / L));
            } else {
                double L = bstar == 0.0 ? (double)INFINITY : us * us / (kappa * bstar);
                chiu = kappa / similarity_profile(p, false, dh, lu, L);
                chiq = kappa / similarity_profile(p, true, dh, lq, L);
                chit = shared_scalar ? chiq : kappa / similarity_profile(p, true, dh, lt, L);
            }
            double nus = chiu * U, nts = chit * dtheta, nqs = chiq * dq;
            it += 1;
            bool go;
            if (converge) {
                double drift = fabs(nus - us) + fabs(nts - ts) + fabs(nqs - qs);
                bool conv = drift < tol;
                reached = it >= maxit;
                go = !(conv | reached);
            } else {
                go = it < maxit;
            }
            us = nus; ts = nts; qs = nqs;
            if (!go) break;
        }
        sT = Ts;
    }
    if (inactive) {  // zero_interface_state (interface_states.jl:800-803)
        if (!converge) it = p.maxiter > 0 ? p.maxiter : 0;  // FixedIterations iterates, then zeroes
        us = 0.0; ts = 0.0; qs = 0.0; su = 0.0; sv = 0.0; sT = NE_CELSIUS_TO_KELVIN;
    }

    double du, dv;
    if (p.velocity_difference == NE_VELOCITY_RELATIVE) { du = ua - su; dv = va - sv; }
    else { du = ua; dv = va; }
    double dU = sqrt(du * du + dv * dv);
    double taux = dU == 0.0 ? 0.0 : -(us * us) * du / dU;
    double tauy = dU == 0.0 ? 0.0 : -(us * us) * dv / dU;
    double rho = air_density(p, Ta, pa, qa);
    double cpm = cp_m(p, qa);
    double Lv = latent_heat_vapor(p, Ta);
    ColumnFluxes r;
    r.Qv = -rho * Lv * us * qs;
    r.Qt = -rho * cpm * us * ts;
    r.Jv = -rho * us * qs;
    r.rtx = rho * taux;
    r.rty = rho * tauy;
    r.us = us;
    r.ts = ts;
    r.qs = qs;
    r.Ts_C = sT - NE_CELSIUS_TO_KELVIN;
    r.iterations = it;
    r.converged = !(converge && reached);
    r.deferred = false;
    return r;
}

// Fractional index (0-based) of x in a source axis given by its cell-centre coordinates
// nodes[0 … n−1] (ascending): k + (x − x_k)/(x_{k+1} − x_k) in the node interval that holds x.
// Periodic axes (longitude, period 360°) wrap from the last node to the first one a period
// on; bounded axes extend their first and last interval linearly (the interpolation then
// holds the edge value, as it does for a regular grid). Flat axes and single nodes give 0.
__host__ __device__ inline double fractional_index_nodes(double x, const double* nodes, int n, int topology) {
    if (topology == NE_FLAT || n <= 1) return 0.0;
    if (topology == NE_PERIODIC) {
        const double x0 = nodes[0];
        if (x < x0 || x >= x0 + 360.0) {   // only longitudes outside the record's turn are moved
            double t = fmod(x - x0, 360.0);
            if (t < 0.0) t += 360.0;
            x = x0 + t;
        }
        if (x >= nodes[n - 1]) return (double)(n - 1) + (x - nodes[n - 1]) / (x0 + 360.0 - nodes[n - 1]);
    } else {
        if (x < nodes[0]) return (x - nodes[0]) / (nodes[1] - nodes[0]);
        if (x >= nodes[n - 1]) return (double)(n - 1) + (x - nodes[n - 1]) / (nodes[n - 1] - nodes[n - 2]);
    }
    int lo = 0, hi = n - 1;   // nodes[lo] ≤ x < nodes[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (nodes[mid] <= x) lo = mid;
        else hi = mid;
    }
    return (double)lo + (x - nodes[lo]) / (nodes[lo + 1] - nodes[lo]);
}

// ---------------------------------------------------------------------------
// Interpolation (interpolate_atmospheric_state.jl:91-182)
// ---------------------------------------------------------------------------

struct Neighbours {
    int lo, hi;
    double w;
};

__device__ __forceinline__ Neighbours neighbours(double f, int n, int topo) {
    Neighbours r;
    if (topo == NE_FLAT || n == 1) { r.lo = 0; r.hi = 0; r.w = 0.0; return r; }
    double fl = floor(f);
    int i0 = (int)fl;
    r.w = f - fl;
    int i1 = i0 + 1;
    if (topo == NE_PERIODIC) {
        i0 = wrap_near(i0, n);
        i1 = i0 == n - 1 ? 0 : i0 + 1;
    }
    else {
        i0 = i0 < 0 ? 0 : (i0 > n - 1 ? n - 1 : i0);
        i1 = i1 < 0 ? 0 : (i1 > n - 1 ? n - 1 : i1);
    }
    r.lo = i0; r.hi = i1;
    return r;
}

// Bilinear in (i, j), linear in time; Real is the storage type of the source field.
template <typename Real>
__device__ __forceinline__ double interp_field(const View3<const Real>& d, const Neighbours& x,
                                               const Neighbours& y, const ne_time_interp& t) {
    if (d.p == nullptr) return 0.0;
    double xi = x.w, eta = y.w;
    double w00 = (1.0 - xi) * (1.0 - eta);
    double w01 = (1.0 - xi) * eta;
    double w10 = xi * (1.0 - eta);
    double w11 = xi * eta;
    double a = w00 * (double)d(x.lo, y.lo, t.n1) + w01 * (double)d(x.lo, y.hi, t.n1) +
               w10 * (double)d(x.hi, y.lo, t.n1) + w11 * (double)d(x.hi, y.hi, t.n1);
    double b = w00 * (double)d(x.lo, y.lo, t.n2) + w01 * (double)d(x.lo, y.hi, t.n2) +
               w10 * (double)d(x.hi, y.lo, t.n2) + w11 * (double)d(x.hi, y.hi, t.n2);
    return (1.0 - t.tilde_n) * a + t.tilde_n * b;
}

// cpu_interpolating_time_indices (interpolate_atmospheric_state.jl:57-60): linear between
// the two snapshots around t inside the series, clamped to the first / last snapshot
// outside it. The same arithmetic on the host and on the device, so a device-side clock
// reproduces the host's indices exactly.
__host__ __device__ inline ne_time_interp time_indices(const double* times, int n, double t) {
    ne_time_interp r = {0, 0, 0.0};
    if (n <= 1 || t <= times[0]) return r;
    if (t >= times[n - 1]) {
        r.n1 = r.n2 = n - 1;
        return r;
    }
    int k = 0;
    while (k + 1 < n && times[k + 1] <= t) k++;
    r.n1 = k;
    r.n2 = k + 1;
    r.tilde_n = (t - times[k]) / (times[k + 1] - times[k]);
    return r;
}

// Oceananigans Cyclical(period) indexing: τ = mod(t − t¹, period) (Julia's mod: the result
// has the sign of the period), t' = t¹ + τ; inside the series as above, and between the last
// snapshot and the first one a period later, ñ = (t' − tᴺ)/(period − (tᴺ − t¹)) from the last
// snapshot to the first.
__host__ __device__ inline ne_time_interp time_indices(const double* times, int n, double t, int indexing,
                                                       double period) {
    if (indexing != NE_TIME_CYCLICAL || n <= 1) return time_indices(times, n, t);
    const double t1 = times[0], tN = times[n - 1];
    double tau = fmod(t - t1, period);
    if (tau != 0.0 && ((tau > 0.0) != (period > 0.0))) tau += period;
    const double tm = tau + t1;
    if (tm <= tN) return time_indices(times, n, tm);
    ne_time_interp r;
    r.n1 = n - 1;
    r.n2 = 0;
    r.tilde_n = (tm - tN) / (period - (tN - t1));
    return r;
}

// ---------------------------------------------------------------------------
// WENO reconstructions: Jiang & Shu (1996) candidate stencils and smoothness
// indicators with Borges et al. (2008) WENO-Z weights (exponent 2), ε = 1e-8, written
// in terms of the differences δ_n = c_{n+1} − c_n of the values along the upwind
// direction (the face lies after the third of five values, or the second of three):
//
//   β_r = 13/12 s_r² + 1/4 e_r², s = (δ₁−δ₀, δ₂−δ₁, δ₃−δ₂), e = (3δ₁−δ₀, δ₁+δ₂, 3δ₂−δ₃)
//   q_r − c₂ = (5δ₁−2δ₀, δ₁+2δ₂, 4δ₂−δ₃)/6
//
// The face value is c₂ + Σ_r ω_r (q_r − c₂): a uniform field gives c₂ exactly, and the
// reconstruction of the mirrored values is the negated reconstruction of the negated
// differences, so the upwind side is chosen by selecting differences. Scaling every β_r
// and ε by the same factor leaves the weights unchanged; the FP64 form uses 4β_r =
// (13/3)s_r² + e_r² with 4ε. The FP64 weights are normalised with one division:
// multiplying every α_r = d_r (1 + (τ/(β_r+ε))²) by Π_s (β_s+ε)² leaves α_r/Σα unchanged;
// the division is a reciprocal refined by two Newton steps (within 2 ulp). In FP32 the
// Π(β+ε)² products could underflow, so FP32 forms the three 1/(β_r+ε) from one reciprocal of
// Π(β+ε) (at least ε³) and the other two factors, and normalises with a second.
// Fused multiply-adds are allowed here (contract(fast)).
// ---------------------------------------------------------------------------


// Σ_r ω_r (q_r − c₂) of the 5-point WENO-Z reconstruction from the four differences.
__device__ __forceinline__ double weno5z_core(double d0, double d1, double d2, double d3) {
#pragma clang fp contract(fast)
    const double eps4 = 4.0 * NE_WENO_EPS;
    const double k = 2.0816659994661326;   // √(13/3)
    const double s0 = k * (d1 - d0), s1 = k * (d2 - d1), s2 = k * (d3 - d2);
    const double e0 = fma(3.0, d1, -d0), e1 = d1 + d2, e2 = fma(3.0, d2, -d3);
    const double b0 = fma(s0, s0, e0 * e0);
    const double b1 = fma(s1, s1, e1 * e1);
    const double b2 = fma(s2, s2, e2 * e2);
    const double tau = fabs(b0 - b2);
    const double t2 = tau * tau;
    const double p0 = (b0 + eps4) * (b0 + eps4), p1 = (b1 + eps4) * (b1 + eps4), p2 = (b2 + eps4) * (b2 + eps4);
    const double a0 = (0.1 * (p0 + t2)) * (p1 * p2);
    const double a1 = (0.6 * (p1 + t2)) * (p0 * p2);
    const double a2 = (0.3 * (p2 + t2)) * (p0 * p1);
    const double r0 = fma(5.0, d1, -2.0 * d0), r1 = fma(2.0, d2, d1), r2 = fma(4.0, d2, -d3);
    return quotient(fma(a0, r0, fma(a1, r1, a2 * r2)), 6.0 * (a0 + a1 + a2));
}

__device__ __forceinline__ float weno5z_core(float d0, float d1, float d2, float d3) {
#pragma clang fp contract(fast)
    const float eps = (float)NE_WENO_EPS;
    const float s0 = d1 - d0, s1 = d2 - d1, s2 = d3 - d2;
    const float e0 = fmaf(3.0f, d1, -d0), e1 = d1 + d2, e2 = fmaf(3.0f, d2, -d3);
    const float b0 = (13.0f / 12.0f) * s0 * s0 + 0.25f * e0 * e0 + eps;
    const float b1 = (13.0f / 12.0f) * s1 * s1 + 0.25f * e1 * e1 + eps;
    const float b2 = (13.0f / 12.0f) * s2 * s2 + 0.25f * e2 * e2 + eps;
    // 1/(βᵣ+ε) from one reciprocal of the product (≥ ε³, no underflow) and the other two factors
    const float b12 = b1 * b2;
    const float q = __builtin_amdgcn_rcpf(b0 * b12);
    const float tau = fabsf(b0 - b2);
    const float tq = tau * q;
    const float g0 = tq * b12;
    const float g1 = tq * (b0 * b2);
    const float g2 = tq * (b0 * b1);
    const float a0 = 0.1f * (1.0f + g0 * g0);
    const float a1 = 0.6f * (1.0f + g1 * g1);
    const float a2 = 0.3f * (1.0f + g2 * g2);
    const float r0 = fmaf(5.0f, d1, -2.0f * d0), r1 = fmaf(2.0f, d2, d1), r2 = fmaf(4.0f, d2, -d3);
    return (a0 * r0 + a1 * r1 + a2 * r2) * __builtin_amdgcn_rcpf(6.0f * (a0 + a1 + a2));
}

// Σ_r ω_r (q_r − c₁) of the 3-point WENO-Z reconstruction (values c₀, c₁, c₂; face after
// c₁): β_r = δ_r², q_r − c₁ = δ_r/2, linear weights 1/3, 2/3.
__device__ __forceinline__ double weno3z_core(double d0, double d1) {
#pragma clang fp contract(fast)
    const double eps = NE_WENO_EPS;
    const double b0 = d0 * d0, b1 = d1 * d1;
    const double tau = fabs(b0 - b1);
    const double t2 = tau * tau;
    const double p0 = (b0 + eps) * (b0 + eps), p1 = (b1 + eps) * (b1 + eps);
    const double a0 = (1.0 / 3.0) * (p0 + t2) * p1;
    const double a1 = (2.0 / 3.0) * (p1 + t2) * p0;
    return quotient(fma(a0, d0, a1 * d1), 2.0 * (a0 + a1));
}

__device__ __forceinline__ float weno3z_core(float d0, float d1) {
#pragma clang fp contract(fast)
    const float eps = (float)NE_WENO_EPS;
    const float b0 = d0 * d0, b1 = d1 * d1;
    const float tau = fabsf(b0 - b1);
    const float g0 = tau * __builtin_amdgcn_rcpf(b0 + eps);
    const float g1 = tau * __builtin_amdgcn_rcpf(b1 + eps);
    const float a0 = (1.0f / 3.0f) * (1.0f + g0 * g0);
    const float a1 = (2.0f / 3.0f) * (1.0f + g1 * g1);
    return (a0 * d0 + a1 * d1) * __builtin_amdgcn_rcpf(2.0f * (a0 + a1));
}

// Reduced order at a face between cells m-1 and m from the activity bits of cells
// m-3 … m+2 (bit n ↔ cell m-3+n): 0 peripheral, 1 upwind, 2 WENO3, 3 WENO5.
__device__ __forceinline__ int face_order(unsigned act) {
    if ((act & 0x0Cu) != 0x0Cu) return 0;
    if ((act & 0x1Eu) != 0x1Eu) return 1;
    if ((act & 0x3Fu) != 0x3Fu) return 2;
    return 3;
}

// Upwind-biased face value from the six cells m-3 … m+2 around the face (order 3: WENO5,
// 2: WENO3, else the upwind cell). With the flow towards +: c₂ + core(δ₀, δ₁, δ₂, δ₃);
// towards −, the mirrored values c₅ … c₁: c₃ − core(δ₄, δ₃, δ₂, δ₁) (δ_n = c_{n+1} − c_n).
// Only the cells the order uses are read, and one reconstruction is evaluated per lane.
template <typename R>
__device__ __forceinline__ R face_value(R c0, R c1, R c2, R c3, R c4, R c5, int order, R vel) {
    const bool left = vel >= (R)0;
    const R base = left ? c2 : c3;
    R core = (R)0;
    if (order >= 2) {
        const R D1 = c2 - c1, D2 = c3 - c2, D3 = c4 - c3;
        if (order == 3) {
            const R D0 = c1 - c0, D4 = c5 - c4;
            core = weno5z_core(left ? D0 : D4, left ? D1 : D3, D2, left ? D3 : D1);
        } else {
            core = weno3z_core(left ? D1 : D3, D2);
        }
    }
    return left ? base + core : base - core;
}

}  // namespace ne
