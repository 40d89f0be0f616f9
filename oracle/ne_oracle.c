/*
 * ne_oracle.c — FP64 CPU restatement of NumericalEarth.jl's coupled surface
 * time-step path (TEST INFRASTRUCTURE: checker only; never linked by the product).
 *
 * Each function cites the reference file:line whose arithmetic it restates. Where a
 * Julia expression fixes an evaluation order, the C expression keeps that order
 * (Julia's binary operators are left-associative; unary minus binds tighter than *,
 * ^ tighter than unary minus), so the restatement can be compared bit-for-bit
 * where the operations involved are correctly rounded. Build with
 * -ffp-contract=off: no fused multiply-adds may be introduced.
 *
 * Third-party functions the reference calls but does not vendor are restated from
 * their published forms (Thermodynamics.jl, Oceananigans WENO / interpolation /
 * metrics / split RK3, ClimaSeaIce LinearLiquidus); their ulp-level arrangement is
 * "parity unpinned" and is pinned instead by closed forms, special cases and
 * invariants in the test suite.
 */
#include "ne_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CELSIUS_TO_KELVIN 273.15 /* src/EarthSystemModels/components.jl:10 */

static int g_threads = 0;
void neo_set_threads(int n) { g_threads = n; }
int neo_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#ifdef _OPENMP
#define NEO_THREADS num_threads(neo_get_threads())
#endif

/* Julia's min/max on floats: NaN-propagating, -0.0 < +0.0. */
static inline double jmin(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a < b) return a;
    if (b < a) return b;
    return signbit(a) ? a : b;
}
static inline double jmax(double a, double b) {
    if (a != a || b != b) return a + b;
    if (a > b) return a;
    if (b > a) return b;
    return signbit(a) ? b : a;
}

#define V3(v, i, j, k) (((double*)(v).data)[(int64_t)(i) * (v).sx + (int64_t)(j) * (v).sy + (int64_t)(k) * (v).sz])
#define V2(v, i, j) V3(v, i, j, 0)

/* ------------------------------------------------------------------------- */
/* Thermodynamics (Thermodynamics.jl on AtmosphereThermodynamicsParameters,   */
/* src/Atmospheres/thermodynamic_parameters.jl:61-258).                      */
/* ------------------------------------------------------------------------- */

static inline double Rv_over_Rd(const ne_params* p) { return p->dry_air_molar_mass / p->water_molar_mass; } /* :76 */
static inline double R_v(const ne_params* p) { return p->gas_constant / p->water_molar_mass; }               /* :77 */
static inline double R_d(const ne_params* p) { return p->gas_constant / p->dry_air_molar_mass; }             /* :78 */
static inline double cp_d(const ne_params* p) { return R_d(p) / p->dry_air_adiabatic_exponent; }            /* :256 */

/* gas_constant_air = R_d (1 + (Rv_over_Rd - 1) q_tot - Rv_over_Rd (q_liq + q_ice)), q_liq = q_ice = 0 */
static inline double gas_constant_air(const ne_params* p, double q) {
    double eps = Rv_over_Rd(p);
    return R_d(p) * ((1.0 + (eps - 1.0) * q) - eps * (0.0 + 0.0));
}

double neo_cp_m(const ne_params* p, double q) {
    /* cp_m = cp_d + (cp_v - cp_d) q_tot + (cp_l - cp_v) q_liq + (cp_i - cp_v) q_ice, q_liq = q_ice = 0 */
    double cpd = cp_d(p);
    return cpd + (p->water_vapor_heat_capacity - cpd) * q;
}

double neo_air_density(const ne_params* p, double T, double pa, double q) {
    return pa / (gas_constant_air(p, q) * T);
}

double neo_latent_heat_vapor(const ne_params* p, double T) {
    double dcp = p->water_vapor_heat_capacity - p->liquid_water_heat_capacity;
    return p->reference_vaporization_enthalpy + dcp * (T - p->reference_temperature);
}

/* latent_heat_sublim = ℒₛ₀ + (cₚᵥ − cₚᵢ)(T − T₀), the sublimation counterpart of the above;
 * the SkinTemperature balance takes it for the latent heat (interface_states.jl:544). */
double neo_latent_heat_sublim(const ne_params* p, double T) {
    double dcp = p->water_vapor_heat_capacity - p->water_ice_heat_capacity;
    return p->reference_sublimation_enthalpy + dcp * (T - p->reference_temperature);
}

/* x^n for a literal integer n > 3 as Julia evaluates Float64^Integer: power by squaring
 * carried in double-double (error-free products through fma), rounded once at the end,
 * so the result is the correctly rounded power up to a rare last-ulp tie. */
double neo_pow_int(double x, int n) {
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

double neo_virtual_temperature(const ne_params* p, double T, double q) {
    return gas_constant_air(p, q) / R_d(p) * T;
}

/* Clausius-Clapeyron with a linearised latent heat, liquid phase:
 * pₛₐₜ = pₜᵣ (T/Tₜᵣ)^(Δcₚ/Rᵥ) exp((ℒᵥ₀ - Δcₚ T₀)/Rᵥ (1/Tₜᵣ - 1/T)), Δcₚ = cₚᵥ - cₚₗ.
 * Called at interface_states.jl:63 and :83. */
double neo_saturation_vapor_pressure(const ne_params* p, double T) {
    double dcp = p->water_vapor_heat_capacity - p->liquid_water_heat_capacity;
    double Rv = R_v(p);
    double Ttr = p->triple_point_temperature;
    return p->triple_point_pressure * pow(T / Ttr, dcp / Rv) *
           exp((p->reference_vaporization_enthalpy - dcp * p->reference_temperature) / Rv *
               (1.0 / Ttr - 1.0 / T));
}

/* surface_specific_humidity(::ImpureSaturationSpecificHumidity) (interface_states.jl:55-74) */
double neo_surface_specific_humidity(const ne_params* p, double pa, double Ts) {
    if (p->specific_humidity == NE_HUMIDITY_FIXED) return p->fixed_specific_humidity;
    double pv = p->water_mole_fraction * neo_saturation_vapor_pressure(p, Ts);
    double edv = 1.0 / Rv_over_Rd(p);
    pv = jmin(pv, 0.999 * pa);
    return edv * pv / (pa - (1.0 - edv) * pv);
}

/* saturation_specific_humidity (interface_states.jl:79-90), used by the input generator. */
double neo_saturation_specific_humidity(const ne_params* p, double T, double pa) {
    double pv = neo_saturation_vapor_pressure(p, T);
    double edv = 1.0 / Rv_over_Rd(p);
    pv = jmin(pv, 0.999 * pa);
    return edv * pv / (pa - (1.0 - edv) * pv);
}

/* ------------------------------------------------------------------------- */
/* Similarity theory (similarity_theory_turbulent_fluxes.jl)                 */
/* ------------------------------------------------------------------------- */

/* EdsonMomentumStabilityFunction (:501-532) */
double neo_psi_momentum(const ne_params* p, double z) {
    if (p->stability_functions == NE_STABILITY_NONE) return 0.0;
    const double* c = p->edson_momentum;
    double zmax = c[0], Ap = c[1], Bp = c[2], Cp = c[3], Dp = c[4];
    double Am = c[5], Bm = c[6], Cm = c[7], Dm = c[8], Em = c[9], Fm = c[10];
    double zm = jmin(0.0, z);
    double zp = jmax(0.0, z);
    double dz = jmin(zmax, Ap * zp);
    double psi_p = -Bp * zp - Cp * (zp - Dp) * exp(-dz) - Cp * Dp;
    double f1 = sqrt(sqrt(1.0 - Am * zm));
    double psi_m1 = Bm * log((1.0 + f1) / Bm) + log((1.0 + f1 * f1) / Bm) - Bm * atan(f1) + Cm;
    double f2 = cbrt(1.0 - Dm * zm);
    double psi_m2 = Em / 2.0 * log((1.0 + f2 + f2 * f2) / Em) - sqrt(Em) * atan((1.0 + 2.0 * f2) / sqrt(Em)) + Fm;
    double f = zm * zm / (1.0 + zm * zm);
    double psi_m = (1.0 - f) * psi_m1 + f * psi_m2;
    return z < 0.0 ? psi_m : psi_p;
}

/* EdsonScalarStabilityFunction (:586-618) */
double neo_psi_scalar(const ne_params* p, double z) {
    if (p->stability_functions == NE_STABILITY_NONE) return 0.0;
    const double* c = p->edson_scalar;
    double zmax = c[0], Ap = c[1], Bp = c[2], Cp = c[3], Dp = c[4], Ep = c[5];
    double Am = c[6], Bm = c[7], Cm = c[8], Dm = c[9], Em = c[10], Fm = c[11];
    double zm = jmin(0.0, z);
    double zp = jmax(0.0, z);
    double dz = jmin(zmax, Ap * zp);
    double psi_p = -pow(1.0 + Bp * zp, Cp) - Bp * (zp - Dp) * exp(-dz) - Ep;
    double f1 = sqrt(1.0 - Am * zm);
    double psi_m1 = Bm * log((1.0 + f1) / Bm) + Cm;
    double f2 = cbrt(1.0 - Dm * zm);
    double psi_m2 = Em / 2.0 * log((1.0 + f2 + f2 * f2) / Em) - sqrt(Em) * atan((1.0 + 2.0 * f2) / sqrt(Em)) + Fm;
    double f = zm * zm / (1.0 + zm * zm);
    double psi_m = (1.0 - f) * psi_m1 + f * psi_m2;
    return z < 0.0 ? psi_m : psi_p;
}

/* vsgs²(::ConvectiveGustiness) (:91-95), composed as SubgridVelocityCorrection (:97-98) */
double neo_gustiness_squared(const ne_params* p, double ustar, double bstar, double hbl) {
    double v2 = 0.0;
    if (p->convective_gustiness) {
        double Jb = -ustar * bstar;
        double Ug = jmax(p->minimum_gustiness, p->gustiness_parameter * cbrt(jmax(0.0, Jb) * hbl));
        v2 = Ug * Ug;
    }
    if (p->mesoscale_gustiness) v2 = v2 + p->mesoscale_velocity * p->mesoscale_velocity;
    return v2;
}

/* roughness_length(::MomentumRoughnessLength) (roughness_lengths.jl:197-210) */
double neo_momentum_roughness(const ne_params* p, double ustar) {
    if (p->momentum_roughness == NE_ROUGHNESS_CONSTANT) return p->momentum_roughness_length;
    double nu = p->air_kinematic_viscosity;
    double g = p->roughness_gravitational_acceleration;
    double Cg = p->charnock_parameter;
    double Cnu = p->smooth_wall_parameter;
    double lW = Cg * (ustar * ustar) / g;
    double lR = Cnu == 0.0 ? 0.0 : Cnu * nu / ustar;
    return jmin(lW + lR, p->maximum_momentum_roughness);
}

/* roughness_length(::ScalarRoughnessLength) (roughness_lengths.jl:231-246); which: 0 T, 1 q */
double neo_scalar_roughness(const ne_params* p, double lu, double ustar, int which) {
    int kind = which == 0 ? p->temperature_roughness : p->water_vapor_roughness;
    if (kind == NE_ROUGHNESS_CONSTANT)
        return which == 0 ? p->temperature_roughness_length : p->water_vapor_roughness_length;
    double Rstar = lu * ustar / p->air_kinematic_viscosity;
    double ls = Rstar == 0.0 ? 0.0 : p->reynolds_scaling_A / pow(Rstar, p->reynolds_scaling_b);
    return jmin(ls, p->maximum_scalar_roughness);
}

/* similarity_profile (:242-253) */
static double similarity_profile(const ne_params* p, int scalar, double h, double l, double L) {
    double zeta = h / L;
    double psih = scalar ? neo_psi_scalar(p, zeta) : neo_psi_momentum(p, zeta);
    if (p->similarity_form == NE_SIMILARITY_COARE_LOG) return log(h / l) - psih;
    double psil = scalar ? neo_psi_scalar(p, l / L) : neo_psi_momentum(p, l / L);
    return log(h / l) - psih + psil;
}

/* ------------------------------------------------------------------------- */
/* Coefficient-based fluxes (coefficient_based_turbulent_fluxes.jl)           */
/* ------------------------------------------------------------------------- */

/* PolynomialNeutralDragCoefficient functor (:48-54) */
double neo_neutral_drag(const ne_params* p, double U) {
    const double* c = p->drag_polynomial;
    U = jmax(U, p->minimum_wind_speed);
    double poly = (c[0] / U + c[1] + c[2] * U - c[3] * neo_pow_int(U, 6)) / 1000.0;
    return U < p->high_wind_speed_threshold ? poly : p->high_wind_drag_coefficient;
}

/* stability_profile of the SplitStabilityFunction(LinearStable, Paulson) pairs of
 * large_yeager_stability_functions (similarity_theory_turbulent_fluxes.jl:772-778):
 * LinearStableStabilityFunction (:742-752) for ζ > 0, Paulson (:688-710) otherwise.
 * scalar = 0: momentum, 1: temperature / water vapour. */
double neo_psi_large_yeager(const ne_params* p, int scalar, double zeta) {
    const double* ls = scalar ? p->linear_stable_scalar : p->linear_stable_momentum;
    double zp = jmax(0.0, zeta);
    double stable = -ls[0] * jmin(zp, ls[1]);
    double a = scalar ? p->paulson_scalar : p->paulson_momentum[0];
    double zm = jmin(0.0, zeta);
    double z = sqrt(sqrt(1.0 - a * zm));
    double unstable;
    if (scalar) {
        unstable = 2.0 * log((1.0 + z * z) / 2.0);
    } else {
        double psi1 = 2.0 * log((1.0 + z) / 2.0);
        double psi2 = log((1.0 + z * z) / 2.0);
        double psi3 = -2.0 * atan(z);
        unstable = psi1 + psi2 + psi3 + p->paulson_momentum[1];
    }
    return zeta > 0.0 ? stable : unstable;
}

/* evaluate_coefficients(::LargeYeagerTransferCoefficients) (:288-340). us, ts, qs: the
 * previous iterate's scales. */
void neo_large_yeager_coefficients(const ne_params* p, double dU, double Ts, double qsurf, double dh,
                                   double us, double ts, double qs, double C[3]) {
    double kappa = p->large_yeager_von_karman_constant;
    double g = p->gravitational_acceleration;
    double h0 = p->large_yeager_reference_height;
    double Umin = p->minimum_wind_speed;
    dU = jmax(dU, Umin);
    /* buoyancy_scale (similarity:417-425) */
    double Tv = neo_virtual_temperature(p, Ts, qsurf);
    double delta = Rv_over_Rd(p) - 1.0;
    double bstar = g / Tv * (ts * (1.0 + delta * qsurf) + delta * Tv * qs);
    double L = bstar == 0.0 ? INFINITY : us * us / (kappa * bstar);
    double zeta = dh / L;
    double psim = neo_psi_large_yeager(p, 0, zeta);
    double psih = neo_psi_large_yeager(p, 1, zeta);
    double Cdp = us == 0.0 ? neo_neutral_drag(p, dU) : us * us / (dU * dU);
    double UN10 = dU / (1.0 + sqrt(Cdp) / kappa * (log(dh / h0) - psim));
    UN10 = jmax(UN10, Umin);
    double CdN = neo_neutral_drag(p, UN10);
    int stable = zeta > 0.0;
    double ChN = sqrt(CdN) / 1000.0 *
                 (stable ? p->stable_heat_transfer_coefficient : p->unstable_heat_transfer_coefficient);
    double CqN = sqrt(CdN) / 1000.0 * p->moisture_transfer_coefficient;
    double xim = sqrt(CdN) / kappa * (log(dh / h0) - psim);
    double Cd = CdN / ((1.0 + xim) * (1.0 + xim));
    double xih = sqrt(CdN) / kappa * (log(dh / h0) - psih);
    double ratio = sqrt(Cd) / sqrt(CdN);
    C[0] = Cd;
    C[1] = ChN * ratio / (1.0 + ChN * xih);
    C[2] = CqN * ratio / (1.0 + CqN * xih);
}

/* ------------------------------------------------------------------------- */
/* Interface state iteration (compute_interface_state.jl)                    */
/* ------------------------------------------------------------------------- */

typedef struct {
    double us, ts, qs;   /* u★ θ★ q★ */
    double u, v;         /* interface velocities */
    double T, q;         /* interface temperature (K), specific humidity */
} iface_state;

typedef struct { double z, u, v, T, p, q, hbl; } atmos_state;

/* air_sea_interface_radiation_state (Radiations/air_sea_interface_radiation_state.jl:8-33):
 * zero when there is no radiation. */
typedef struct { double sigma, alpha, eps, SW, LW; } rad_state;

/* surface_atmosphere_temperature (interface_states.jl:308-317) */
static double surface_atmosphere_temperature(const ne_params* p, const atmos_state* A) {
    return A->T + p->gravitational_acceleration * A->z / neo_cp_m(p, A->q);
}

/* compute_interface_temperature(::SkinTemperature) (interface_states.jl:536-577) with
 * flux_balance_temperature(::SkinTemperature{<:DiffusiveFlux}) (:425-448) for a constant
 * diffusivity. Ti: interior (top-cell) temperature in K. */
static double skin_temperature(const ne_params* p, const atmos_state* A, const rad_state* R, double Ti,
                               const iface_state* s) {
    double rhoa = neo_air_density(p, A->T, A->p, A->q);
    double ca = neo_cp_m(p, A->q);
    double Li = neo_latent_heat_sublim(p, A->T);
    double Tm = s->T;
    double Iup = R->sigma * R->eps * neo_pow_int(Tm, 4);
    double Qd = -(1.0 - R->alpha) * R->SW - R->eps * R->LW;
    double QT = -rhoa * ca * s->us * s->ts;
    double Qv = -rhoa * Li * s->us * s->qs;
    /* flux_balance_temperature */
    double kappa = p->skin_diffusivity;
    double delta = p->skin_depth;
    double rho = p->ocean_reference_density;
    double c = p->ocean_heat_capacity;
    double Qa = Qv + Iup + Qd;
    double lambda = 1.0 / (rho * c);
    double JT = Qa * lambda;
    double Ta = surface_atmosphere_temperature(p, A);
    double dT = Ta - Tm;
    double OmT = QT * lambda;
    double D = kappa * dT - OmT * delta;
    double Tstar = (Ti * kappa * dT - (JT * dT + OmT * Ta) * delta) / D;
    Tstar = D == 0.0 ? Tm : Tstar;
    double maxdT = p->skin_max_temperature_change;
    /* clamp(x, lo, hi) = x > hi ? hi : x < lo ? lo : x (Julia Base) */
    double x = Tstar - Ti;
    double cl = x > maxdT ? maxdT : (x < -maxdT ? -maxdT : x);
    return Ti + cl;
}

/* iterate_interface_fluxes(::SimilarityTheoryFluxes) (similarity:315-385) */
static void similarity_scales(const ne_params* p, const atmos_state* A, const iface_state* s, double Ts,
                              double qs, double dtheta, double dq, double dh, double out[3]) {
    double g = p->gravitational_acceleration;
    double ustar = s->us, tstar = s->ts, qstar = s->qs;
    /* buoyancy_scale (similarity:417-425) */
    double Tv = neo_virtual_temperature(p, Ts, qs);
    double delta = Rv_over_Rd(p) - 1.0;
    double bstar = g / Tv * (tstar * (1.0 + delta * qs) + delta * Tv * qstar);
    double Usg2 = neo_gustiness_squared(p, ustar, bstar, A->hbl);
    double du, dv;
    if (p->velocity_difference == NE_VELOCITY_RELATIVE) { du = A->u - s->u; dv = A->v - s->v; }
    else { du = A->u; dv = A->v; }
    double U = sqrt(du * du + dv * dv + Usg2);
    double lu = neo_momentum_roughness(p, ustar);
    double lq = neo_scalar_roughness(p, lu, ustar, 1);
    double lt = neo_scalar_roughness(p, lu, ustar, 0);
    dh = jmax(dh - p->zero_plane_displacement, 2.0 * lu); /* displaced_profile_height, :313 */
    double kappa = p->von_karman_constant;
    double L = bstar == 0.0 ? INFINITY : ustar * ustar / (kappa * bstar);
    double chiu = kappa / similarity_profile(p, 0, dh, lu, L);
    double chit = kappa / similarity_profile(p, 1, dh, lt, L);
    double chiq = kappa / similarity_profile(p, 1, dh, lq, L);
    out[0] = chiu * U;
    out[1] = chit * dtheta;
    out[2] = chiq * dq;
}

/* iterate_interface_fluxes(::CoefficientBasedFluxes) (coefficient_based_turbulent_fluxes.jl:346-371)
 * with evaluate_coefficients (:269-340). */
static void coefficient_scales(const ne_params* p, const atmos_state* A, const iface_state* s, double Ts,
                               double qs, double dtheta, double dq, double dh, double out[3]) {
    double du, dv;
    if (p->velocity_difference == NE_VELOCITY_RELATIVE) { du = A->u - s->u; dv = A->v - s->v; }
    else { du = A->u; dv = A->v; }
    int ly = p->transfer_coefficients == NE_TRANSFER_LARGE_YEAGER;
    double Umin = ly ? p->minimum_wind_speed : 0.0; /* minimum_wind_speed (:265-267) */
    double dU = jmax(sqrt(du * du + dv * dv), Umin);
    double C[3];
    if (ly) {
        neo_large_yeager_coefficients(p, dU, Ts, qs, dh, s->us, s->ts, s->qs, C);
    } else {
        for (int n = 0; n < 3; n++)
            C[n] = p->transfer_kind[n] == NE_COEFFICIENT_POLYNOMIAL ? neo_neutral_drag(p, dU)
                                                                    : p->transfer_coefficient[n];
    }
    double Cd = C[0];
    out[0] = sqrt(Cd) * dU;
    out[1] = Cd == 0.0 ? 0.0 : C[1] / sqrt(Cd) * dtheta;
    out[2] = Cd == 0.0 ? 0.0 : C[2] / sqrt(Cd) * dq;
}

/* iterate_interface_state (compute_interface_state.jl:69-122) with BulkTemperature
 * (interface_states.jl:333) or SkinTemperature, ImpureSaturationSpecificHumidity
 * (:587-588) or a fixed humidity, and the similarity-theory or coefficient-based fluxes. */
static iface_state iterate_state(const ne_params* p, const atmos_state* A, const rad_state* R, double Ti,
                                 iface_state s) {
    double Ts = p->interface_temperature == NE_TEMPERATURE_SKIN ? skin_temperature(p, A, R, Ti, &s) : s.T;
    double qs = neo_surface_specific_humidity(p, A->p, Ts);
    double dq = A->q - qs;
    double thetaa = surface_atmosphere_temperature(p, A);
    double dtheta = thetaa - Ts;
    double dh = A->z;
    double scales[3];
    if (p->turbulent_fluxes == NE_FLUXES_COEFFICIENT_BASED)
        coefficient_scales(p, A, &s, Ts, qs, dtheta, dq, dh, scales);
    else
        similarity_scales(p, A, &s, Ts, qs, dtheta, dq, dh, scales);
    iface_state n = s;
    n.us = scales[0];
    n.ts = scales[1];
    n.qs = scales[2];
    n.T = Ts;
    n.q = qs;
    return n;
}

/* One column of _compute_atmosphere_ocean_interface_state! (atmosphere_ocean_fluxes.jl:80-197). */
int neo_flux_column(const ne_params* p, const double in[11], int inactive, double out[9]) {
    atmos_state A;
    A.u = in[0]; A.v = in[1]; A.T = in[2]; A.p = in[3]; A.q = in[4];
    A.z = p->surface_layer_height;
    A.hbl = p->boundary_layer_height;
    double To = in[5] + CELSIUS_TO_KELVIN; /* convert_to_kelvin, components.jl:11 */
    double uo = in[7], vo = in[8];
    rad_state R = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (p->radiation) {
        R.sigma = p->stefan_boltzmann_constant;
        R.alpha = p->ocean_albedo;
        R.eps = p->ocean_emissivity;
        R.SW = in[9];
        R.LW = in[10];
    }

    double qs0 = neo_surface_specific_humidity(p, A.p, To);
    iface_state s0 = {1e-4, 1e-4, 1e-4, uo, vo, To, qs0};
    iface_state zero = {0.0, 0.0, 0.0, 0.0, 0.0, CELSIUS_TO_KELVIN, 0.0}; /* interface_states.jl:800-803 */
    iface_state s = s0;
    int it = 0;
    int converge = p->stop_criteria == NE_STOP_CONVERGENCE;
    if (converge && inactive) {
        s = zero;
    } else {
        /* compute_interface_state (compute_interface_state.jl:31-58) */
        iface_state prev = s0;
        for (;;) {
            int iterating;
            if (converge) {
                /* iterating(::ConvergenceStopCriteria) (:10-18) */
                double drift = fabs(s.us - prev.us) + fabs(s.ts - prev.ts) + fabs(s.qs - prev.qs);
                int converged = drift < p->tolerance;
                int reached = it >= p->maxiter;
                iterating = !(converged | reached) | (it == 0);
            } else {
                iterating = it < p->maxiter; /* FixedIterations (:24) */
            }
            if (!iterating) break;
            prev = s;
            s = iterate_state(p, &A, &R, To, prev);
            it += 1;
        }
    }
    if (inactive) s = zero; /* :158 */

    double du, dv;
    if (p->velocity_difference == NE_VELOCITY_RELATIVE) { du = A.u - s.u; dv = A.v - s.v; }
    else { du = A.u; dv = A.v; }
    double dU = sqrt(du * du + dv * dv);
    double taux = dU == 0.0 ? 0.0 : -(s.us * s.us) * du / dU;
    double tauy = dU == 0.0 ? 0.0 : -(s.us * s.us) * dv / dU;
    double rho = neo_air_density(p, A.T, A.p, A.q);
    double cpm = neo_cp_m(p, A.q);
    double Lv = neo_latent_heat_vapor(p, A.T);
    out[0] = -rho * Lv * s.us * s.qs;       /* latent heat 𝒬ᵛ   */
    out[1] = -rho * cpm * s.us * s.ts;      /* sensible heat 𝒬ᵀ */
    out[2] = -rho * s.us * s.qs;            /* water vapor Jᵛ   */
    out[3] = rho * taux;                    /* ρτˣ              */
    out[4] = rho * tauy;                    /* ρτʸ              */
    out[5] = s.us;
    out[6] = s.ts;
    out[7] = s.qs;
    out[8] = s.T - CELSIUS_TO_KELVIN;       /* convert_from_kelvin, components.jl:14 */
    return it;
}

/* ------------------------------------------------------------------------- */
/* Grid helpers                                                               */
/* ------------------------------------------------------------------------- */

static inline int wrap(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }

/* Column (i, j) inside the domain after periodic wrapping; -1 outside (bounded). */
static inline int64_t column(const ne_grid* g, int i, int j) {
    if (g->topology_x == NE_PERIODIC) i = wrap(i, g->Nx);
    else if (i < 0 || i >= g->Nx) return -1;
    if (g->topology_y == NE_PERIODIC) j = wrap(j, g->Ny);
    else if (j < 0 || j >= g->Ny) return -1;
    return (int64_t)i + (int64_t)j * g->Nx;
}

/* inactive_node complement for (Center, Center, Center): inside the domain and
 * above the immersed bottom. */
static inline int active(const ne_grid* g, int i, int j, int k) {
    if (k < 0 || k >= g->Nz) return 0;
    int64_t c = column(g, i, j);
    return c >= 0 && k >= g->kmin[c];
}

/* TwoColorRadiation in use (radiative_forcing.jl:84-91): shortwave penetrates below the surface. */
static int two_color(const ne_params* p, const ne_net_ocean_fluxes* n) {
    return p && p->radiative_forcing == NE_RADIATIVE_FORCING_TWO_COLOR && n->shortwave_surface_flux.data;
}

static inline int wx(const ne_grid* g, int i) { return g->topology_x == NE_PERIODIC ? wrap(i, g->Nx) : i; }
static inline int wy(const ne_grid* g, int j) { return g->topology_y == NE_PERIODIC ? wrap(j, g->Ny) : j; }

/* GridFittedBottom: immersed iff z_center <= bottom height. */
void neo_immersed_levels(const ne_grid* g, ne_view b, int32_t* kmin) {
    for (int j = 0; j < g->Ny; j++)
        for (int i = 0; i < g->Nx; i++) {
            double h = V2(b, i, j);
            int k = 0;
            while (k < g->Nz && g->z_center[k] <= h) k++;
            kmin[i + (int64_t)j * g->Nx] = k;
        }
}

/* ------------------------------------------------------------------------- */
/* Interpolation (prescribed_atmosphere_regridder.jl:51-71;                   */
/* interpolate_atmospheric_state.jl:91-182; interpolate_radiation_state.jl)   */
/* ------------------------------------------------------------------------- */

void neo_fractional_indices(const ne_grid* g, const ne_source_grid* s, double* fi, double* fj) {
    double lc0 = s->lambda0 + 0.5 * s->dlambda;
    double pc0 = s->phi0 + 0.5 * s->dphi;
    for (int i = 0; i < g->Nx; i++)
        fi[i] = s->topology_x == NE_FLAT ? 0.0 : (g->lambda_c[i] - lc0) / s->dlambda;
    for (int j = 0; j < g->Ny; j++)
        fj[j] = s->topology_y == NE_FLAT ? 0.0 : (g->phi_c[j] - pc0) / s->dphi;
}

/* Fractional index of x among cell-centre nodes x_0 < … < x_{n-1} (the stretched-direction
 * FractionalIndices of a LatitudeLongitudeGrid built from coordinate arrays, as for the JRA55
 * grid, DataWrangling/JRA55/JRA55_metadata.jl:27-40): k + (x − x_k)/(x_{k+1} − x_k) in the
 * interval holding x, found here by a linear scan. A periodic axis (360°) is the node list
 * continued by x_0 + 360, x being brought into [x_0, x_0 + 360) first; a bounded axis
 * continues its first and last interval. */
static double fractional_index_on_nodes(double x, const double* xn, int n, int topo) {
    if (topo == NE_FLAT || n <= 1) return 0.0;
    if (topo == NE_PERIODIC) {
        if (x < xn[0] || x >= xn[0] + 360.0) {
            double t = fmod(x - xn[0], 360.0);
            if (t < 0.0) t += 360.0;
            x = xn[0] + t;
        }
        for (int k = 0; k < n; k++) {
            double right = k + 1 < n ? xn[k + 1] : xn[0] + 360.0;
            if (k + 1 == n || x < right) return (double)k + (x - xn[k]) / (right - xn[k]);
        }
    }
    if (x < xn[0]) return (x - xn[0]) / (xn[1] - xn[0]);
    for (int k = 0; k + 1 < n; k++)
        if (x < xn[k + 1]) return (double)k + (x - xn[k]) / (xn[k + 1] - xn[k]);
    return (double)(n - 1) + (x - xn[n - 1]) / (xn[n - 1] - xn[n - 2]);
}

void neo_fractional_indices_nodes(const ne_grid* g, const double* lambda_nodes, int nx, int topology_x,
                                  const double* phi_nodes, int ny, int topology_y, double* fi, double* fj) {
    for (int i = 0; i < g->Nx; i++) fi[i] = fractional_index_on_nodes(g->lambda_c[i], lambda_nodes, nx, topology_x);
    for (int j = 0; j < g->Ny; j++) fj[j] = fractional_index_on_nodes(g->phi_c[j], phi_nodes, ny, topology_y);
}

/* Integer neighbours and weight of a fractional index: periodic wrap or a
 * zero-gradient (Neumann) halo at bounded edges. */
static inline void neighbours(double f, int n, int topo, int* lo, int* hi, double* xi) {
    if (topo == NE_FLAT || n == 1) { *lo = 0; *hi = 0; *xi = 0.0; return; }
    double fl = floor(f);
    int i0 = (int)fl;
    *xi = f - fl;
    int i1 = i0 + 1;
    if (topo == NE_PERIODIC) { i0 = wrap(i0, n); i1 = wrap(i1, n); }
    else {
        i0 = i0 < 0 ? 0 : (i0 > n - 1 ? n - 1 : i0);
        i1 = i1 < 0 ? 0 : (i1 > n - 1 ? n - 1 : i1);
    }
    *lo = i0; *hi = i1;
}

static inline double bilinear(ne_view d, int n, int i0, int i1, int j0, int j1, double xi, double eta) {
    double w00 = (1.0 - xi) * (1.0 - eta);
    double w01 = (1.0 - xi) * eta;
    double w10 = xi * (1.0 - eta);
    double w11 = xi * eta;
    return w00 * V3(d, i0, j0, n) + w01 * V3(d, i0, j1, n) + w10 * V3(d, i1, j0, n) + w11 * V3(d, i1, j1, n);
}

/* Width of the ring the interface kernels also cover: the reference launches them over
 * 0:N+1 in every non-Flat direction, except on single-column grids
 * (InterfaceComputations.jl:99-116); here when the caller's views have a halo. */
static inline int ring_width(const ne_grid* g) {
    return (g->halo >= 1 && !(g->Nx == 1 && g->Ny == 1)) ? 1 : 0;
}

/* Coordinate of exchange node i of n (centres x[0 … n−1]): interior nodes, and the halo nodes
 * −1 and n of the equally spaced grid one spacing beyond the edge nodes. */
static inline double halo_node(const double* x, int n, int i) {
    if (i >= 0 && i < n) return x[i];
    if (n == 1) return x[0];
    return i < 0 ? x[0] - (x[1] - x[0]) : x[n - 1] + (x[n - 1] - x[n - 2]);
}

/* Fractional index of exchange node i: the given interior index, or at a ring node the index of
 * its coordinate in a regular source grid (FractionalIndices of the halo node,
 * prescribed_atmosphere_regridder.jl:55-58); a source given by node coordinates (no regular
 * spacing) extends the edge indices linearly. */
static inline double node_index(const double* f, const double* x, int n, int i, int topology, double origin,
                                double spacing) {
    if (i >= 0 && i < n) return f[i];
    if (topology == NE_FLAT) return 0.0;
    if (spacing > 0.0) return (halo_node(x, n, i) - (origin + 0.5 * spacing)) / spacing;
    if (n == 1) return f[0];
    return i < 0 ? f[0] - (f[1] - f[0]) : f[n - 1] + (f[n - 1] - f[n - 2]);
}

void neo_interp_field(const ne_grid* g, const ne_source_grid* s, const double* fi, const double* fj,
                      ne_view src, ne_time_interp t, ne_view out) {
    const int R = ring_width(g);
#pragma omp parallel for NEO_THREADS
    for (int j = -R; j < g->Ny + R; j++) {
        int j0, j1; double eta;
        neighbours(node_index(fj, g->phi_c, g->Ny, j, s->topology_y, s->phi0, s->dphi), s->Ny, s->topology_y, &j0,
                   &j1, &eta);
        for (int i = -R; i < g->Nx + R; i++) {
            if (src.data == NULL) { V2(out, i, j) = 0.0; continue; }
            int i0, i1; double xi;
            neighbours(node_index(fi, g->lambda_c, g->Nx, i, s->topology_x, s->lambda0, s->dlambda), s->Nx,
                       s->topology_x, &i0, &i1, &xi);
            double a = bilinear(src, t.n1, i0, i1, j0, j1, xi, eta);
            double b = bilinear(src, t.n2, i0, i1, j0, j1, xi, eta);
            V2(out, i, j) = (1.0 - t.tilde_n) * a + t.tilde_n * b;
        }
    }
}

void neo_interp_atmos(const ne_grid* g, const ne_prescribed_atmosphere* a, const double* fi,
                      const double* fj, const ne_exchange_state* out) {
    neo_interp_field(g, &a->grid, fi, fj, a->u, a->time, out->u);
    neo_interp_field(g, &a->grid, fi, fj, a->v, a->time, out->v);
    neo_interp_field(g, &a->grid, fi, fj, a->T, a->time, out->T);
    neo_interp_field(g, &a->grid, fi, fj, a->p, a->time, out->p);
    neo_interp_field(g, &a->grid, fi, fj, a->q, a->time, out->q);
    if (out->rain.data) neo_interp_field(g, &a->grid, fi, fj, a->rain, a->time, out->rain);
    if (out->snow.data) neo_interp_field(g, &a->grid, fi, fj, a->snow, a->time, out->snow);
}

void neo_interp_radiation(const ne_grid* g, const ne_prescribed_radiation* r, const double* fi,
                          const double* fj, ne_view sw, ne_view lw) {
    neo_interp_field(g, &r->grid, fi, fj, r->downwelling_shortwave, r->time, sw);
    neo_interp_field(g, &r->grid, fi, fj, r->downwelling_longwave, r->time, lw);
}

/* ------------------------------------------------------------------------- */
/* Flux solve over the grid                                                  */
/* ------------------------------------------------------------------------- */

/* On the ring (ring_width) the ocean surface is read from the caller's halo, which holds
 * what Oceananigans' fill_halo_regions! leaves there, and a ring cell beyond a bounded edge
 * is inactive (inactive_node outside the domain). */
void neo_ao_flux_solve(const ne_grid* g, const ne_params* p, const ne_exchange_state* a,
                       const ne_ocean_state* o, const ne_interface_fluxes* f) {
    int kN = g->Nz - 1;
    const int R = ring_width(g);
#pragma omp parallel for schedule(dynamic, 4) NEO_THREADS
    for (int j = -R; j < g->Ny + R; j++) {
        for (int i = -R; i < g->Nx + R; i++) {
            double in[11], out[9];
            in[0] = V2(a->u, i, j); in[1] = V2(a->v, i, j); in[2] = V2(a->T, i, j);
            in[3] = V2(a->p, i, j); in[4] = V2(a->q, i, j);
            in[5] = V3(o->T, i, j, kN);
            in[6] = V3(o->S, i, j, kN);
            /* ℑxᶜᵃᵃ u, ℑyᵃᶜᵃ v at k = Nz (atmosphere_ocean_fluxes.jl:62-71) */
            int ip = g->topology_x == NE_PERIODIC ? wrap(i + 1, g->Nx) : i + 1;
            int jp = g->topology_y == NE_PERIODIC ? wrap(j + 1, g->Ny) : j + 1;
            in[7] = (V3(o->u, i, j, kN) + V3(o->u, ip, j, kN)) / 2.0;
            in[8] = (V3(o->v, i, j, kN) + V3(o->v, i, jp, kN)) / 2.0;
            in[9] = a->downwelling_shortwave.data ? V2(a->downwelling_shortwave, i, j) : 0.0;
            in[10] = a->downwelling_longwave.data ? V2(a->downwelling_longwave, i, j) : 0.0;
            int inactive = !active(g, i, j, kN);
            int it = neo_flux_column(p, in, inactive, out);
            V2(f->latent_heat, i, j) = out[0];
            V2(f->sensible_heat, i, j) = out[1];
            V2(f->water_vapor, i, j) = out[2];
            V2(f->x_momentum, i, j) = out[3];
            V2(f->y_momentum, i, j) = out[4];
            V2(f->friction_velocity, i, j) = out[5];
            V2(f->temperature_scale, i, j) = out[6];
            V2(f->water_vapor_scale, i, j) = out[7];
            V2(f->interface_temperature, i, j) = out[8];
            if (f->iterations.data)
                ((int32_t*)f->iterations.data)[i * f->iterations.sx + j * f->iterations.sy] = it;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Freeze clamp (freezing_limited_ocean_temperature.jl:95-118)               */
/* ------------------------------------------------------------------------- */

void neo_freeze_clamp(const ne_grid* g, const ne_params* p, ne_view T, ne_view S, ne_view frazil, double dt) {
    double rho = p->ocean_reference_density, c = p->ocean_heat_capacity;
#pragma omp parallel for NEO_THREADS
    for (int j = 0; j < g->Ny; j++)
        for (int i = 0; i < g->Nx; i++) {
            double dQ = 0.0;
            for (int k = g->Nz - 1; k >= 0; k--) {
                double dz = g->dz[k];
                double Tk = V3(T, i, j, k), Sk = V3(S, i, j, k);
                double Tm = p->freshwater_melting_temperature - p->liquidus_slope * Sk; /* LinearLiquidus */
                int freezing = Tk < Tm;
                double dE = freezing ? rho * c * (Tm - Tk) : 0.0;
                V3(T, i, j, k) = freezing ? Tm : Tk;
                dQ = dQ - dE * dz / dt;
            }
            V2(frazil, i, j) = dQ;
        }
}

/* ------------------------------------------------------------------------- */
/* Net ocean fluxes (assemble_net_ocean_fluxes.jl:74-153), ℵ = 0            */
/* ------------------------------------------------------------------------- */

void neo_assemble_net_ocean_fluxes(const ne_grid* g, const ne_params* p, const ne_interface_fluxes* f,
                                   const ne_exchange_state* a, ne_view sst, const ne_net_ocean_fluxes* n) {
    int kN = g->Nz - 1;
    double rinv = 1.0 / p->ocean_reference_density;
    double cinv = 1.0 / p->ocean_heat_capacity;
#pragma omp parallel for NEO_THREADS
    for (int j = 0; j < g->Ny; j++)
        for (int i = 0; i < g->Nx; i++) {
            double aice = 0.0;
            double To = V3(sst, i, j, kN);
            double Jr = a->rain.data ? V2(a->rain, i, j) : 0.0;
            double Js = a->snow.data ? V2(a->snow, i, j) : 0.0;
            double Jl = 0.0, Ps = 0.0;
            double Qt = V2(f->sensible_heat, i, j), Qv = V2(f->latent_heat, i, j), Jv = V2(f->water_vapor, i, j);
            double SQao = (Qt + Qv) * (1.0 - aice);
            double SFao = -(Jr + Jl + Js - Ps) * rinv + (1.0 - aice) * Jv * rinv;
            double Jwao = -SFao;
            int inactive = !active(g, i, j, kN);
            double Qin = 0.0, Jsio = 0.0, Jwio = 0.0;
            double JTao = SQao * rinv * cinv;
            double JTio = Qin * rinv * cinv;
            /* ℑxᶠᵃᵃ of τᶜᶜᶜ = ρₒ⁻¹ (1 - ℵ) ρτ; outside-domain neighbours are inactive (zero state) */
            int64_t cw = column(g, i - 1, j), cs = column(g, i, j - 1);
            double rtx_w = cw >= 0 ? V2(f->x_momentum, wx(g, i - 1), j) : 0.0;
            double rty_s = cs >= 0 ? V2(f->y_momentum, i, wy(g, j - 1)) : 0.0;
            double tx_ao = (rinv * (1.0 - aice) * V2(f->x_momentum, i, j) + rinv * (1.0 - aice) * rtx_w) / 2.0;
            double ty_ao = (rinv * (1.0 - aice) * V2(f->y_momentum, i, j) + rinv * (1.0 - aice) * rty_s) / 2.0;
            double tx_io = 0.0 * rinv * 0.0, ty_io = 0.0 * rinv * 0.0;
            V2(n->tau_x, i, j) = inactive ? 0.0 : tx_ao + tx_io;
            V2(n->tau_y, i, j) = inactive ? 0.0 : ty_ao + ty_io;
            V2(n->J_T, i, j) = inactive ? 0.0 : JTao + JTio;
            V2(n->J_S, i, j) = inactive ? 0.0 : Jsio;
            V2(n->J_w, i, j) = inactive ? 0.0 : Jwao + Jwio;
            V2(n->J_H, i, j) = inactive ? 0.0 : To * Jwao;
        }
}

/* ------------------------------------------------------------------------- */
/* Radiation (apply_air_sea_radiative_fluxes.jl:62-111; radiation_kernels.jl) */
/* ------------------------------------------------------------------------- */

void neo_apply_air_sea_radiation(const ne_grid* g, const ne_params* p, ne_view sw, ne_view lw,
                                 ne_view Tsv, const ne_net_ocean_fluxes* n, const ne_radiative_fluxes* r) {
    int kN = g->Nz - 1;
    double rinv = 1.0 / p->ocean_reference_density;
    double cinv = 1.0 / p->ocean_heat_capacity;
    double sigma = p->stefan_boltzmann_constant, alpha = p->ocean_albedo, eps = p->ocean_emissivity;
#pragma omp parallel for NEO_THREADS
    for (int j = 0; j < g->Ny; j++)
        for (int i = 0; i < g->Nx; i++) {
            double aice = 0.0;
            double Ts = V2(Tsv, i, j) + CELSIUS_TO_KELVIN;
            double SW = V2(sw, i, j), LW = V2(lw, i, j);
            /* emitted_longwave_radiation σϵT^4 (radiation_kernels.jl:3); T^4 as Julia evaluates
               Float64^Integer, the compensated power (as in the skin-temperature balance) */
            double Iup = sigma * eps * neo_pow_int(Ts, 4);
            double Ila = -eps * LW;
            double Itsw = -(1.0 - alpha) * SW;
            Ila = Ila * (1.0 - aice);
            Itsw = Itsw * (1.0 - aice);
            double Iup_ocean = Iup * (1.0 - aice);
            /* shortwave_radiative_forcing (radiative_forcing.jl:84-91): with TwoColorRadiation the
               transmitted shortwave becomes its surface flux J₀ and leaves Jᵀ */
            double Qss = Itsw;
            if (two_color(p, n)) {
                V2(n->shortwave_surface_flux, i, j) = -Itsw / (p->ocean_reference_density * p->ocean_heat_capacity);
                Qss = 0.0;
            }
            double SQ = Iup_ocean + Ila + Qss;
            double JTrad = SQ * rinv * cinv;
            int inactive = !active(g, i, j, kN);
            V2(n->J_T, i, j) = V2(n->J_T, i, j) + (inactive ? 0.0 : JTrad);
            if (r->upwelling_longwave.data) V2(r->upwelling_longwave, i, j) = Iup;
            if (r->downwelling_longwave.data) V2(r->downwelling_longwave, i, j) = -Ila;
            if (r->downwelling_shortwave.data) V2(r->downwelling_shortwave, i, j) = -Itsw;
        }
}

/* ------------------------------------------------------------------------- */
/* WENO reconstructions (Jiang & Shu 1996 smoothness indicators, Borges et al. */
/* 2008 WENO-Z weights, ε = 1e-8). Values ordered along the upwind direction;  */
/* the face lies between c2 and c3 (WENO5) or c1 and c2 (WENO3).               */
/* ------------------------------------------------------------------------- */

#define WENO_EPS 1e-8

double neo_weno5z(double c0, double c1, double c2, double c3, double c4) {
    double q0 = (2.0 * c0 - 7.0 * c1 + 11.0 * c2) / 6.0;
    double q1 = (-c1 + 5.0 * c2 + 2.0 * c3) / 6.0;
    double q2 = (2.0 * c2 + 5.0 * c3 - c4) / 6.0;
    double d0 = c0 - 2.0 * c1 + c2, e0 = c0 - 4.0 * c1 + 3.0 * c2;
    double d1 = c1 - 2.0 * c2 + c3, e1 = c1 - c3;
    double d2 = c2 - 2.0 * c3 + c4, e2 = 3.0 * c2 - 4.0 * c3 + c4;
    double b0 = 13.0 / 12.0 * (d0 * d0) + 0.25 * (e0 * e0);
    double b1 = 13.0 / 12.0 * (d1 * d1) + 0.25 * (e1 * e1);
    double b2 = 13.0 / 12.0 * (d2 * d2) + 0.25 * (e2 * e2);
    double tau = fabs(b0 - b2);
    double r0 = tau / (b0 + WENO_EPS), r1 = tau / (b1 + WENO_EPS), r2 = tau / (b2 + WENO_EPS);
    double a0 = 0.1 * (1.0 + r0 * r0);
    double a1 = 0.6 * (1.0 + r1 * r1);
    double a2 = 0.3 * (1.0 + r2 * r2);
    double s = a0 + a1 + a2;
    return (a0 * q0 + a1 * q1 + a2 * q2) / s;
}

double neo_weno3z(double c0, double c1, double c2) {
    double q0 = (-c0 + 3.0 * c1) / 2.0;
    double q1 = (c1 + c2) / 2.0;
    double d0 = c1 - c0, d1 = c2 - c1;
    double b0 = d0 * d0, b1 = d1 * d1;
    double tau = fabs(b0 - b1);
    double r0 = tau / (b0 + WENO_EPS), r1 = tau / (b1 + WENO_EPS);
    double a0 = (1.0 / 3.0) * (1.0 + r0 * r0);
    double a1 = (2.0 / 3.0) * (1.0 + r1 * r1);
    return (a0 * q0 + a1 * q1) / (a0 + a1);
}

/* Order (0: peripheral face, 1, 2, 3 ↦ WENO order 2N-1) at a face between cells
 * m-1 and m along a line, from the activity of the symmetric stencil m-3 … m+2. */
static inline int face_order(const int act[6]) {
    /* act[n] = activity of cell m-3+n */
    if (!act[2] || !act[3]) return 0;
    if (!act[1] || !act[4]) return 1;
    if (!act[0] || !act[5]) return 2;
    return 3;
}

/* Reconstructed face value from the six cells m-3 … m+2 (c[n]), upwind side chosen by vel. */
static inline double face_value(const double c[6], int order, double vel) {
    int left = vel >= 0.0;
    if (order == 3) return left ? neo_weno5z(c[0], c[1], c[2], c[3], c[4]) : neo_weno5z(c[5], c[4], c[3], c[2], c[1]);
    if (order == 2) return left ? neo_weno3z(c[1], c[2], c[3]) : neo_weno3z(c[4], c[3], c[2]);
    return left ? c[2] : c[3];
}

/* Advective fluxes Aₓuc̃, A_y vc̃, A_z wc̃ through the west/south/bottom face of cell (i,j,k). */
static double flux_x(const ne_grid* g, ne_view u, ne_view c, int i, int j, int k) {
    int act[6];
    for (int n = 0; n < 6; n++) act[n] = active(g, i - 3 + n, j, k);
    int order = face_order(act);
    if (order == 0) return 0.0;
    double cv[6];
    for (int n = 0; n < 6; n++) cv[n] = act[n] ? V3(c, wx(g, i - 3 + n), j, k) : 0.0;
    double vel = V3(u, wx(g, i), j, k);
    double A = g->dy_xface[j] * g->dz[k];
    return A * vel * face_value(cv, order, vel);
}

static double flux_y(const ne_grid* g, ne_view v, ne_view c, int i, int j, int k) {
    int act[6];
    for (int n = 0; n < 6; n++) act[n] = active(g, i, j - 3 + n, k);
    int order = face_order(act);
    if (order == 0) return 0.0;
    double cv[6];
    for (int n = 0; n < 6; n++) cv[n] = act[n] ? V3(c, i, wy(g, j - 3 + n), k) : 0.0;
    int jf = wy(g, j);
    double vel = V3(v, i, jf, k);
    double A = g->dx_yface[jf] * g->dz[k];
    return A * vel * face_value(cv, order, vel);
}

static double flux_z(const ne_grid* g, ne_view w, ne_view c, int i, int j, int k) {
    int act[6];
    for (int n = 0; n < 6; n++) act[n] = active(g, i, j, k - 3 + n);
    int order = face_order(act);
    if (order == 0) return 0.0;
    double cv[6];
    for (int n = 0; n < 6; n++) cv[n] = act[n] ? V3(c, i, j, k - 3 + n) : 0.0;
    double vel = V3(w, i, j, k);
    double A = g->Az[j];
    return A * vel * face_value(cv, order, vel);
}

/* Gc = −(1/V)(δx(Aₓuc̃) + δy(A_y vc̃) + δz(A_z wc̃)) − top boundary flux·Az/V at k = Nz.
 * If `scale` is given it receives Σ|terms|/V, the magnitude of the terms summed into G
 * (the scale against which the rounding of G is measured). */
/* TwoColorRadiation (Oceans/radiative_forcing.jl:52-71). Flux of one colour through face kf
   (0-based; face kf is the bottom of cell kf, face Nz the surface η): J₀ exp(κ (z − η)), and
   zero where the cell below the face is inactive, so the radiation that reaches the bottom
   is absorbed by the lowest active cell. */
static double beers_law_radiation(const ne_grid* g, int i, int j, int kf, double J0, double kappa) {
    double z = g->z_face[kf];
    double eta = g->z_face[g->Nz];
    double J = J0 * exp(kappa * (z - eta));
    return active(g, i, j, kf - 1) ? J : 0.0;
}

/* ε₁ ∂z J₁ + (1 − ε₁) ∂z J₂ at cell k (∂zᶜᶜᶜ: (J(k+1) − J(k)) / Δz(k)); *mag receives the
   magnitude of the terms. */
static double two_color_forcing(const ne_grid* g, const ne_params* p, int i, int j, int k, double J0,
                                double* mag) {
    double e1 = p->first_color_fraction, k1 = p->first_absorption_coefficient, k2 = p->second_absorption_coefficient;
    double dz = g->dz[k];
    double a1 = beers_law_radiation(g, i, j, k + 1, J0, k1), b1 = beers_law_radiation(g, i, j, k, J0, k1);
    double a2 = beers_law_radiation(g, i, j, k + 1, J0, k2), b2 = beers_law_radiation(g, i, j, k, J0, k2);
    double dJ1dz = (a1 - b1) / dz;
    double dJ2dz = (a2 - b2) / dz;
    if (mag) *mag = (e1 * (fabs(a1) + fabs(b1)) + (1.0 - e1) * (fabs(a2) + fabs(b2))) / dz;
    return e1 * dJ1dz + (1.0 - e1) * dJ2dz;
}

static double tendency(const ne_grid* g, ne_view u, ne_view v, ne_view w, ne_view c,
                       int i, int j, int k, double top_flux, double forcing, double forcing_mag, double* scale) {
    double div = 0.0, mag = 0.0;
    if (g->advection) {
        double fxe = flux_x(g, u, c, i + 1, j, k), fxw = flux_x(g, u, c, i, j, k);
        double fyn = flux_y(g, v, c, i, j + 1, k), fys = flux_y(g, v, c, i, j, k);
        double fzt = flux_z(g, w, c, i, j, k + 1), fzb = flux_z(g, w, c, i, j, k);
        double dx = fxe - fxw;
        double dy = fyn - fys;
        double dzf = fzt - fzb;
        div = dx + dy + dzf;
        mag = fabs(fxe) + fabs(fxw) + fabs(fyn) + fabs(fys) + fabs(fzt) + fabs(fzb);
    }
    double V = g->Az[j] * g->dz[k];
    double G = -(1.0 / V * div);
    G = G + forcing;
    mag = mag + forcing_mag * V;
    if (k == g->Nz - 1) {
        G = G - top_flux * g->Az[j] / V;
        mag = mag + fabs(top_flux * g->Az[j]);
    }
    if (scale) *scale = mag / V;
    return G;
}

void neo_tracer_tendency(const ne_grid* g, const ne_params* p, const ne_ocean_state* o,
                         const ne_net_ocean_fluxes* n, ne_view GT, ne_view GS) {
    ne_view none = {NULL, 0, 0, 0};
    neo_tracer_tendency_scaled(g, p, o, n, GT, GS, none, none);
}

void neo_tracer_tendency_scaled(const ne_grid* g, const ne_params* p, const ne_ocean_state* o,
                                const ne_net_ocean_fluxes* n, ne_view GT, ne_view GS, ne_view sT, ne_view sS) {
    int tc = two_color(p, n);
    int kN = g->Nz - 1;
#pragma omp parallel for NEO_THREADS
    for (int j = 0; j < g->Ny; j++)
        for (int i = 0; i < g->Nx; i++)
            for (int k = 0; k < g->Nz; k++) {
                double scT = 0.0, scS = 0.0;
                if (!active(g, i, j, k)) {
                    V3(GT, i, j, k) = 0.0;
                    V3(GS, i, j, k) = 0.0;
                } else {
                    double FT = 0.0, FS = 0.0;
                    if (k == kN) {
                        /* MultipleFluxes(Jᵀ, FreshwaterExchange{:T}) on a static grid = Jᵀ + 0;
                           MultipleFluxes(Jˢ, FreshwaterExchange{:S}) = Jˢ + (Sᴺ Jʷ − 0)
                           (multiple_surface_fluxes.jl:27-31; ocean_simulation.jl:80-91) */
                        FT = V2(n->J_T, i, j) + (0.0 + 0.0);
                        FS = V2(n->J_S, i, j) + ((V3(o->S, i, j, kN) * V2(n->J_w, i, j) - 0.0) + 0.0);
                    }
                    double F = 0.0, Fm = 0.0;
                    if (tc) F = two_color_forcing(g, p, i, j, k, V2(n->shortwave_surface_flux, i, j), &Fm);
                    V3(GT, i, j, k) = tendency(g, o->u, o->v, o->w, o->T, i, j, k, FT, F, Fm, &scT);
                    V3(GS, i, j, k) = tendency(g, o->u, o->v, o->w, o->S, i, j, k, FS, 0.0, 0.0, &scS);
                }
                if (sT.data) V3(sT, i, j, k) = scT;
                if (sS.data) V3(sS, i, j, k) = scS;
            }
}

void neo_rk_stage(const ne_grid* g, const ne_params* p, ne_view u, ne_view v, ne_view w, ne_view T0, ne_view S0,
                  ne_view Tin, ne_view Sin, ne_view Tout, ne_view Sout,
                  const ne_net_ocean_fluxes* n, double beta_dt) {
    int kN = g->Nz - 1;
    int tc = two_color(p, n);
    /* Tout may alias T0: each cell reads T0 only at itself, and writes after all
       stencil reads of Tin, which never aliases Tout. Rows are independent. */
#pragma omp parallel for NEO_THREADS
    for (int j = 0; j < g->Ny; j++)
        for (int i = 0; i < g->Nx; i++)
            for (int k = 0; k < g->Nz; k++) {
                if (!active(g, i, j, k)) continue;
                double FT = 0.0, FS = 0.0;
                if (k == kN) {
                    FT = V2(n->J_T, i, j) + (0.0 + 0.0);
                    FS = V2(n->J_S, i, j) + ((V3(Sin, i, j, kN) * V2(n->J_w, i, j) - 0.0) + 0.0);
                }
                double F = tc ? two_color_forcing(g, p, i, j, k, V2(n->shortwave_surface_flux, i, j), NULL) : 0.0;
                double GTv = tendency(g, u, v, w, Tin, i, j, k, FT, F, 0.0, NULL);
                double GSv = tendency(g, u, v, w, Sin, i, j, k, FS, 0.0, 0.0, NULL);
                V3(Tout, i, j, k) = V3(T0, i, j, k) + beta_dt * GTv;
                V3(Sout, i, j, k) = V3(S0, i, j, k) + beta_dt * GSv;
            }
}

/* Neumaier-compensated Σ weight·V·c over active cells, fixed (k, j, i) order. */
void neo_budget(const ne_grid* g, ne_view c, double weight, double out[2]) {
    double sum = 0.0, comp = 0.0;
    for (int k = 0; k < g->Nz; k++)
        for (int j = 0; j < g->Ny; j++)
            for (int i = 0; i < g->Nx; i++) {
                if (!active(g, i, j, k)) continue;
                double x = weight * (g->Az[j] * g->dz[k]) * V3(c, i, j, k);
                double t = sum + x;
                if (fabs(sum) >= fabs(x)) comp += (sum - t) + x;
                else comp += (x - t) + sum;
                sum = t;
            }
    out[0] = sum;
    out[1] = comp;
}

/* ------------------------------------------------------------------------- */
/* Whole coupled step (time_step_earth_system_model.jl:14-83)                */
/* ------------------------------------------------------------------------- */

void neo_model_update_state(neo_model* m, double dt_frazil) {
    const ne_grid* g = &m->grid;
    if (m->params.radiation)
        neo_interp_radiation(g, &m->radiation, m->radiation_fi, m->radiation_fj, m->sw, m->lw);
    neo_interp_atmos(g, &m->atmosphere, m->atmos_fi, m->atmos_fj, &m->exchange);
    /* the radiation exchanger's state, read by the SkinTemperature balance */
    ne_view none = {NULL, 0, 0, 0};
    m->exchange.downwelling_shortwave = m->params.radiation ? m->sw : none;
    m->exchange.downwelling_longwave = m->params.radiation ? m->lw : none;
    neo_ao_flux_solve(g, &m->params, &m->exchange, &m->ocean, &m->fluxes);
    neo_freeze_clamp(g, &m->params, m->ocean.T, m->ocean.S, m->frazil_heat, dt_frazil);
    neo_assemble_net_ocean_fluxes(g, &m->params, &m->fluxes, &m->exchange, m->ocean.T, &m->net);
    if (m->params.radiation)
        neo_apply_air_sea_radiation(g, &m->params, m->sw, m->lw, m->fluxes.interface_temperature,
                                    &m->net, &m->radiative);
}

/* compute_w_from_continuity! of the HydrostaticFreeSurfaceModel (Oceananigans, not vendored;
 * ocean_simulation.jl:487-499 builds the model): w = 0 at the bottom face of every column,
 * w[k+1] = w[k] − (δx(Ax u) + δy(Ay v))/Az on active cells, Ax = Δy·Δz and Ay = Δx·Δz of the
 * cell's faces; faces above an inactive cell carry 0. */
void neo_compute_w_from_continuity(const ne_grid* g, ne_view u, ne_view v, ne_view w) {
#pragma omp parallel for schedule(static) NEO_THREADS
    for (int j = 0; j < g->Ny; j++) {
        for (int i = 0; i < g->Nx; i++) {
            int ip = g->topology_x == NE_PERIODIC ? wrap(i + 1, g->Nx) : i + 1;
            int jp = g->topology_y == NE_PERIODIC ? wrap(j + 1, g->Ny) : j + 1;
            double wk = 0.0;
            V3(w, i, j, 0) = 0.0;
            for (int k = 0; k < g->Nz; k++) {
                double Ax = g->dy_xface[j] * g->dz[k];
                double Ays = g->dx_yface[j] * g->dz[k];
                double Ayn = g->dx_yface[jp] * g->dz[k];
                double div = (Ax * V3(u, ip, j, k) - Ax * V3(u, i, j, k)) +
                             (Ayn * V3(v, i, jp, k) - Ays * V3(v, i, j, k));
                wk = active(g, i, j, k) ? wk - div / g->Az[j] : 0.0;
                V3(w, i, j, k + 1) = wk;
            }
        }
    }
}

/* SplitRungeKutta3: c⁽ᵐ⁾ = c⁰ + βₘ Δt G(c⁽ᵐ⁻¹⁾), β = (1/3, 1/2, 1). */
void neo_model_ocean_step(neo_model* m, double dt) {
    const ne_grid* g = &m->grid;
    ne_ocean_state* o = &m->ocean;
    const double beta[3] = {1.0 / 3.0, 0.5, 1.0};
    neo_rk_stage(g, &m->params, o->u, o->v, o->w, o->T, o->S, o->T, o->S, m->T1, m->S1, &m->net, beta[0] * dt);
    neo_rk_stage(g, &m->params, o->u, o->v, o->w, o->T, o->S, m->T1, m->S1, m->T2, m->S2, &m->net, beta[1] * dt);
    neo_rk_stage(g, &m->params, o->u, o->v, o->w, o->T, o->S, m->T2, m->S2, o->T, o->S, &m->net, beta[2] * dt);
}
