/*
 * ne_oracle.h — CPU restatement (FP64, OpenMP) of the reference's coupled surface
 * time-step path. TEST INFRASTRUCTURE: only the test suite, the smoke check and
 * the benchmark's CPU-baseline leg may load it; the product never does.
 *
 * The oracle takes the same descriptors as the C ABI (include/numericalearth_hip.h)
 * but every pointer inside them is HOST memory.
 */
#ifndef NE_ORACLE_H
#define NE_ORACLE_H

#include "numericalearth_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Thermodynamics (Thermodynamics.jl forms on AtmosphereThermodynamicsParameters). */
double neo_saturation_vapor_pressure(const ne_params* p, double T);
double neo_surface_specific_humidity(const ne_params* p, double pa, double Ts);
double neo_saturation_specific_humidity(const ne_params* p, double T, double pa);
double neo_cp_m(const ne_params* p, double q);
double neo_air_density(const ne_params* p, double T, double pa, double q);
double neo_latent_heat_vapor(const ne_params* p, double T);
double neo_virtual_temperature(const ne_params* p, double T, double q);

/* Similarity-theory pieces. */
double neo_psi_momentum(const ne_params* p, double zeta);
double neo_psi_scalar(const ne_params* p, double zeta);
double neo_gustiness_squared(const ne_params* p, double ustar, double bstar, double hbl);
double neo_momentum_roughness(const ne_params* p, double ustar);
double neo_scalar_roughness(const ne_params* p, double lu, double ustar, int which);

/* Coefficient-based fluxes: the polynomial neutral drag, the Large-Yeager stability
 * profiles (scalar = 0 momentum, 1 scalar) and transfer coefficients C = (Cd, Ch, Cq). */
double neo_neutral_drag(const ne_params* p, double U);
double neo_psi_large_yeager(const ne_params* p, int scalar, double zeta);
void neo_large_yeager_coefficients(const ne_params* p, double dU, double Ts, double qsurf, double dh,
                                   double us, double ts, double qs, double C[3]);
double neo_latent_heat_sublim(const ne_params* p, double T);
double neo_pow_int(double x, int n);

/* One column of the flux solve. in: ua va Ta pa qa To_C So uo vo SW LW (the downwelling
 * radiation, read by the skin-temperature balance when params->radiation is set);
 * inactive flag. out[9]: Qv Qt Jv rho_tau_x rho_tau_y ustar thetastar qstar Ts_C.
 * Returns iterations. */
int neo_flux_column(const ne_params* p, const double in[11], int inactive, double out[9]);

/* Phase restatements over a grid (host views). */
void neo_immersed_levels(const ne_grid* g, ne_view bottom_height, int32_t* kmin);
void neo_fractional_indices(const ne_grid* g, const ne_source_grid* s, double* fi, double* fj);
void neo_fractional_indices_nodes(const ne_grid* g, const double* lambda_nodes, int nx, int topology_x,
                                  const double* phi_nodes, int ny, int topology_y, double* fi, double* fj);
void neo_interp_field(const ne_grid* g, const ne_source_grid* s, const double* fi, const double* fj,
                      ne_view src, ne_time_interp t, ne_view out);
void neo_interp_atmos(const ne_grid* g, const ne_prescribed_atmosphere* a, const double* fi,
                      const double* fj, const ne_exchange_state* out);
void neo_interp_radiation(const ne_grid* g, const ne_prescribed_radiation* r, const double* fi,
                          const double* fj, ne_view sw, ne_view lw);
void neo_ao_flux_solve(const ne_grid* g, const ne_params* p, const ne_exchange_state* a,
                       const ne_ocean_state* o, const ne_interface_fluxes* out);
void neo_freeze_clamp(const ne_grid* g, const ne_params* p, ne_view T, ne_view S,
                      ne_view frazil, double dt);
void neo_assemble_net_ocean_fluxes(const ne_grid* g, const ne_params* p, const ne_interface_fluxes* f,
                                   const ne_exchange_state* a, ne_view sst, const ne_net_ocean_fluxes* net);
void neo_apply_air_sea_radiation(const ne_grid* g, const ne_params* p, ne_view sw, ne_view lw,
                                 ne_view Ts, const ne_net_ocean_fluxes* net, const ne_radiative_fluxes* rad);
/* params may be NULL (no tracer forcing); with params->radiative_forcing two-colour the T
 * tendency includes the TwoColorRadiation forcing of net->shortwave_surface_flux. */
void neo_tracer_tendency(const ne_grid* g, const ne_params* params, const ne_ocean_state* o,
                         const ne_net_ocean_fluxes* net, ne_view GT, ne_view GS);
/* As above, and writes Σ|terms|/V of every cell (the magnitude of the flux terms summed into G). */
void neo_tracer_tendency_scaled(const ne_grid* g, const ne_params* params, const ne_ocean_state* o,
                                const ne_net_ocean_fluxes* net, ne_view GT, ne_view GS, ne_view scale_T,
                                ne_view scale_S);
void neo_rk_stage(const ne_grid* g, const ne_params* params, ne_view u, ne_view v, ne_view w, ne_view T0, ne_view S0,
                  ne_view Tin, ne_view Sin, ne_view Tout, ne_view Sout,
                  const ne_net_ocean_fluxes* net, double beta_dt);
void neo_budget(const ne_grid* g, ne_view c, double weight, double out[2]);
void neo_compute_w_from_continuity(const ne_grid* g, ne_view u, ne_view v, ne_view w);

/* WENO reconstructions (left-biased: upwind value is the middle-left one). */
double neo_weno5z(double c0, double c1, double c2, double c3, double c4);
double neo_weno3z(double c0, double c1, double c2);

/* Whole coupled step on the CPU (OceanOnlyModel time_step!). Workspace views hold the
 * two intermediate RK stages. `iteration` is the ocean clock iteration before the step. */
typedef struct neo_model {
    ne_grid grid;
    ne_params params;
    ne_prescribed_atmosphere atmosphere;
    ne_prescribed_radiation radiation;
    const double *atmos_fi, *atmos_fj, *radiation_fi, *radiation_fj;
    ne_ocean_state ocean;
    ne_exchange_state exchange;           /* interpolated atmosphere (workspace) */
    ne_view sw, lw;                       /* interpolated radiation (workspace)  */
    ne_interface_fluxes fluxes;
    ne_net_ocean_fluxes net;
    ne_radiative_fluxes radiative;
    ne_view frazil_heat;
    ne_view T1, S1, T2, S2;               /* RK workspace */
} neo_model;

void neo_model_update_state(neo_model* m, double dt_frazil);
void neo_model_ocean_step(neo_model* m, double dt);

/* Number of OpenMP threads the oracle uses (0 = OpenMP default). */
void neo_set_threads(int n);
int neo_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
