/*
 * numericalearth_hip.h — C ABI of the MI355X (gfx950) coupled surface time-step path.
 *
 * The path replaces four phases of NumericalEarth.jl's coupled step
 * (src/EarthSystemModels/time_step_earth_system_model.jl:14-83) for an
 * OceanOnlyModel on an immersed LatitudeLongitudeGrid:
 *
 *   phase 1  prescribed atmosphere / radiation interpolation
 *   phase 2  SimilarityTheoryFluxes fixed-point air-sea flux solve (+ freeze clamp)
 *   phase 3  net ocean flux assembly, phase 4 radiative contribution
 *   ocean    WENO5 tracer-advection tendency with the T/S top flux boundary
 *            condition, advanced by a three-stage split Runge-Kutta step
 *
 * Conventions (mirroring the reference's kernels, which are allocation-free and
 * write in place into caller-owned Fields):
 *   - every pointer handed to an entry point is a DEVICE pointer unless the
 *     parameter is documented as host memory; the library never frees or
 *     reallocates caller memory;
 *   - arrays are described by ne_view: `data` points at interior element
 *     (i, j, k) = (1, 1, 1) in Julia's 1-based numbering and sx/sy/sz are element
 *     strides, so a Julia `parent(field)` with any halo maps directly
 *     (data = pointer(parent) + offset of the first interior point);
 *   - 3-D tracer and velocity views must have sx == 1 (i fastest);
 *   - every entry point returns an ne_status (0 = success) and never throws;
 *   - work is enqueued on `stream` (a hipStream_t passed as void*; NULL is the
 *     null stream) and entry points do not synchronise unless documented.
 */
#ifndef NUMERICALEARTH_HIP_H
#define NUMERICALEARTH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NE_ABI_VERSION 1

typedef enum ne_status {
    NE_OK = 0,
    NE_ERR_NULL = 1,         /* a required pointer is NULL                          */
    NE_ERR_SHAPE = 2,        /* sizes, strides or topology are inconsistent          */
    NE_ERR_UNSUPPORTED = 3,  /* the option is not implemented by this library        */
    NE_ERR_DEVICE = 4,       /* a HIP runtime call failed                            */
    NE_ERR_CAPACITY = 5,     /* a documented maximum is exceeded                     */
    NE_ERR_STATE = 6         /* object used in the wrong state                       */
} ne_status;

/* Topology of a horizontal direction (Oceananigans Periodic / Bounded / Flat). */
enum { NE_PERIODIC = 0, NE_BOUNDED = 1, NE_FLAT = 2 };

/* Arithmetic of a model / batch member. */
enum { NE_FP64 = 0, NE_MIXED_FP32 = 1 };

/* ---------------------------------------------------------------------------
 * Views and descriptors
 * ------------------------------------------------------------------------- */

typedef struct ne_view {
    void*   data;            /* interior origin (1,1,1); element type set by the entry point */
    int64_t sx, sy, sz;      /* element strides (sz doubles as the snapshot stride of a time series) */
} ne_view;

/* Immersed LatitudeLongitudeGrid (exchange grid == ocean grid), device arrays.
 * Reference: Oceananigans LatitudeLongitudeGrid + ImmersedBoundaryGrid(GridFittedBottom)
 * as built by the reference's examples (README.md:78-108). Metrics are per
 * latitude row because LatLon metrics depend on j only. Nz ≤ NE_MAX_LEVELS (grids with
 * more levels are refused with NE_ERR_UNSUPPORTED). */
#define NE_MAX_LEVELS 1023
typedef struct ne_grid {
    int32_t Nx, Ny, Nz;
    int32_t topology_x;          /* NE_PERIODIC or NE_BOUNDED                       */
    int32_t topology_y;          /* NE_BOUNDED (NE_PERIODIC accepted)               */
    int32_t advection;           /* 0 disables tracer advection (single column grids,
                                    src/Oceans/ocean_simulation.jl:367-379)         */
    const double* dz;            /* [Nz]   Δzᶜᶜᶜ per level (m)                       */
    const double* z_center;      /* [Nz]   cell-centre height (m, negative down)     */
    const double* dy_xface;      /* [Ny]   Δyᶠᶜᵃ, meridional length of x-faces (m)   */
    const double* dx_yface;      /* [Ny+1] Δxᶜᶠᵃ, zonal length of y-faces (m)        */
    const double* Az;            /* [Ny]   Azᶜᶜᵃ, horizontal cell area (m²)          */
    const double* lambda_c;      /* [Nx]   cell-centre longitude (deg)               */
    const double* phi_c;         /* [Ny]   cell-centre latitude (deg)                */
    const int32_t* kmin;         /* [Nx*Ny] i-fastest: first active level (0-based)
                                    of each column; Nz marks a land column          */
    const double* z_face;        /* [Nz+1] cell-face height (m); face k is the bottom of
                                    cell k, face Nz the surface                      */
    int32_t halo;                /* halo width of the caller's 2-D surface views (0: the
                                    views cover the interior only). With halo ≥ 1 the
                                    interpolations and the flux solve also write the ring
                                    i = −1, Nx and j = −1, Ny, as the reference launches
                                    them over 0:N+1 in every non-Flat direction
                                    (InterfaceComputations.jl:99-116); single-column grids
                                    have no ring                                      */
} ne_grid;

/* Regular longitude-latitude source grid of a prescribed atmosphere or radiation
 * (Oceananigans LatitudeLongitudeGrid, topology (Periodic|Bounded, Bounded, Flat)).
 * Cell i (0-based) has centre longitude lambda0 + (i + 1/2) dlambda. */
typedef struct ne_source_grid {
    int32_t Nx, Ny;
    int32_t topology_x, topology_y;
    double  lambda0, dlambda;
    double  phi0, dphi;
} ne_source_grid;

/* Time interpolation between snapshots n1 and n2 with weight tilde_n
 * (Oceananigans cpu_interpolating_time_indices, computed on the host;
 * src/Atmospheres/interpolate_atmospheric_state.jl:57-60). 0-based snapshots. */
typedef struct ne_time_interp {
    int32_t n1, n2;
    double  tilde_n;
} ne_time_interp;

/* PrescribedAtmosphere FieldTimeSeries data (src/Atmospheres/prescribed_atmosphere.jl:219-255).
 * Each view addresses data[i*sx + j*sy + n*sz] on the source grid. rain/snow may have
 * data == NULL, which reads as zero flux. */
typedef struct ne_prescribed_atmosphere {
    ne_source_grid grid;
    ne_view u, v, T, q, p, rain, snow;
    ne_time_interp time;
} ne_prescribed_atmosphere;

/* PrescribedRadiation downwelling shortwave / longwave (src/Radiations/prescribed_radiation.jl:55-92). */
typedef struct ne_prescribed_radiation {
    ne_source_grid grid;
    ne_view downwelling_shortwave, downwelling_longwave;
    ne_time_interp time;
} ne_prescribed_radiation;

/* Atmosphere state interpolated onto the exchange grid (2-D (C,C,Nothing) fields,
 * src/Atmospheres/prescribed_atmosphere_regridder.jl:7-13), and the radiation
 * exchanger's interpolated downwelling shortwave / longwave
 * (atmosphere_ocean_fluxes.jl:12-15), which only the SkinTemperature flux balance reads
 * (interface_states.jl:547-553); with no radiation, or with BulkTemperature, they may
 * have data == NULL. */
typedef struct ne_exchange_state {
    ne_view u, v, T, p, q, rain, snow;
    ne_view downwelling_shortwave, downwelling_longwave;
} ne_exchange_state;

/* Ocean state: u (Face,Center,Center), v (Center,Face,Center), w (Center,Center,Face),
 * T, S (Center,Center,Center). Face index i addresses the west face of cell i. */
typedef struct ne_ocean_state {
    ne_view u, v, w, T, S;
} ne_ocean_state;

/* AtmosphereSurfaceFluxes + interface temperature
 * (src/EarthSystemModels/InterfaceComputations/atmosphere_ocean_fluxes.jl:176-196).
 * `iterations` (int32) may have data == NULL; when set it receives the number of
 * fixed-point iterations each column took. */
typedef struct ne_interface_fluxes {
    ne_view latent_heat, sensible_heat, water_vapor, x_momentum, y_momentum;
    ne_view friction_velocity, temperature_scale, water_vapor_scale;
    ne_view interface_temperature;
    ne_view iterations;
} ne_interface_fluxes;

/* Net ocean fluxes (src/Oceans/assemble_net_ocean_fluxes.jl:119-151): τˣ (Face,Center),
 * τʸ (Center,Face), Jᵀ, Jˢ, Jʷ (= net_fluxes.η), Jᴴ (freshwater heat content). */
typedef struct ne_net_ocean_fluxes {
    ne_view tau_x, tau_y, J_T, J_S, J_w, J_H;
    ne_view shortwave_surface_flux;  /* J₀ = −ℐₜˢʷ/(ρₒcₒ) (K m/s) of TwoColorRadiation
                                        (radiative_forcing.jl:86-91); read and written only when
                                        params.radiative_forcing is two-color, else may be NULL */
} ne_net_ocean_fluxes;

/* InterfaceRadiationFlux diagnostics (src/Radiations/apply_air_sea_radiative_fluxes.jl:107-109). */
typedef struct ne_radiative_fluxes {
    ne_view upwelling_longwave, downwelling_longwave, downwelling_shortwave;
} ne_radiative_fluxes;

/* ---------------------------------------------------------------------------
 * Parameters of the flux formulation, interfaces and components (host struct,
 * passed by pointer and copied by value into kernels). ne_params_default fills
 * the reference defaults.
 * ------------------------------------------------------------------------- */

enum { NE_STABILITY_NONE = 0, NE_STABILITY_EDSON = 1 };
enum { NE_SIMILARITY_LOG = 0, NE_SIMILARITY_COARE_LOG = 1 };
enum { NE_ROUGHNESS_CONSTANT = 0, NE_ROUGHNESS_EDSON = 1 };
enum { NE_STOP_CONVERGENCE = 0, NE_STOP_FIXED_ITERATIONS = 1 };
enum { NE_HUMIDITY_IMPURE_SATURATION = 0, NE_HUMIDITY_FIXED = 1 };
enum { NE_VELOCITY_RELATIVE = 0, NE_VELOCITY_WIND = 1 };

typedef struct ne_params {
    /* SimilarityTheoryFluxes (similarity_theory_turbulent_fluxes.jl:174-214) */
    double  von_karman_constant;          /* 0.4 */
    double  turbulent_prandtl_number;     /* 1; stored, unused by the iteration (as in the reference) */
    int32_t convective_gustiness;         /* 1: ConvectiveGustiness (:45-48, :91-95); 0: nothing */
    int32_t mesoscale_gustiness;          /* 1: SubgridVelocityCorrection mesoscale Number (:97-98) */
    double  gustiness_parameter;          /* β = 1.2 */
    double  minimum_gustiness;            /* 0.01 m/s */
    double  mesoscale_velocity;           /* m/s */
    int32_t stability_functions;          /* NE_STABILITY_EDSON (:487-625) or NE_STABILITY_NONE (ψ ≡ 0) */
    int32_t similarity_form;              /* NE_SIMILARITY_LOG (:242-247) or NE_SIMILARITY_COARE_LOG (:249-253) */
    double  edson_momentum[11];           /* ζmax A⁺ B⁺ C⁺ D⁺ A⁻ B⁻ C⁻ D⁻ E⁻ F⁻ (:487-499) */
    double  edson_scalar[12];             /* ζmax A⁺ B⁺ C⁺ D⁺ E⁺ A⁻ B⁻ C⁻ D⁻ E⁻ F⁻ (:571-584) */
    /* roughness lengths (roughness_lengths.jl) */
    int32_t momentum_roughness;           /* NE_ROUGHNESS_EDSON: MomentumRoughnessLength (:197-210) */
    int32_t temperature_roughness;        /* NE_ROUGHNESS_EDSON: ScalarRoughnessLength (:234-246) */
    int32_t water_vapor_roughness;
    int32_t reserved0;
    double  momentum_roughness_length;    /* value used when the kind is NE_ROUGHNESS_CONSTANT */
    double  temperature_roughness_length;
    double  water_vapor_roughness_length;
    double  roughness_gravitational_acceleration; /* MomentumRoughnessLength g (default 9.80665) */
    double  air_kinematic_viscosity;      /* ν = 1.5e-5 */
    double  charnock_parameter;           /* wave_formulation = 0.02 */
    double  smooth_wall_parameter;        /* 0.11 */
    double  maximum_momentum_roughness;   /* 1 */
    double  reynolds_scaling_A;           /* 5.85e-5 */
    double  reynolds_scaling_b;           /* 0.72 */
    double  maximum_scalar_roughness;     /* 1.6e-4 */
    double  zero_plane_displacement;      /* 0 m */
    /* solver stop criteria (compute_interface_state.jl:5-24) */
    int32_t stop_criteria;                /* NE_STOP_CONVERGENCE or NE_STOP_FIXED_ITERATIONS */
    int32_t maxiter;                      /* maxiter (100) or the fixed iteration count */
    double  tolerance;                    /* 1e-8 */
    /* interface properties (component_interfaces.jl:353-358, :399-401) */
    int32_t specific_humidity;            /* NE_HUMIDITY_IMPURE_SATURATION (x_H₂O) or NE_HUMIDITY_FIXED */
    int32_t velocity_difference;          /* NE_VELOCITY_RELATIVE or NE_VELOCITY_WIND */
    double  water_mole_fraction;          /* 0.98 */
    double  fixed_specific_humidity;      /* qₛ when NE_HUMIDITY_FIXED */
    /* atmosphere / interface properties (prescribed_atmosphere.jl:222-223; component_interfaces.jl:483-484) */
    double  surface_layer_height;         /* 10 m */
    double  boundary_layer_height;        /* 512 m */
    double  gravitational_acceleration;   /* 9.80665 */
    /* AtmosphereThermodynamicsParameters (src/Atmospheres/thermodynamic_parameters.jl) */
    double  gas_constant;                 /* 8.3144598 */
    double  dry_air_molar_mass;           /* 0.02897 */
    double  water_molar_mass;             /* 0.018015 */
    double  dry_air_adiabatic_exponent;   /* 2/7 */
    double  water_vapor_heat_capacity;    /* 1859 */
    double  liquid_water_heat_capacity;   /* 4181 */
    double  reference_vaporization_enthalpy; /* 2500800 */
    double  reference_temperature;        /* 273.16 */
    double  triple_point_temperature;     /* 273.16 */
    double  triple_point_pressure;        /* 611.657 */
    /* ocean properties (component_interfaces.jl:433-436; ocean_simulation.jl:341, :527-530) */
    double  ocean_reference_density;      /* 1020 */
    double  ocean_heat_capacity;          /* TEOS-10 cₚ⁰ = 3991.86795711963 */
    double  freshwater_density;           /* 1000 */
    /* radiation (src/Radiations/Radiations.jl:13-16, prescribed_radiation.jl:68) */
    int32_t radiation;                    /* 1: PrescribedRadiation present */
    int32_t radiative_forcing;            /* NE_RADIATIVE_FORCING_NONE: transmitted shortwave goes
                                             into Jᵀ; NE_RADIATIVE_FORCING_TWO_COLOR: it penetrates
                                             (TwoColorRadiation, Oceans/radiative_forcing.jl:5-106) */
    double  stefan_boltzmann_constant;    /* 5.670374419e-8 */
    double  ocean_albedo;                 /* 0.05 */
    double  ocean_emissivity;             /* 0.97 */
    /* FreezingLimitedOceanTemperature LinearLiquidus (freezing_limited_ocean_temperature.jl:25) */
    double  freshwater_melting_temperature; /* 0 °C */
    double  liquidus_slope;               /* 0.054 °C/psu */
    /* TwoColorRadiation (radiative_forcing.jl:36-46; defaults of default_radiative_forcing,
     * ocean_simulation.jl:158-167: Paulson & Simpson Type I) */
    double  first_color_fraction;         /* 0.58 */
    double  first_absorption_coefficient; /* 1/0.35 m⁻¹ */
    double  second_absorption_coefficient;/* 1/23 m⁻¹ */
    /* Turbulent flux formulation and interface temperature. CoefficientBasedFluxes
     * (InterfaceComputations/coefficient_based_turbulent_fluxes.jl:142-371) shares the stop
     * criteria fields above (its defaults: tolerance 1e-8, maxiter 20, :215-229). */
    int32_t turbulent_fluxes;             /* NE_FLUXES_SIMILARITY_THEORY or NE_FLUXES_COEFFICIENT_BASED */
    int32_t transfer_coefficients;        /* NE_TRANSFER_SCALES: one coefficient each for momentum,
                                             temperature, water vapour (:273-285);
                                             NE_TRANSFER_LARGE_YEAGER (:288-340) */
    int32_t transfer_kind[3];             /* NE_TRANSFER_SCALES entries: NE_COEFFICIENT_CONSTANT or
                                             NE_COEFFICIENT_POLYNOMIAL (the polynomial below, :269-271) */
    int32_t interface_temperature;        /* NE_TEMPERATURE_BULK (interface_states.jl:333) or
                                             NE_TEMPERATURE_SKIN (SkinTemperature(DiffusiveFlux(κ, δ)), :355-577) */
    double  transfer_coefficient[3];      /* constant Cd, Cₕ, Cq (default 1e-3 each) */
    /* PolynomialNeutralDragCoefficient (:20-54): C_dN = (a/U + b + cU − dU⁶)/1000 */
    double  drag_polynomial[4];           /* a b c d: 2.7, 0.142, 1/13.09, 3.14807e-10 */
    double  high_wind_speed_threshold;    /* 33 m/s */
    double  high_wind_drag_coefficient;   /* 2.34e-3 */
    double  minimum_wind_speed;           /* 0.5 m/s */
    /* LargeYeagerTransferCoefficients (:82-108) */
    double  large_yeager_von_karman_constant;   /* 0.4 */
    double  large_yeager_reference_height;      /* h₀ = 10 m */
    double  stable_heat_transfer_coefficient;   /* 18 */
    double  unstable_heat_transfer_coefficient; /* 32.7 */
    double  moisture_transfer_coefficient;      /* 34.6 */
    /* large_yeager_stability_functions (similarity_theory_turbulent_fluxes.jl:772-778):
     * SplitStabilityFunction(stable = LinearStable, unstable = Paulson) */
    double  paulson_momentum[2];          /* a b: 16, π/2 (:683-699) */
    double  paulson_scalar;               /* a: 16 (:701-710) */
    double  linear_stable_momentum[2];    /* c ζmax: 5, 10 (:742-752) */
    double  linear_stable_scalar[2];      /* c ζmax: 5, 10 */
    /* SkinTemperature(DiffusiveFlux(κ, δ); max_ΔT) (interface_states.jl:355-444, :536-577) */
    double  skin_diffusivity;             /* κ, m² s⁻¹ */
    double  skin_depth;                   /* δ, m */
    double  skin_max_temperature_change;  /* max_ΔT = 5 K */
    double  reference_sublimation_enthalpy; /* ℒₛ₀ = 2834400 (thermodynamic_parameters.jl:150) */
    double  water_ice_heat_capacity;      /* cₚᵢ = 2100 (thermodynamic_parameters.jl:102) */
} ne_params;

enum { NE_RADIATIVE_FORCING_NONE = 0, NE_RADIATIVE_FORCING_TWO_COLOR = 1 };
enum { NE_FLUXES_SIMILARITY_THEORY = 0, NE_FLUXES_COEFFICIENT_BASED = 1 };
enum { NE_TRANSFER_SCALES = 0, NE_TRANSFER_LARGE_YEAGER = 1 };
enum { NE_COEFFICIENT_CONSTANT = 0, NE_COEFFICIENT_POLYNOMIAL = 1 };
enum { NE_TEMPERATURE_BULK = 0, NE_TEMPERATURE_SKIN = 1 };

/* Counters written by the flux-solve kernels (reference: non-convergence is silent
 * and NaNs propagate to the NaNChecker, earth_system_model.jl:483-495). */
typedef struct ne_counters {
    uint64_t total_iterations;            /* Σ iterations over solved columns       */
    uint32_t solved_columns;              /* active columns solved                   */
    uint32_t nonconverged_columns;        /* columns that hit maxiter                */
    uint32_t nan_columns;                 /* columns whose u★, θ★ or q★ is not finite */
    uint32_t max_iterations;              /* largest per-column iteration count      */
} ne_counters;

typedef struct ne_ctx ne_ctx;

/* ---------------------------------------------------------------------------
 * Library and context
 * ------------------------------------------------------------------------- */

/* ABI version (NE_ABI_VERSION). Host only. */
int ne_version(void);

/* Fill `p` with the reference defaults: SimilarityTheoryFluxes(Float64)
 * (similarity_theory_turbulent_fluxes.jl:174-214), ComponentInterfaces defaults
 * (component_interfaces.jl:390-419), PrescribedAtmosphere heights
 * (prescribed_atmosphere.jl:222-223), AtmosphereThermodynamicsParameters,
 * TEOS-10 ocean properties, PrescribedRadiation ocean surface, LinearLiquidus. Host only. */
int ne_params_default(ne_params* p);

/* Create a context on HIP device `device`. The context owns the status counters
 * and no other device memory. Replaces the reference's implicit architecture state. */
int ne_ctx_create(int device, ne_ctx** ctx);
int ne_ctx_destroy(ne_ctx* ctx);

/* Copy the counters to host memory `out` (synchronises `stream`) and, if `reset`
 * is non-zero, zero them. */
int ne_ctx_counters(ne_ctx* ctx, ne_counters* out, int reset, void* stream);

/* ---------------------------------------------------------------------------
 * Per-phase entry points (FP64). Each replaces one reference kernel launch.
 * ------------------------------------------------------------------------- */

/* First active level of every column of an ImmersedBoundaryGrid(GridFittedBottom):
 * cell (i,j,k) is immersed iff z_center[k] <= bottom_height[i,j]. Writes kmin
 * (the grid's own array may be passed). bottom_height is FP64 2-D.
 * Replaces the immersed-cell predicate behind Oceananigans inactive_node
 * (used at atmosphere_ocean_fluxes.jl:142, assemble_net_ocean_fluxes.jl:127). */
int ne_grid_immersed_levels(ne_ctx* ctx, const ne_grid* grid, ne_view bottom_height,
                            int32_t* kmin, void* stream);

/* Fractional indices of the exchange-grid centres in a source grid
 * (_compute_fractional_indices!, prescribed_atmosphere_regridder.jl:51-71;
 * radiation twin prescribed_radiation_regridder.jl:36-51). For LatLon grids the
 * fields are separable: fi[i] (Nx) and fj[j] (Ny), 0-based fractional positions. */
int ne_fractional_indices_f64(ne_ctx* ctx, const ne_grid* grid, const ne_source_grid* source,
                              double* fi, double* fj, void* stream);

/* Fractional indices in a source grid given by its cell-centre coordinates instead of a
 * regular spacing: lambda_nodes[nx] (longitudes, degrees) and phi_nodes[ny] (latitudes), DEVICE
 * arrays in ascending order. The reference builds the JRA55 forcing grid from the files'
 * own cell bounds (DataWrangling/JRA55/JRA55_metadata.jl:38-48), so its latitude centres (the
 * midpoints of bounds around 320 Gaussian latitudes) are unequally spaced; a fractional
 * position in such a stretched direction (FractionalIndices on a stretched
 * LatitudeLongitudeGrid) is k + (x − x_k)/(x_{k+1} − x_k) in the node interval holding x.
 * Periodic axes (period 360°) wrap from the last node to the first one a period on; bounded
 * axes extend their first and last interval (the interpolation then holds the edge value).
 * fi, fj as for ne_fractional_indices_f64; the source grid passed to the interpolation keeps
 * its Nx, Ny and topologies (its spacing fields are then unused). */
int ne_fractional_indices_nodes_f64(ne_ctx* ctx, const ne_grid* grid, const double* lambda_nodes, int32_t nx,
                                    int32_t topology_x, const double* phi_nodes, int32_t ny, int32_t topology_y,
                                    double* fi, double* fj, void* stream);

/* _interpolate_primary_atmospheric_state! (interpolate_atmospheric_state.jl:91-137):
 * bilinear in space, linear in time, onto the exchange grid interior. */
int ne_interp_atmos_f64(ne_ctx* ctx, const ne_grid* grid, const ne_prescribed_atmosphere* atmos,
                        const double* fi, const double* fj,
                        const ne_exchange_state* out, void* stream);

/* _interpolate_radiation_state! (interpolate_radiation_state.jl:43-69). */
int ne_interp_radiation_f64(ne_ctx* ctx, const ne_grid* grid, const ne_prescribed_radiation* rad,
                            const double* fi, const double* fj,
                            ne_view downwelling_shortwave, ne_view downwelling_longwave,
                            void* stream);

/* _compute_atmosphere_ocean_interface_state! (atmosphere_ocean_fluxes.jl:80-197):
 * per-column Monin-Obukhov fixed point (compute_interface_state.jl:31-58). Reads
 * the exchange state and the ocean surface (k = Nz) and writes the interface fluxes. */
int ne_ao_flux_solve_f64(ne_ctx* ctx, const ne_grid* grid, const ne_params* params,
                         const ne_exchange_state* atmos, const ne_ocean_state* ocean,
                         const ne_interface_fluxes* out, void* stream);

/* _freeze_ocean_temperature! (freezing_limited_ocean_temperature.jl:95-118): clamps T
 * to the liquidus in place and writes the frazil heat flux. dt = +INFINITY on the
 * first (iteration 0) call, as at :88. */
int ne_freeze_clamp_f64(ne_ctx* ctx, const ne_grid* grid, const ne_params* params,
                        ne_view T, ne_view S, ne_view frazil_heat, double dt, void* stream);

/* _assemble_net_ocean_fluxes! (assemble_net_ocean_fluxes.jl:74-153) for an ocean-only
 * model with FreezingLimitedOceanTemperature (ℵ = 0, no sea-ice fluxes). `sst` is the
 * ocean temperature field (read at k = Nz), `rain`/`snow` come from the exchange state. */
int ne_assemble_net_ocean_fluxes_f64(ne_ctx* ctx, const ne_grid* grid, const ne_params* params,
                                     const ne_interface_fluxes* fluxes, const ne_exchange_state* atmos,
                                     ne_view sst, const ne_net_ocean_fluxes* out, void* stream);

/* _apply_air_sea_radiative_fluxes! (apply_air_sea_radiative_fluxes.jl:62-111): Jᵀ += radiative
 * contribution, and the three diagnostics. */
int ne_apply_air_sea_radiation_f64(ne_ctx* ctx, const ne_grid* grid, const ne_params* params,
                                   ne_view downwelling_shortwave, ne_view downwelling_longwave,
                                   ne_view interface_temperature,
                                   const ne_net_ocean_fluxes* net, const ne_radiative_fluxes* out,
                                   void* stream);

/* Tracer tendency of T and S: −∇·(U c) with WENO(order=5) on the immersed grid plus the
 * top flux boundary condition (Gᵀ[Nz] −= Jᵀ Az/V, Gˢ[Nz] −= (Jˢ + S Jʷ) Az/V;
 * ocean_simulation.jl:65-94 + Oceananigans apply_z_bcs!). Inactive cells get G = 0.
 * params may be NULL; with params->radiative_forcing two-colour, Gᵀ also receives the
 * TwoColorRadiation forcing of net->shortwave_surface_flux (radiative_forcing.jl:52-71). */
int ne_tracer_tendency_weno5_f64(ne_ctx* ctx, const ne_grid* grid, const ne_params* params,
                                 const ne_ocean_state* ocean, const ne_net_ocean_fluxes* net,
                                 ne_view G_T, ne_view G_S, void* stream);

/* One Quasi-Adams-Bashforth 2 step (the reference's timestepper = :QuasiAdamsBashforth2,
 * ocean_simulation.jl:323; Oceananigans' QuasiAdamsBashforth2TimeStepper, whose χ is 0.1 there,
 * NE_AB2_CHI) from G, the tendency of the state T, S (ne_tracer_tendency_weno5_f64), and G_prev,
 * the last step's: on active cells c ← c + Δt·((3/2 + χ)·G − (1/2 + χ)·G_prev) for T and S,
 * inactive cells left untouched. euler ≠ 0 (the first step, or a Δt other than the last
 * step's) takes χ = −1/2 and does not read G_prev (its views may then have NULL data). The
 * caller keeps G as the next step's G_prev. All views FP64; T and S are updated in place and
 * may share storage with none of the others (NE_ERR_SHAPE). */
#define NE_AB2_CHI 0.1
int ne_tracer_ab2_update_f64(ne_ctx* ctx, const ne_grid* grid, ne_view T, ne_view S, ne_view G_T, ne_view G_S,
                             ne_view G_prev_T, ne_view G_prev_S, double dt, double chi, int32_t euler,
                             void* stream);

/* One split Runge-Kutta stage, tendency fused: c_out = c0 + beta_dt·G(c_in) for T and S
 * on active cells (inactive cells are left untouched). c_out may alias c0 but not c_in. */
int ne_tracer_rk_stage_f64(ne_ctx* ctx, const ne_grid* grid, const ne_params* params,
                           ne_view u, ne_view v, ne_view w,
                           ne_view T0, ne_view S0, ne_view T_in, ne_view S_in,
                           ne_view T_out, ne_view S_out,
                           const ne_net_ocean_fluxes* net, double beta_dt, void* stream);

/* FP32 forms for models on Float32 storage (the mixed form's convention, the precision form of
 * the tracer stencil and the interpolation): prescribed series and 3-D fields (u, v, w, T, S, G,
 * stage buffers) are float; fractional indices, the exchange state and every 2-D flux are double.
 * The interpolation is formed in FP64 from the float series; the WENO5 reconstructions run in FP32;
 * the stage update c0 + beta_dt·G is formed in FP64 and rounded to float to nearest.
 * ne_interp_atmos_f32 / ne_interp_radiation_f32 replace the kernels of
 * interpolate_atmospheric_state.jl:91-137 and interpolate_radiation_state.jl:43-69 on Float32
 * series; ne_tracer_tendency_weno5_f32 / ne_tracer_rk_stage_f32 the tendency and stage of
 * ocean_simulation.jl:65-94, :347 (Oceananigans) on a Float32 grid. */
int ne_interp_atmos_f32(ne_ctx* ctx, const ne_grid* grid, const ne_prescribed_atmosphere* atmos,
                        const double* fi, const double* fj,
                        const ne_exchange_state* out, void* stream);
int ne_interp_radiation_f32(ne_ctx* ctx, const ne_grid* grid, const ne_prescribed_radiation* rad,
                            const double* fi, const double* fj,
                            ne_view downwelling_shortwave, ne_view downwelling_longwave,
                            void* stream);
int ne_tracer_tendency_weno5_f32(ne_ctx* ctx, const ne_grid* grid, const ne_params* params,
                                 const ne_ocean_state* ocean, const ne_net_ocean_fluxes* net,
                                 ne_view G_T, ne_view G_S, void* stream);
int ne_tracer_rk_stage_f32(ne_ctx* ctx, const ne_grid* grid, const ne_params* params,
                           ne_view u, ne_view v, ne_view w,
                           ne_view T0, ne_view S0, ne_view T_in, ne_view S_in,
                           ne_view T_out, ne_view S_out,
                           const ne_net_ocean_fluxes* net, double beta_dt, void* stream);

/* fill_halo_regions! of a cell-centred 3-D field (T or S) after a step, as Oceananigans leaves
 * the tracers of the HydrostaticFreeSurfaceModel that ocean_simulation builds
 * (ocean_simulation.jl:487-499; Oceananigans not vendored): the halo of widths Hx, Hy, Hz
 * (0 … 64) around the view's interior takes, at each point, the value of the interior point it
 * maps to — wrapped in a periodic direction, the nearest interior point in a bounded one and
 * vertically (the zero-gradient halo of the flux and no-flux boundary conditions of the tracers).
 * The caller's memory must hold the halo. The step kernels read neighbours by index and need no
 * halo; this is for consumers of the caller's halo-padded arrays. */
int ne_fill_halo_f64(ne_ctx* ctx, const ne_grid* grid, ne_view c, int32_t Hx, int32_t Hy, int32_t Hz,
                     void* stream);
int ne_fill_halo_f32(ne_ctx* ctx, const ne_grid* grid, ne_view c, int32_t Hx, int32_t Hy, int32_t Hz,
                     void* stream);

/* Vertical velocity from continuity, for callers that evolve u and v: replaces Oceananigans'
 * compute_w_from_continuity! of the HydrostaticFreeSurfaceModel that ocean_simulation builds
 * (ocean_simulation.jl:487-499; Oceananigans not vendored). Per column, w = 0 at the bottom
 * face and w[k+1] = w[k] − (δx(Ax u) + δy(Ay v))/Az on active cells with Ax = Δy·Δz,
 * Ay = Δx·Δz of the cell's faces; faces above inactive cells carry w = 0. u and v are zero on
 * immersed and peripheral faces (the model's convention, as Oceananigans masks them).
 * Writes every face of w (Nz+1 levels). */
int ne_compute_w_from_continuity_f64(ne_ctx* ctx, const ne_grid* grid, ne_view u, ne_view v, ne_view w,
                                     void* stream);

/* Compensated (Neumaier, double-double combine) sum over active cells of
 * weight·V·c, V = Az·Δz. Writes {sum, compensation} (2 doubles) to device `out`. */
int ne_budget_f64(ne_ctx* ctx, const ne_grid* grid, ne_view c, double weight,
                  double* out, void* stream);
int ne_budget_f32(ne_ctx* ctx, const ne_grid* grid, ne_view c, double weight,
                  double* out, void* stream);

/* ne_budget_f64 without any allocation, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). It
 * has no counterpart in the reference. Same kernels and result as ne_budget_f64, bit for
 * bit; the block partials live in a fixed array of the library on each device (1 MiB, held
 * until the process ends) instead of the context's workspace. Grids that need more than
 * NE_BUDGET_NOALLOC_MAX_BLOCKS blocks, ceil(Nx/64)·ceil(Ny/4), get NE_ERR_CAPACITY. Every
 * call on one device shares that array: calls that overlap on one device (other streams,
 * other threads) race, so the caller orders them. */
#define NE_BUDGET_NOALLOC_MAX_BLOCKS 65536
int ne_budget_f64_noalloc(ne_ctx* ctx, const ne_grid* grid, ne_view c, double weight,
                          double* out, void* stream);

/* ---------------------------------------------------------------------------
 * Model: OceanOnlyModel(ocean; atmosphere, radiation) + time_step!
 * (earth_system_model.jl:399-402; time_step_earth_system_model.jl:14-83).
 * The model binds caller-owned state and output views and allocates its own
 * workspace once, at creation; stepping allocates nothing.
 * ------------------------------------------------------------------------- */

typedef struct ne_model_desc {
    int32_t precision;                    /* NE_FP64, or NE_MIXED_FP32: T, S, u, v, w and the
                                             prescribed fields are float; fluxes stay double */
    int32_t fuse_freeze_clamp;            /* reserved, must be 0 */
    ne_grid grid;
    ne_params params;
    ne_prescribed_atmosphere atmosphere;
    ne_prescribed_radiation radiation;    /* ignored unless params.radiation */
    const double* atmos_fi;               /* [Nx] from ne_fractional_indices_f64 */
    const double* atmos_fj;               /* [Ny] */
    const double* radiation_fi;
    const double* radiation_fj;
    ne_ocean_state ocean;                 /* T, S evolve; u, v, w are held fixed */
    ne_interface_fluxes fluxes;           /* written every update_state */
    ne_net_ocean_fluxes net;
    ne_radiative_fluxes radiative;        /* may have NULL data (not written) */
    ne_view frazil_heat;                  /* double 2-D */
    const double* atmosphere_times;       /* HOST: snapshot times (s) */
    int32_t n_atmosphere_times;
    int32_t n_radiation_times;
    const double* radiation_times;        /* HOST */
    void*   workspace;                    /* optional caller-owned device memory for the RK stage
                                             buffers, freeze flags and rounding residual
                                             (ne_model_workspace_bytes), 256-byte aligned
                                             (NE_ERR_SHAPE otherwise);
                                             NULL: the model allocates them itself */
    int64_t workspace_bytes;
    /* Snapshot time indexing (FieldTimeSeries time_indexing; JRA55 atmospheres are Cyclical,
     * DataWrangling/JRA55/JRA55_prescribed_atmosphere.jl:17) and in-memory windows (the
     * InMemory(N) backend of a FieldTimeSeries, JRA55_prescribed_atmosphere.jl:42). */
    int32_t time_indexing;                /* NE_TIME_CLAMPED or NE_TIME_CYCLICAL */
    int32_t atmosphere_slots;             /* 0: every snapshot resident in the atmosphere views;
                                             N ≥ 2: the views hold N snapshot slots, snapshot n in
                                             slot n mod N, uploaded by the caller
                                             (ne_model_snapshots_at, ne_model_snapshot_resident) */
    int32_t radiation_slots;              /* the same for the radiation views */
    int32_t timestepper;                  /* NE_TIMESTEPPER_SPLIT_RK3 (the reference's default) or
                                             NE_TIMESTEPPER_QUASI_ADAMS_BASHFORTH2 (FP64 models only;
                                             the step's tendency and the last one's are kept in the
                                             stage buffers) */
    double  time_period;                  /* NE_TIME_CYCLICAL: the period (s), ≥ last − first time */
} ne_model_desc;

/* Time indexing of the snapshot series. CLAMPED: linear between the two snapshots around t,
 * the first / last snapshot outside the series (the reference's Linear indexing inside the
 * series). CYCLICAL: Oceananigans Cyclical(period): t is taken modulo the period from the
 * first snapshot time, and between the last snapshot and the first one a period later the
 * two are interpolated across the wrap. */
enum { NE_TIME_CLAMPED = 0, NE_TIME_CYCLICAL = 1 };

/* Time stepping of T and S (ocean_simulation's timestepper, ocean_simulation.jl:323, :347):
 * three split Runge-Kutta stages per step, or one Quasi-Adams-Bashforth 2 update (χ = NE_AB2_CHI;
 * forward Euler on the first step and on a step whose Δt differs from the last step's). */
enum { NE_TIMESTEPPER_SPLIT_RK3 = 0, NE_TIMESTEPPER_QUASI_ADAMS_BASHFORTH2 = 1 };

/* Bytes of device workspace a model of this descriptor needs: four RK stage buffers of T, S,
 * one byte per column of freeze flags (the step's last stage marks the columns holding a
 * cell below its freezing point, and the freeze clamp after the step visits only those), and
 * the rounding residual (ne_model_rounding_residual: four doubles per 63×4 tile), the list of
 * columns whose FP32 flux iteration the mixed form finishes in a second kernel (one int per
 * column) and the two-colour forcing's Beer's-law factors at the Nz+1 faces. */
int ne_model_workspace_bytes(const ne_model_desc* desc, int64_t* bytes);

typedef struct ne_model ne_model;

/* Create the model and run reconcile_state! (fluxes computed at construction,
 * earth_system_model.jl:86-90), i.e. one update_state with dt_frazil = +INFINITY.
 * With desc->workspace set, creation and destruction make no device allocation. */
int ne_model_create(ne_ctx* ctx, const ne_model_desc* desc, ne_model** model, void* stream);
int ne_model_destroy(ne_model* model);

/* time_step!(coupled_model, Δt): ocean split-RK3 step with the current fluxes, clock tick,
 * then update_state! (interpolate, flux solve, freeze clamp, assemble, radiation). */
int ne_model_time_step(ne_model* model, double dt, void* stream);

/* update_state! alone (phases 1-4). */
int ne_model_update_state(ne_model* model, void* stream);

/* Recompute the model's w from its u and v by continuity (ne_compute_w_from_continuity_f64,
 * in the model's precision, accumulated in FP64), for callers that set new horizontal
 * velocities between steps. */
int ne_model_update_vertical_velocity(ne_model* model, void* stream);

/* Snapshot time interpolation of a series (Oceananigans.OutputReaders
 * interpolating_time_indices for Linear / Cyclical time indexing, the function a
 * FieldTimeSeries evaluates at a time; cpu_interpolating_time_indices,
 * src/Atmospheres/interpolate_atmospheric_state.jl:57-60): `interp` receives the two 0-based
 * snapshots and the weight of the second, `positions` their stream positions (see
 * ne_model_snapshots_at). The same function the models and the device clock evaluate.
 * NE_ERR_SHAPE for an empty series or a CYCLICAL period that is not positive or shorter
 * than the series. Host only; needs no device. */
int ne_snapshot_time_indices(const double* times, int32_t n_times, int32_t time_indexing, double time_period,
                             double time, ne_time_interp* interp, int64_t positions[2]);

/* Stream positions of the snapshots the model interpolates between at model time `time`:
 * out[0], out[1] of the atmosphere, out[2], out[3] of the radiation (INT64_MIN for a series
 * with no snapshots). A position p is snapshot p mod n_times of its series (floor modulo) and
 * fills slot p mod slots of windowed views. With CLAMPED indexing p is the snapshot index; with
 * CYCLICAL indexing p = cycle · n_times + index, so the interpolation across the wrap reads two
 * consecutive positions (the last snapshot and the first one of the next cycle) and a ring of
 * two slots never has both in one slot. A caller streaming snapshots into windowed views
 * (atmosphere_slots / radiation_slots) asks this for the time after the next step and uploads
 * the missing ones first. Host only.
 * (FieldTimeSeries time_indices / update_field_time_series!, Oceananigans.OutputReaders;
 * the reference's InMemory windows are refilled through set!, JRA55_field_time_series.jl:64-80.) */
int ne_model_snapshots_at(const ne_model* model, double time, int64_t out[4]);

/* Record that the snapshot at stream position p of the atmosphere (which = 0) or the
 * radiation (which = 1) now fills its slot p mod slots of the windowed views; the caller
 * orders the upload before the next step on that step's stream. A step (or update_state)
 * whose snapshots are not resident returns NE_ERR_STATE before enqueueing anything. At
 * creation the slots are taken to hold the N consecutive positions from p₁, the first one
 * read at time 0 (with CLAMPED indexing the run is moved to fit inside the series), as an
 * in-memory window is filled from the snapshot the current time interpolates from; the
 * caller fills them before ne_model_create (ne_snapshot_time_indices at time 0 gives p₁).
 * Windowed models cannot join a batch
 * (NE_ERR_UNSUPPORTED). NE_ERR_STATE for a series that is not windowed, NE_ERR_SHAPE for a
 * CLAMPED position outside 0 … n_times − 1. Host only. */
int ne_model_snapshot_resident(ne_model* model, int32_t which, int64_t position);

/* Clock (seconds) and iteration count of the model. Host only. */
int ne_model_clock(const ne_model* model, double* time, int64_t* iteration);

/* Rounding residual of a NE_MIXED_FP32 model's state: since creation, Σ V·(x − stored) over the
 * active cells of every step's last RK stage, for T (out[0] + out[1], K·m³) and S (out[2] +
 * out[3], psu·m³), each a compensated (sum, compensation) pair, where x is the stage's FP64
 * update and `stored` the float written (the last stage differences its FP32 face fluxes and forms
 * x in FP64, so Σ V·(x − c⁰) telescopes to the boundary fluxes). The budget of the FP32 state
 * plus this residual is conserved to FP64 round-off: ρ₀c₀·(Σ V·T + out[0] + out[1]) changes by
 * −Δt·ρ₀c₀·Σ Az·Jᵀ per step, as the FP64 state's heat content does
 * (test/test_tracer_budget.jl:66-76 checks that closure for the reference). Zero for an FP64
 * model, whose update is stored exactly. The last stage adds each tile's share to the tile's
 * running totals, so stepping carries no reduction; this call sums the tiles (one kernel, in a
 * fixed order) into `out`, device memory of 4 doubles, in stream order on `stream`. Steps taken
 * in a batch (ne_batch_time_step) add to the same totals. */
int ne_model_rounding_residual(const ne_model* model, double* out, void* stream);

/* Per-kernel timing of ne_model_time_step with HIP events recorded on the step's own
 * stream (no counterpart in the reference, whose examples time whole steps in callbacks,
 * examples/near_global_ocean_simulation.jl:109-132). begin creates the events for up to
 * max_steps steps; end waits for them and writes the summed milliseconds of each kernel
 * to ms[NE_TIMED_KERNELS] and the number of steps recorded to *steps. Kernels, in order:
 * RK stage 1, RK stage 2, RK stage 3, surface update (interpolation + flux solve, then per
 * column the freeze clamp, net fluxes other than τ and radiation), surface τ (the momentum
 * fluxes averaged onto the faces). */
#define NE_TIMED_KERNELS 5
int ne_model_timing_begin(ne_model* model, int max_steps);
int ne_model_timing_end(ne_model* model, double* ms, int* steps);

/* ---------------------------------------------------------------------------
 * Batches: ensembles of models stepped together (no counterpart in the reference,
 * which steps one model per time_step! call; this is the form for calibration
 * ensembles and request serving). A batch binds existing models of one precision;
 * each ne_batch_time_step is time_step!(model, Δt) for every member, issued as six
 * persistent-grid launches for the whole ensemble (optionally one HIP graph launch).
 * Results are identical to stepping each member with ne_model_time_step. While a
 * model belongs to a batch, ne_model_time_step / ne_model_update_state /
 * ne_model_destroy on it return NE_ERR_STATE; ne_batch_destroy releases the members.
 * ------------------------------------------------------------------------- */

typedef struct ne_batch ne_batch;

/* Launch shape of a batch step. tile_rows ∈ {4, 8, 16} is the tile height of the WENO5
 * stage kernel (tiles are 63 columns wide, one wavefront per row); stage_blocks (a
 * multiple of 8) is its persistent grid; surface_block ∈ {64, 128, 256} columns per work
 * item and surface_blocks the grid of the two surface kernels; a grid of 0 means one
 * block per work item. stage_ms / surface_ms are the timings the autotuner measured
 * (0 when set by hand). */
typedef struct ne_batch_tuning {
    int32_t tile_rows;
    int32_t stage_blocks;
    int32_t surface_block;
    int32_t surface_blocks;
    double  stage_ms;
    double  surface_ms;
} ne_batch_tuning;

/* A batch of `precision` (NE_FP64 or NE_MIXED_FP32) holding the n_models given models
 * (n_models may be 0; members step with split RK3: a Quasi-Adams-Bashforth 2 model is refused
 * with NE_ERR_UNSUPPORTED, as is one with windowed snapshot views). All calls on one batch must
 * be ordered on one stream. */
int ne_batch_create(ne_ctx* ctx, int32_t precision, int n_models, ne_model* const* models, ne_batch** batch,
                    void* stream);
/* Waits for the device, then releases the batch's device memory; the members are released. */
int ne_batch_destroy(ne_batch* batch);
/* Membership changes take effect at the next step (the member table is re-uploaded with one
 * stream-ordered copy; nothing synchronises the device). A model added joins at its own clock. */
int ne_batch_add(ne_batch* batch, ne_model* model);
int ne_batch_remove(ne_batch* batch, ne_model* model);
int ne_batch_time_step(ne_batch* batch, double dt, void* stream);
/* enabled != 0: capture the step once per Δt into a HIP graph and replay it. */
int ne_batch_set_graph(ne_batch* batch, int enabled);
/* Times every candidate shape on the batch itself (RK stage 1 into the stage buffers and
 * the idempotent surface solve; the model state is unchanged) and keeps the fastest. */
int ne_batch_autotune(ne_batch* batch, ne_batch_tuning* chosen, void* stream);
int ne_batch_get_tuning(const ne_batch* batch, ne_batch_tuning* tuning);
int ne_batch_set_tuning(ne_batch* batch, const ne_batch_tuning* tuning);
int ne_batch_info(const ne_batch* batch, int* members, int64_t* cells, int* stage_items, int* surface_items);

#ifdef __cplusplus
}
#endif

#endif /* NUMERICALEARTH_HIP_H */
