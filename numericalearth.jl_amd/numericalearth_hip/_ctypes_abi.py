"""ctypes mirror of the C ABI types declared in include/numericalearth_hip.h.

The structures here are plain data: they carry device (or, for the CPU oracle,
host) pointers and sizes, never torch objects. ``tests/test_abi_layout.py``
checks every size and field offset against the C header compiled by gcc.
"""
from __future__ import annotations

import ctypes as C
import math

NE_OK = 0
NE_ERR_NULL = 1
NE_ERR_SHAPE = 2
NE_ERR_UNSUPPORTED = 3
NE_ERR_DEVICE = 4
NE_ERR_CAPACITY = 5
NE_ERR_STATE = 6

STATUS_NAMES = {
    NE_OK: "NE_OK",
    NE_ERR_NULL: "NE_ERR_NULL",
    NE_ERR_SHAPE: "NE_ERR_SHAPE",
    NE_ERR_UNSUPPORTED: "NE_ERR_UNSUPPORTED",
    NE_ERR_DEVICE: "NE_ERR_DEVICE",
    NE_ERR_CAPACITY: "NE_ERR_CAPACITY",
    NE_ERR_STATE: "NE_ERR_STATE",
}

NE_PERIODIC, NE_BOUNDED, NE_FLAT = 0, 1, 2
NE_FP64, NE_MIXED_FP32 = 0, 1
NE_RADIATIVE_FORCING_NONE, NE_RADIATIVE_FORCING_TWO_COLOR = 0, 1
NE_FLUXES_SIMILARITY_THEORY, NE_FLUXES_COEFFICIENT_BASED = 0, 1
NE_TRANSFER_SCALES, NE_TRANSFER_LARGE_YEAGER = 0, 1
NE_COEFFICIENT_CONSTANT, NE_COEFFICIENT_POLYNOMIAL = 0, 1
NE_TEMPERATURE_BULK, NE_TEMPERATURE_SKIN = 0, 1
NE_STABILITY_NONE, NE_STABILITY_EDSON = 0, 1
NE_SIMILARITY_LOG, NE_SIMILARITY_COARE_LOG = 0, 1
NE_ROUGHNESS_CONSTANT, NE_ROUGHNESS_EDSON = 0, 1
NE_STOP_CONVERGENCE, NE_STOP_FIXED_ITERATIONS = 0, 1
NE_HUMIDITY_IMPURE_SATURATION, NE_HUMIDITY_FIXED = 0, 1
NE_VELOCITY_RELATIVE, NE_VELOCITY_WIND = 0, 1
NE_TIME_CLAMPED, NE_TIME_CYCLICAL = 0, 1
NE_TIMESTEPPER_SPLIT_RK3, NE_TIMESTEPPER_QUASI_ADAMS_BASHFORTH2 = 0, 1
NE_AB2_CHI = 0.1
NE_BUDGET_NOALLOC_MAX_BLOCKS = 65536   # blocks of 64×4 columns the allocation-free budget holds partials for
NO_POSITION = -(2 ** 63)   # ne_model_snapshots_at: a series with no snapshots

_d = C.c_double
_i32 = C.c_int32
_i64 = C.c_int64
_p = C.c_void_p
_pd = C.POINTER(C.c_double)
_pi32 = C.POINTER(C.c_int32)


class View(C.Structure):
    _fields_ = [("data", _p), ("sx", _i64), ("sy", _i64), ("sz", _i64)]


class Grid(C.Structure):
    _fields_ = [
        ("Nx", _i32), ("Ny", _i32), ("Nz", _i32),
        ("topology_x", _i32), ("topology_y", _i32), ("advection", _i32),
        ("dz", _p), ("z_center", _p), ("dy_xface", _p), ("dx_yface", _p), ("Az", _p),
        ("lambda_c", _p), ("phi_c", _p), ("kmin", _p), ("z_face", _p),
        ("halo", _i32),
    ]


class SourceGrid(C.Structure):
    _fields_ = [
        ("Nx", _i32), ("Ny", _i32), ("topology_x", _i32), ("topology_y", _i32),
        ("lambda0", _d), ("dlambda", _d), ("phi0", _d), ("dphi", _d),
    ]


class TimeInterp(C.Structure):
    _fields_ = [("n1", _i32), ("n2", _i32), ("tilde_n", _d)]


class PrescribedAtmosphereDesc(C.Structure):
    _fields_ = [
        ("grid", SourceGrid),
        ("u", View), ("v", View), ("T", View), ("q", View), ("p", View), ("rain", View), ("snow", View),
        ("time", TimeInterp),
    ]


class PrescribedRadiationDesc(C.Structure):
    _fields_ = [
        ("grid", SourceGrid),
        ("downwelling_shortwave", View), ("downwelling_longwave", View),
        ("time", TimeInterp),
    ]


class ExchangeState(C.Structure):
    _fields_ = [("u", View), ("v", View), ("T", View), ("p", View), ("q", View), ("rain", View), ("snow", View),
                ("downwelling_shortwave", View), ("downwelling_longwave", View)]


class OceanState(C.Structure):
    _fields_ = [("u", View), ("v", View), ("w", View), ("T", View), ("S", View)]


class InterfaceFluxes(C.Structure):
    _fields_ = [
        ("latent_heat", View), ("sensible_heat", View), ("water_vapor", View),
        ("x_momentum", View), ("y_momentum", View),
        ("friction_velocity", View), ("temperature_scale", View), ("water_vapor_scale", View),
        ("interface_temperature", View), ("iterations", View),
    ]


class NetOceanFluxes(C.Structure):
    _fields_ = [("tau_x", View), ("tau_y", View), ("J_T", View), ("J_S", View), ("J_w", View), ("J_H", View),
                ("shortwave_surface_flux", View)]


class RadiativeFluxes(C.Structure):
    _fields_ = [("upwelling_longwave", View), ("downwelling_longwave", View), ("downwelling_shortwave", View)]


class Params(C.Structure):
    _fields_ = [
        ("von_karman_constant", _d),
        ("turbulent_prandtl_number", _d),
        ("convective_gustiness", _i32),
        ("mesoscale_gustiness", _i32),
        ("gustiness_parameter", _d),
        ("minimum_gustiness", _d),
        ("mesoscale_velocity", _d),
        ("stability_functions", _i32),
        ("similarity_form", _i32),
        ("edson_momentum", _d * 11),
        ("edson_scalar", _d * 12),
        ("momentum_roughness", _i32),
        ("temperature_roughness", _i32),
        ("water_vapor_roughness", _i32),
        ("reserved0", _i32),
        ("momentum_roughness_length", _d),
        ("temperature_roughness_length", _d),
        ("water_vapor_roughness_length", _d),
        ("roughness_gravitational_acceleration", _d),
        ("air_kinematic_viscosity", _d),
        ("charnock_parameter", _d),
        ("smooth_wall_parameter", _d),
        ("maximum_momentum_roughness", _d),
        ("reynolds_scaling_A", _d),
        ("reynolds_scaling_b", _d),
        ("maximum_scalar_roughness", _d),
        ("zero_plane_displacement", _d),
        ("stop_criteria", _i32),
        ("maxiter", _i32),
        ("tolerance", _d),
        ("specific_humidity", _i32),
        ("velocity_difference", _i32),
        ("water_mole_fraction", _d),
        ("fixed_specific_humidity", _d),
        ("surface_layer_height", _d),
        ("boundary_layer_height", _d),
        ("gravitational_acceleration", _d),
        ("gas_constant", _d),
        ("dry_air_molar_mass", _d),
        ("water_molar_mass", _d),
        ("dry_air_adiabatic_exponent", _d),
        ("water_vapor_heat_capacity", _d),
        ("liquid_water_heat_capacity", _d),
        ("reference_vaporization_enthalpy", _d),
        ("reference_temperature", _d),
        ("triple_point_temperature", _d),
        ("triple_point_pressure", _d),
        ("ocean_reference_density", _d),
        ("ocean_heat_capacity", _d),
        ("freshwater_density", _d),
        ("radiation", _i32),
        ("radiative_forcing", _i32),
        ("stefan_boltzmann_constant", _d),
        ("ocean_albedo", _d),
        ("ocean_emissivity", _d),
        ("freshwater_melting_temperature", _d),
        ("liquidus_slope", _d),
        ("first_color_fraction", _d),
        ("first_absorption_coefficient", _d),
        ("second_absorption_coefficient", _d),
        ("turbulent_fluxes", C.c_int32),
        ("transfer_coefficients", C.c_int32),
        ("transfer_kind", C.c_int32 * 3),
        ("interface_temperature", C.c_int32),
        ("transfer_coefficient", _d * 3),
        ("drag_polynomial", _d * 4),
        ("high_wind_speed_threshold", _d),
        ("high_wind_drag_coefficient", _d),
        ("minimum_wind_speed", _d),
        ("large_yeager_von_karman_constant", _d),
        ("large_yeager_reference_height", _d),
        ("stable_heat_transfer_coefficient", _d),
        ("unstable_heat_transfer_coefficient", _d),
        ("moisture_transfer_coefficient", _d),
        ("paulson_momentum", _d * 2),
        ("paulson_scalar", _d),
        ("linear_stable_momentum", _d * 2),
        ("linear_stable_scalar", _d * 2),
        ("skin_diffusivity", _d),
        ("skin_depth", _d),
        ("skin_max_temperature_change", _d),
        ("reference_sublimation_enthalpy", _d),
        ("water_ice_heat_capacity", _d),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("total_iterations", C.c_uint64),
        ("solved_columns", C.c_uint32),
        ("nonconverged_columns", C.c_uint32),
        ("nan_columns", C.c_uint32),
        ("max_iterations", C.c_uint32),
    ]


class ModelDesc(C.Structure):
    _fields_ = [
        ("precision", _i32),
        ("fuse_freeze_clamp", _i32),
        ("grid", Grid),
        ("params", Params),
        ("atmosphere", PrescribedAtmosphereDesc),
        ("radiation", PrescribedRadiationDesc),
        ("atmos_fi", _p), ("atmos_fj", _p), ("radiation_fi", _p), ("radiation_fj", _p),
        ("ocean", OceanState),
        ("fluxes", InterfaceFluxes),
        ("net", NetOceanFluxes),
        ("radiative", RadiativeFluxes),
        ("frazil_heat", View),
        ("atmosphere_times", _pd),
        ("n_atmosphere_times", _i32),
        ("n_radiation_times", _i32),
        ("radiation_times", _pd),
        ("workspace", _p),
        ("workspace_bytes", C.c_int64),
        ("time_indexing", _i32),
        ("atmosphere_slots", _i32),
        ("radiation_slots", _i32),
        ("timestepper", _i32),
        ("time_period", C.c_double),
    ]


# Default constants of the reference (cited where each is defined).
EDSON_MOMENTUM = (  # similarity_theory_turbulent_fluxes.jl:487-499
    50.0, 0.35, 0.7, 0.75, 5 / 0.35, 15.0, 2.0, math.pi / 2, 10.15, 3.0, math.pi / math.sqrt(3),
)
EDSON_SCALAR = (  # similarity_theory_turbulent_fluxes.jl:571-584
    50.0, 0.35, 2 / 3, 3 / 2, 14.28, 8.525, 15.0, 2.0, 0.0, 34.15, 3.0, math.pi / math.sqrt(3),
)
TEOS10_REFERENCE_HEAT_CAPACITY = 3991.86795711963  # SeawaterPolynomials TEOS-10 cₚ⁰ (ocean_simulation.jl:527-530)
GRAVITATIONAL_ACCELERATION = 9.80665  # Oceananigans.defaults (EarthSystemModels.jl:70)


class BatchTuning(C.Structure):
    _fields_ = [
        ("tile_rows", _i32),
        ("stage_blocks", _i32),
        ("surface_block", _i32),
        ("surface_blocks", _i32),
        ("stage_ms", C.c_double),
        ("surface_ms", C.c_double),
    ]
