"""Flux-formulation types of the reference, mirrored for the HIP path.

Each class carries the reference type's name, keyword arguments and defaults
(file:line of the reference given per class) and lowers to the flat ``ne_params``
block of the C ABI via :func:`to_params`. Formulations the HIP path does not
implement raise ``NotImplementedError`` when lowered, never silently.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Any

from . import _ctypes_abi as abi

GRAVITATIONAL_ACCELERATION = abi.GRAVITATIONAL_ACCELERATION


@dataclass
class ConvergenceStopCriteria:  # compute_interface_state.jl:5-8
    tolerance: float = 1e-8
    maxiter: int = 100


@dataclass
class FixedIterations:  # compute_interface_state.jl:20-22
    iterations: int


@dataclass
class ConvectiveGustiness:  # similarity_theory_turbulent_fluxes.jl:45-48
    gustiness_parameter: float = 1.2
    minimum_gustiness: float = 0.01


@dataclass
class SubgridVelocityCorrection:  # similarity_theory_turbulent_fluxes.jl:69-80
    convective: Any = field(default_factory=ConvectiveGustiness)
    mesoscale: Any = None


def mahrt_sun_subgrid_velocity(dx: float, threshold: float = 5e3) -> float:
    """Mahrt & Sun (1995) mesoscale velocity (similarity_theory_turbulent_fluxes.jl:114-117)."""
    d = max(dx / threshold - 1, 0)
    return 0.32 * d ** 0.33


@dataclass
class EdsonMomentumStabilityFunction:  # similarity_theory_turbulent_fluxes.jl:487-499
    ζmax: float = 50.0
    A_plus: float = 0.35
    B_plus: float = 0.7
    C_plus: float = 0.75
    D_plus: float = 5 / 0.35
    A_minus: float = 15.0
    B_minus: float = 2.0
    C_minus: float = math.pi / 2
    D_minus: float = 10.15
    E_minus: float = 3.0
    F_minus: float = math.pi / math.sqrt(3)

    def values(self):
        return (self.ζmax, self.A_plus, self.B_plus, self.C_plus, self.D_plus, self.A_minus, self.B_minus, self.C_minus, self.D_minus, self.E_minus, self.F_minus)


@dataclass
class EdsonScalarStabilityFunction:  # similarity_theory_turbulent_fluxes.jl:571-584
    ζmax: float = 50.0
    A_plus: float = 0.35
    B_plus: float = 2 / 3
    C_plus: float = 3 / 2
    D_plus: float = 14.28
    E_plus: float = 8.525
    A_minus: float = 15.0
    B_minus: float = 2.0
    C_minus: float = 0.0
    D_minus: float = 34.15
    E_minus: float = 3.0
    F_minus: float = math.pi / math.sqrt(3)

    def values(self):
        return (self.ζmax, self.A_plus, self.B_plus, self.C_plus, self.D_plus, self.E_plus, self.A_minus, self.B_minus, self.C_minus, self.D_minus, self.E_minus,
                self.F_minus)


@dataclass
class SimilarityScales:  # similarity_theory_turbulent_fluxes.jl:432-436
    momentum: Any
    temperature: Any
    water_vapor: Any


def atmosphere_ocean_stability_functions():  # similarity_theory_turbulent_fluxes.jl:621-625
    ψc = EdsonScalarStabilityFunction()
    return SimilarityScales(EdsonMomentumStabilityFunction(), ψc, ψc)


def zero_stability_function(ζ):
    """ψ ≡ 0 (the neutral-limit mock of test_surface_fluxes.jl:160)."""
    return 0.0


@dataclass
class MomentumRoughnessLength:  # roughness_lengths.jl:123-139
    gravitational_acceleration: float = GRAVITATIONAL_ACCELERATION
    maximum_roughness_length: float = 1.0
    air_kinematic_viscosity: float = 1.5e-5
    wave_formulation: float = 0.02
    smooth_wall_parameter: float = 0.11


@dataclass
class ReynoldsScalingFunction:  # roughness_lengths.jl:228-231
    A: float = 5.85e-5
    b: float = 0.72


@dataclass
class ScalarRoughnessLength:  # roughness_lengths.jl:93-101
    air_kinematic_viscosity: float = 1.5e-5
    reynolds_number_scaling_function: ReynoldsScalingFunction = field(default_factory=ReynoldsScalingFunction)
    maximum_roughness_length: float = 1.6e-4


class LogarithmicSimilarityProfile:  # similarity_theory_turbulent_fluxes.jl:239
    pass


class COARELogarithmicSimilarityProfile:  # similarity_theory_turbulent_fluxes.jl:240
    pass


@dataclass
class SimilarityTheoryFluxes:  # similarity_theory_turbulent_fluxes.jl:174-214
    von_karman_constant: float = 0.4
    turbulent_prandtl_number: float = 1.0
    subgrid_velocities: Any = field(default_factory=ConvectiveGustiness)
    stability_functions: Any = field(default_factory=atmosphere_ocean_stability_functions)
    momentum_roughness_length: Any = field(default_factory=MomentumRoughnessLength)
    temperature_roughness_length: Any = field(default_factory=ScalarRoughnessLength)
    water_vapor_roughness_length: Any = field(default_factory=ScalarRoughnessLength)
    zero_plane_displacement: float = 0.0
    similarity_form: Any = field(default_factory=LogarithmicSimilarityProfile)
    solver_stop_criteria: Any = None
    solver_tolerance: float = 1e-8
    solver_maxiter: int = 100

    def __post_init__(self):
        if self.solver_stop_criteria is None:
            self.solver_stop_criteria = ConvergenceStopCriteria(self.solver_tolerance, self.solver_maxiter)
        if self.stability_functions is None:
            z = zero_stability_function
            self.stability_functions = SimilarityScales(z, z, z)


# ---------------------------------------------------------------------------
# Coefficient-based fluxes (coefficient_based_turbulent_fluxes.jl) and the Large-Yeager
# stability functions (similarity_theory_turbulent_fluxes.jl:680-778)
# ---------------------------------------------------------------------------

@dataclass
class PaulsonMomentumStabilityFunction:  # similarity_theory_turbulent_fluxes.jl:683-686
    a: float = 16.0
    b: float = math.pi / 2


@dataclass
class PaulsonScalarStabilityFunction:  # similarity_theory_turbulent_fluxes.jl:701-703
    a: float = 16.0


@dataclass
class LinearStableStabilityFunction:  # similarity_theory_turbulent_fluxes.jl:742-745
    coefficient: float = 5.0
    maximum_stability_parameter: float = 10.0


@dataclass
class SplitStabilityFunction:  # similarity_theory_turbulent_fluxes.jl:712-715
    stable: Any
    unstable: Any


def large_yeager_stability_functions():  # similarity_theory_turbulent_fluxes.jl:772-778
    """Paulson (1970) for unstable, ψ = −5ζ bounded at ζ ≤ 10 for stable stratification."""
    stable = LinearStableStabilityFunction()
    momentum = SplitStabilityFunction(stable, PaulsonMomentumStabilityFunction())
    scalar = SplitStabilityFunction(stable, PaulsonScalarStabilityFunction())
    return SimilarityScales(momentum, scalar, scalar)


@dataclass
class PolynomialNeutralDragCoefficient:  # coefficient_based_turbulent_fluxes.jl:20-46
    """C_dN(U) = (a/U + b + cU − dU⁶)/1000 below the high-wind threshold (Large & Yeager 2004)."""
    a: float = 2.7
    b: float = 0.142
    c: float = 1 / 13.09
    d: float = 3.14807e-10
    high_wind_speed_threshold: float = 33.0
    high_wind_drag_coefficient: float = 2.34e-3
    minimum_wind_speed: float = 0.5

    def values(self):
        return (self.a, self.b, self.c, self.d, self.high_wind_speed_threshold, self.high_wind_drag_coefficient,
                self.minimum_wind_speed)


@dataclass
class LargeYeagerTransferCoefficients:  # coefficient_based_turbulent_fluxes.jl:82-108
    von_karman_constant: float = 0.4
    neutral_drag_coefficient: Any = field(default_factory=PolynomialNeutralDragCoefficient)
    stability_functions: Any = field(default_factory=large_yeager_stability_functions)
    reference_height: float = 10.0
    stable_heat_transfer_coefficient: float = 18.0
    unstable_heat_transfer_coefficient: float = 32.7
    moisture_transfer_coefficient: float = 34.6


_TRANSFER_KEYS = ("momentum", "temperature", "water_vapor")


@dataclass
class CoefficientBasedFluxes:  # coefficient_based_turbulent_fluxes.jl:142-229
    """Bulk fluxes from transfer coefficients: a 3-tuple, a dict with keys momentum,
    temperature, water_vapor, a SimilarityScales (entries constant numbers or a
    PolynomialNeutralDragCoefficient) or a LargeYeagerTransferCoefficients."""
    transfer_coefficients: Any = (1e-3, 1e-3, 1e-3)
    solver_stop_criteria: Any = None
    solver_tolerance: float = 1e-8
    solver_maxiter: int = 20

    def __post_init__(self):
        tc = self.transfer_coefficients
        if isinstance(tc, dict):          # validate_coefficients(::NamedTuple) (:234-246)
            missing = [k for k in _TRANSFER_KEYS if k not in tc]
            if missing:
                raise ValueError(f"Transfer coefficients must contain keys {_TRANSFER_KEYS}. Missing keys: "
                                 f"{tuple(missing)}. Received: {tuple(tc)}")
        elif isinstance(tc, (tuple, list)):  # validate_coefficients(::Tuple) (:248-258)
            if len(tc) != 3:
                raise ValueError("Transfer coefficients must be a tuple of length 3: (momentum, temperature, "
                                 f"water_vapor). Got length {len(tc)} with value {tuple(tc)}.")
        if self.solver_stop_criteria is None:
            self.solver_stop_criteria = ConvergenceStopCriteria(self.solver_tolerance, self.solver_maxiter)

    def coefficients(self):
        """(momentum, temperature, water vapour) entries of a scales-type coefficient set."""
        tc = self.transfer_coefficients
        if isinstance(tc, dict):
            return tuple(tc[k] for k in _TRANSFER_KEYS)
        if isinstance(tc, SimilarityScales):
            return (tc.momentum, tc.temperature, tc.water_vapor)
        return tuple(tc)


@dataclass
class ImpureSaturationSpecificHumidity:  # interface_states.jl:20-47; default component_interfaces.jl:353-358
    phase: str = "liquid"
    water_mole_fraction: float = 0.98


@dataclass
class FixedSpecificHumidity:
    """Constant interface humidity (the FixedSpecificHumidity mock of test_surface_fluxes.jl:28-32)."""
    q: float


class BulkTemperature:  # interface_states.jl:330-333
    pass


@dataclass
class InteriorDiffusivity:  # interface_states.jl:378-392
    minimum_diffusivity: float = 1.4e-7


@dataclass
class DiffusiveFlux:  # interface_states.jl:366-376
    """Internal flux J = −κ (Tₛ − Tᵢ)/δ between the top-cell temperature, δ below the
    interface, and the interface temperature."""
    κ: Any
    δ: float


@dataclass
class SkinTemperature:  # interface_states.jl:355-364
    """Interface temperature from the flux balance F(Tₛ) − Jᵀ = 0 (semi-implicit in the
    sensible heat flux, each iterate's step clamped to max_ΔT)."""
    internal_flux: Any
    max_ΔT: float = 5.0


class RelativeVelocity:  # interface_states.jl:287
    pass


class WindVelocity:  # interface_states.jl:284
    pass


@dataclass
class AtmosphereThermodynamicsParameters:  # src/Atmospheres/thermodynamic_parameters.jl
    gas_constant: float = 8.3144598
    dry_air_molar_mass: float = 0.02897
    water_molar_mass: float = 0.018015
    dry_air_adiabatic_exponent: float = 2 / 7
    water_vapor_heat_capacity: float = 1859.0
    liquid_water_heat_capacity: float = 4181.0
    reference_vaporization_enthalpy: float = 2500800.0
    reference_sublimation_enthalpy: float = 2834400.0
    water_ice_heat_capacity: float = 2100.0
    reference_temperature: float = 273.16
    triple_point_temperature: float = 273.16
    triple_point_pressure: float = 611.657


@dataclass
class LinearLiquidus:  # ClimaSeaIce LinearLiquidus defaults
    freshwater_melting_temperature: float = 0.0
    slope: float = 0.054


@dataclass
class SurfaceRadiationProperties:  # prescribed_radiation.jl:68 (ocean surface default)
    albedo: float = 0.05
    emissivity: float = 0.97


@dataclass
class TwoColorRadiation:
    """TwoColorRadiation (Oceans/radiative_forcing.jl:5-46): shortwave that penetrates the ocean as two
    colours decaying by Beer's law, I(z) = ε₁ I₀ exp(κ₁ z) + (1 − ε₁) I₀ exp(κ₂ z)."""
    first_color_fraction: float = 0.58
    first_absorption_coefficient: float = 1 / 0.35
    second_absorption_coefficient: float = 1 / 23


def default_radiative_forcing() -> TwoColorRadiation:
    """Paulson & Simpson (1977) Jerlov Type I (ocean_simulation.jl:158-167)."""
    return TwoColorRadiation(first_color_fraction=0.58, first_absorption_coefficient=1 / 0.35,
                             second_absorption_coefficient=1 / 23)


@dataclass
class ComponentInterfaces:
    """The keyword arguments of ComponentInterfaces (component_interfaces.jl:390-419) on this path."""
    atmosphere_ocean_fluxes: Any = field(default_factory=SimilarityTheoryFluxes)
    atmosphere_ocean_interface_temperature: Any = field(default_factory=BulkTemperature)
    atmosphere_ocean_velocity_difference: Any = field(default_factory=RelativeVelocity)
    atmosphere_ocean_interface_specific_humidity: Any = field(default_factory=ImpureSaturationSpecificHumidity)
    ocean_reference_density: float = 1020.0
    ocean_heat_capacity: float = abi.TEOS10_REFERENCE_HEAT_CAPACITY
    freshwater_density: float = 1000.0
    gravitational_acceleration: float = GRAVITATIONAL_ACCELERATION


def _roughness(kind_obj, which: str):
    if isinstance(kind_obj, (int, float)):
        return abi.NE_ROUGHNESS_CONSTANT, float(kind_obj)
    if which == "momentum" and isinstance(kind_obj, MomentumRoughnessLength):
        return abi.NE_ROUGHNESS_EDSON, 0.0
    if which == "scalar" and isinstance(kind_obj, ScalarRoughnessLength):
        return abi.NE_ROUGHNESS_EDSON, 0.0
    raise NotImplementedError(f"roughness length {type(kind_obj).__name__} is not implemented by the HIP path")


def _set_polynomial(p: abi.Params, poly: PolynomialNeutralDragCoefficient) -> None:
    for n in range(4):
        p.drag_polynomial[n] = float(poly.values()[n])
    p.high_wind_speed_threshold = float(poly.high_wind_speed_threshold)
    p.high_wind_drag_coefficient = float(poly.high_wind_drag_coefficient)
    p.minimum_wind_speed = float(poly.minimum_wind_speed)


def _lower_coefficients(p: abi.Params, fluxes: CoefficientBasedFluxes | None) -> None:
    """The coefficient-based block; with fluxes None, its defaults (as ne_params_default)."""
    ly_default = LargeYeagerTransferCoefficients()
    _set_polynomial(p, ly_default.neutral_drag_coefficient)
    p.large_yeager_von_karman_constant = ly_default.von_karman_constant
    p.large_yeager_reference_height = ly_default.reference_height
    p.stable_heat_transfer_coefficient = ly_default.stable_heat_transfer_coefficient
    p.unstable_heat_transfer_coefficient = ly_default.unstable_heat_transfer_coefficient
    p.moisture_transfer_coefficient = ly_default.moisture_transfer_coefficient
    pm, ps, ls = PaulsonMomentumStabilityFunction(), PaulsonScalarStabilityFunction(), LinearStableStabilityFunction()
    p.paulson_momentum[0], p.paulson_momentum[1], p.paulson_scalar = pm.a, pm.b, ps.a
    for arr in (p.linear_stable_momentum, p.linear_stable_scalar):
        arr[0], arr[1] = ls.coefficient, ls.maximum_stability_parameter
    p.turbulent_fluxes = abi.NE_FLUXES_SIMILARITY_THEORY
    p.transfer_coefficients = abi.NE_TRANSFER_SCALES
    for n in range(3):
        p.transfer_kind[n] = abi.NE_COEFFICIENT_CONSTANT
        p.transfer_coefficient[n] = 1e-3
    if fluxes is None:
        return
    p.turbulent_fluxes = abi.NE_FLUXES_COEFFICIENT_BASED
    tc = fluxes.transfer_coefficients
    if isinstance(tc, LargeYeagerTransferCoefficients):
        p.transfer_coefficients = abi.NE_TRANSFER_LARGE_YEAGER
        if not isinstance(tc.neutral_drag_coefficient, PolynomialNeutralDragCoefficient):
            raise NotImplementedError(f"neutral drag {type(tc.neutral_drag_coefficient).__name__}")
        _set_polynomial(p, tc.neutral_drag_coefficient)
        p.large_yeager_von_karman_constant = float(tc.von_karman_constant)
        p.large_yeager_reference_height = float(tc.reference_height)
        p.stable_heat_transfer_coefficient = float(tc.stable_heat_transfer_coefficient)
        p.unstable_heat_transfer_coefficient = float(tc.unstable_heat_transfer_coefficient)
        p.moisture_transfer_coefficient = float(tc.moisture_transfer_coefficient)
        sf = tc.stability_functions
        for name, ψ, unstable_type in (("momentum", sf.momentum, PaulsonMomentumStabilityFunction),
                                       ("temperature", sf.temperature, PaulsonScalarStabilityFunction)):
            if not (isinstance(ψ, SplitStabilityFunction) and isinstance(ψ.stable, LinearStableStabilityFunction)
                    and isinstance(ψ.unstable, unstable_type)):
                raise NotImplementedError(f"Large-Yeager {name} stability function other than "
                                          f"SplitStabilityFunction(LinearStable, {unstable_type.__name__})")
            ls = p.linear_stable_momentum if name == "momentum" else p.linear_stable_scalar
            ls[0], ls[1] = float(ψ.stable.coefficient), float(ψ.stable.maximum_stability_parameter)
        p.paulson_momentum[0] = float(sf.momentum.unstable.a)
        p.paulson_momentum[1] = float(sf.momentum.unstable.b)
        p.paulson_scalar = float(sf.temperature.unstable.a)
        return
    poly = None
    for n, c in enumerate(fluxes.coefficients()):
        if isinstance(c, bool) or not isinstance(c, (int, float, PolynomialNeutralDragCoefficient)):
            raise NotImplementedError(f"{_TRANSFER_KEYS[n]} transfer coefficient of type {type(c).__name__} "
                                      "(the HIP path takes numbers or a PolynomialNeutralDragCoefficient)")
        if isinstance(c, PolynomialNeutralDragCoefficient):
            if poly is not None and c.values() != poly.values():
                raise NotImplementedError("two different PolynomialNeutralDragCoefficient parameter sets")
            poly = c
            p.transfer_kind[n] = abi.NE_COEFFICIENT_POLYNOMIAL
            p.transfer_coefficient[n] = 0.0
        else:
            p.transfer_coefficient[n] = float(c)
    if poly is not None:
        _set_polynomial(p, poly)


def _lower_interface_temperature(p: abi.Params, it) -> None:
    p.interface_temperature = abi.NE_TEMPERATURE_BULK
    p.skin_diffusivity, p.skin_depth, p.skin_max_temperature_change = 0.0, 0.0, 5.0
    if isinstance(it, BulkTemperature):
        return
    if not isinstance(it, SkinTemperature):
        raise NotImplementedError(f"interface temperature {type(it).__name__}")
    F = it.internal_flux
    if not isinstance(F, DiffusiveFlux):
        raise NotImplementedError(f"skin-temperature internal flux {type(F).__name__} (sea-ice conductive fluxes "
                                  "have no counterpart on an ocean-only path)")
    if isinstance(F.κ, bool) or not isinstance(F.κ, (int, float)):
        raise NotImplementedError(f"skin diffusivity {type(F.κ).__name__}: a diffusivity assessed from the "
                                  "interior model needs a turbulence closure, which this path's ocean has not")
    p.interface_temperature = abi.NE_TEMPERATURE_SKIN
    p.skin_diffusivity = float(F.κ)
    p.skin_depth = float(F.δ)
    p.skin_max_temperature_change = float(it.max_ΔT)


def to_params(fluxes: SimilarityTheoryFluxes | CoefficientBasedFluxes | None = None,
              interfaces: ComponentInterfaces | None = None, *,
              surface_layer_height: float = 10.0, boundary_layer_height: float = 512.0,
              thermodynamics: AtmosphereThermodynamicsParameters | None = None,
              radiation: SurfaceRadiationProperties | None = SurfaceRadiationProperties(),
              stefan_boltzmann_constant: float = 5.670374419e-8,
              liquidus: LinearLiquidus | None = None,
              radiative_forcing: TwoColorRadiation | None = None) -> abi.Params:
    """Lower the reference formulation objects to an ``ne_params`` block."""
    interfaces = interfaces or ComponentInterfaces()
    fluxes = fluxes or interfaces.atmosphere_ocean_fluxes
    if isinstance(fluxes, CoefficientBasedFluxes):
        p = _lower_similarity(SimilarityTheoryFluxes())   # the unused similarity block keeps its defaults
        _lower_coefficients(p, fluxes)
    elif isinstance(fluxes, SimilarityTheoryFluxes):
        p = _lower_similarity(fluxes)
        _lower_coefficients(p, None)
    else:
        raise NotImplementedError(f"{type(fluxes).__name__} is not implemented by the HIP path "
                                  "(SimilarityTheoryFluxes or CoefficientBasedFluxes)")
    sc = fluxes.solver_stop_criteria
    if isinstance(sc, ConvergenceStopCriteria):
        p.stop_criteria = abi.NE_STOP_CONVERGENCE
        p.maxiter = int(sc.maxiter)
        p.tolerance = float(sc.tolerance)
    elif isinstance(sc, FixedIterations):
        p.stop_criteria = abi.NE_STOP_FIXED_ITERATIONS
        p.maxiter = int(sc.iterations)
        p.tolerance = 0.0
    else:
        raise NotImplementedError(f"solver stop criteria {type(sc).__name__}")
    _lower_interface_temperature(p, interfaces.atmosphere_ocean_interface_temperature)
    return _lower_interfaces(p, interfaces, surface_layer_height, boundary_layer_height, thermodynamics, radiation,
                             stefan_boltzmann_constant, liquidus, radiative_forcing)


def _lower_similarity(fluxes: SimilarityTheoryFluxes) -> abi.Params:
    p = abi.Params()
    p.von_karman_constant = fluxes.von_karman_constant
    p.turbulent_prandtl_number = fluxes.turbulent_prandtl_number
    sv = fluxes.subgrid_velocities
    conv, meso = None, None
    if isinstance(sv, SubgridVelocityCorrection):
        conv, meso = sv.convective, sv.mesoscale
    else:
        conv = sv
    if conv is None:
        p.convective_gustiness = 0
    elif isinstance(conv, ConvectiveGustiness):
        p.convective_gustiness = 1
        p.gustiness_parameter = conv.gustiness_parameter
        p.minimum_gustiness = conv.minimum_gustiness
    elif isinstance(conv, (int, float)):
        # a Number: vsgs² = v² (similarity:89); expressed as the mesoscale slot
        if meso is not None:
            raise NotImplementedError("two constant subgrid velocities")
        p.convective_gustiness = 0
        meso = float(conv)
    else:
        raise NotImplementedError(f"subgrid velocity {type(conv).__name__}")
    if meso is None:
        p.mesoscale_gustiness = 0
    elif isinstance(meso, (int, float)):
        p.mesoscale_gustiness = 1
        p.mesoscale_velocity = float(meso)
    else:
        raise NotImplementedError(f"mesoscale subgrid velocity {type(meso).__name__}")

    sf = fluxes.stability_functions
    if all(f is zero_stability_function for f in (sf.momentum, sf.temperature, sf.water_vapor)):
        p.stability_functions = abi.NE_STABILITY_NONE
    elif (isinstance(sf.momentum, EdsonMomentumStabilityFunction)
          and isinstance(sf.temperature, EdsonScalarStabilityFunction)
          and isinstance(sf.water_vapor, EdsonScalarStabilityFunction)
          and sf.temperature.values() == sf.water_vapor.values()):
        p.stability_functions = abi.NE_STABILITY_EDSON
        for n, x in enumerate(sf.momentum.values()):
            p.edson_momentum[n] = x
        for n, x in enumerate(sf.temperature.values()):
            p.edson_scalar[n] = x
    else:
        raise NotImplementedError("stability functions other than Edson (2013) or ψ ≡ 0")
    form = fluxes.similarity_form
    if isinstance(form, LogarithmicSimilarityProfile):
        p.similarity_form = abi.NE_SIMILARITY_LOG
    elif isinstance(form, COARELogarithmicSimilarityProfile):
        p.similarity_form = abi.NE_SIMILARITY_COARE_LOG
    else:
        raise NotImplementedError(f"similarity form {type(form).__name__}")

    mr, mval = _roughness(fluxes.momentum_roughness_length, "momentum")
    tr, tval = _roughness(fluxes.temperature_roughness_length, "scalar")
    qr, qval = _roughness(fluxes.water_vapor_roughness_length, "scalar")
    p.momentum_roughness, p.momentum_roughness_length = mr, mval
    p.temperature_roughness, p.temperature_roughness_length = tr, tval
    p.water_vapor_roughness, p.water_vapor_roughness_length = qr, qval
    m = fluxes.momentum_roughness_length
    m = m if isinstance(m, MomentumRoughnessLength) else MomentumRoughnessLength()
    s = fluxes.water_vapor_roughness_length
    s = s if isinstance(s, ScalarRoughnessLength) else ScalarRoughnessLength()
    t = fluxes.temperature_roughness_length
    if isinstance(t, ScalarRoughnessLength) and isinstance(fluxes.water_vapor_roughness_length, ScalarRoughnessLength):
        if (t.air_kinematic_viscosity, t.maximum_roughness_length, t.reynolds_number_scaling_function) != \
           (s.air_kinematic_viscosity, s.maximum_roughness_length, s.reynolds_number_scaling_function):
            raise NotImplementedError("different temperature and water-vapour ScalarRoughnessLength parameters")
    if m.air_kinematic_viscosity != s.air_kinematic_viscosity:
        raise NotImplementedError("different momentum and scalar air kinematic viscosities")
    if not isinstance(m.wave_formulation, (int, float)):
        raise NotImplementedError("wind-dependent wave formulation")
    p.roughness_gravitational_acceleration = m.gravitational_acceleration
    p.air_kinematic_viscosity = m.air_kinematic_viscosity
    p.charnock_parameter = m.wave_formulation
    p.smooth_wall_parameter = m.smooth_wall_parameter
    p.maximum_momentum_roughness = m.maximum_roughness_length
    p.reynolds_scaling_A = s.reynolds_number_scaling_function.A
    p.reynolds_scaling_b = s.reynolds_number_scaling_function.b
    p.maximum_scalar_roughness = s.maximum_roughness_length
    if not isinstance(fluxes.zero_plane_displacement, (int, float)):
        raise NotImplementedError("field-valued zero-plane displacement")
    p.zero_plane_displacement = float(fluxes.zero_plane_displacement)
    return p


def _lower_interfaces(p, interfaces, surface_layer_height, boundary_layer_height, thermodynamics, radiation,
                      stefan_boltzmann_constant, liquidus, radiative_forcing) -> abi.Params:
    th = thermodynamics or AtmosphereThermodynamicsParameters()
    liq = liquidus or LinearLiquidus()
    q = interfaces.atmosphere_ocean_interface_specific_humidity
    if isinstance(q, ImpureSaturationSpecificHumidity):
        if q.phase != "liquid":
            raise NotImplementedError("ice-phase interface humidity")
        p.specific_humidity = abi.NE_HUMIDITY_IMPURE_SATURATION
        p.water_mole_fraction = 1.0 if q.water_mole_fraction is None else float(q.water_mole_fraction)
    elif isinstance(q, FixedSpecificHumidity):
        p.specific_humidity = abi.NE_HUMIDITY_FIXED
        p.fixed_specific_humidity = float(q.q)
    else:
        raise NotImplementedError(f"interface humidity {type(q).__name__}")
    vd = interfaces.atmosphere_ocean_velocity_difference
    if isinstance(vd, RelativeVelocity):
        p.velocity_difference = abi.NE_VELOCITY_RELATIVE
    elif isinstance(vd, WindVelocity):
        p.velocity_difference = abi.NE_VELOCITY_WIND
    else:
        raise NotImplementedError(f"velocity difference {type(vd).__name__}")

    p.surface_layer_height = float(surface_layer_height)
    p.boundary_layer_height = float(boundary_layer_height)
    p.gravitational_acceleration = float(interfaces.gravitational_acceleration)
    for name in ("gas_constant", "dry_air_molar_mass", "water_molar_mass", "dry_air_adiabatic_exponent",
                 "water_vapor_heat_capacity", "liquid_water_heat_capacity", "reference_vaporization_enthalpy",
                 "reference_sublimation_enthalpy", "water_ice_heat_capacity",
                 "reference_temperature", "triple_point_temperature", "triple_point_pressure"):
        setattr(p, name, float(getattr(th, name)))
    p.ocean_reference_density = float(interfaces.ocean_reference_density)
    p.ocean_heat_capacity = float(interfaces.ocean_heat_capacity)
    p.freshwater_density = float(interfaces.freshwater_density)
    if radiation is None:
        p.radiation = 0
    else:
        p.radiation = 1
        p.ocean_albedo = float(radiation.albedo)
        p.ocean_emissivity = float(radiation.emissivity)
    p.stefan_boltzmann_constant = float(stefan_boltzmann_constant)
    p.freshwater_melting_temperature = float(liq.freshwater_melting_temperature)
    p.liquidus_slope = float(liq.slope)
    tcr = default_radiative_forcing()
    p.first_color_fraction = float(tcr.first_color_fraction)
    p.first_absorption_coefficient = float(tcr.first_absorption_coefficient)
    p.second_absorption_coefficient = float(tcr.second_absorption_coefficient)
    if radiative_forcing is None:
        p.radiative_forcing = abi.NE_RADIATIVE_FORCING_NONE
    elif isinstance(radiative_forcing, TwoColorRadiation):
        p.radiative_forcing = abi.NE_RADIATIVE_FORCING_TWO_COLOR
        p.first_color_fraction = float(radiative_forcing.first_color_fraction)
        p.first_absorption_coefficient = float(radiative_forcing.first_absorption_coefficient)
        p.second_absorption_coefficient = float(radiative_forcing.second_absorption_coefficient)
    else:
        raise NotImplementedError(f"radiative forcing {type(radiative_forcing).__name__}")
    return p


def default_params() -> abi.Params:
    """ne_params of the reference defaults (equal, field by field, to ne_params_default)."""
    return to_params()


def params_dict(p: abi.Params) -> dict:
    out = {}
    for name, _ in abi.Params._fields_:
        v = getattr(p, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out
