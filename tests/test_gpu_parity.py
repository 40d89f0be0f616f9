"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same seeded inputs.

Tolerances (FP64): every row is compared elementwise with relative error ≤ 1e-12 (and
≤ 1e-12 relative to the field's largest magnitude for elements near zero). The WENO5
tendency, a sum of face fluxes that may cancel, is compared per cell relative to
max(|G|, Σ|flux terms|/V). The land / immersed indexing (first active level per column)
must match bit for bit, and so must the rows whose GPU arithmetic is the oracle's
sequence of correctly rounded operations (interpolation, freeze clamp, net-flux
assembly, radiation).
"""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import assert_rel, assert_rel_envelope, assert_rel_terms, rel_err

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import make_case
from numericalearth_hip.model import OceanOnlyModel, Phases
from numericalearth_hip.params import (CoefficientBasedFluxes, ComponentInterfaces, DiffusiveFlux, FixedIterations,
                                       FixedSpecificHumidity, LargeYeagerTransferCoefficients,
                                       PolynomialNeutralDragCoefficient, SimilarityScales, SimilarityTheoryFluxes,
                                       SkinTemperature, to_params)

pytestmark = pytest.mark.gpu

RTOL = 1e-12
SMALL = (90, 40, 10)


def _np(t):
    return t.detach().cpu().numpy()


def _pair(shape=SMALL, seed=1, params=None, radiation=True, **kw):
    case = make_case(*shape, seed=seed, **kw)
    print(f"[case] shape={shape} seed={seed} land={case.meta['land_fraction']:.3f}", flush=True)
    p = params if params is not None else to_params()
    om = O.OracleModel(case, params=abi.Params.from_buffer_copy(p), radiation=radiation)
    gm = OceanOnlyModel(case, params=abi.Params.from_buffer_copy(p), radiation=radiation)
    return case, om, gm


def test_immersed_levels_bit_exact(require_gpu):
    case, om, gm = _pair()
    kmin_dev = _np(gm.grid.kmin)
    kmin_orc = np.zeros_like(kmin_dev)
    import ctypes as C
    O.lib().neo_immersed_levels(C.byref(om.grid), O.np_view(case.grid.bottom_height),
                                kmin_orc.ctypes.data_as(C.POINTER(C.c_int32)))
    assert np.array_equal(kmin_dev, kmin_orc)
    assert np.array_equal(kmin_dev, case.grid.kmin)
    assert 0 < np.mean(kmin_dev >= case.grid.spec.Nz) < 0.5   # some land, mostly ocean


def test_interpolation_bit_exact(require_gpu):
    case, om, gm = _pair(n_times=2)
    ph = Phases(gm)
    t = abi.TimeInterp(0, 1, 0.375)
    om._model.atmosphere.time = t
    om._model.radiation.time = t
    import ctypes as C
    L = O.lib()
    L.neo_interp_atmos(C.byref(om.grid), C.byref(om._model.atmosphere), O._dp(om.afi), O._dp(om.afj),
                       C.byref(om.exchange))
    L.neo_interp_radiation(C.byref(om.grid), C.byref(om._model.radiation), O._dp(om.rfi), O._dp(om.rfj),
                           O.np_view(om.sw), O.np_view(om.lw))
    ph.interp(t, t)
    torch.cuda.synchronize()
    for k in O.EXCHANGE_NAMES:
        assert np.array_equal(_np(ph.exchange[k]), om.exchange_arrays[k]), k
    assert np.array_equal(_np(ph.sw), om.sw)
    assert np.array_equal(_np(ph.lw), om.lw)
    assert np.array_equal(_np(gm.afi), om.afi) and np.array_equal(_np(gm.afj), om.afj)


def _solve_both(case, om, gm):
    ph = Phases(gm)
    om.interp()
    ph.interp()
    torch.cuda.synchronize()
    om.flux_solve()
    ph.flux_solve()
    torch.cuda.synchronize()
    return ph


# Amplification allowed over the oracle's own one-ulp response (see _flux_envelope). The
# elements beyond 1e-12 relative use at most 0.92 of that response in these tests (printed by
# _compare_fluxes); 4 leaves room for a different box's transcendental rounding, not for error.
ENVELOPE_FACTOR = 4.0


def _compare_fluxes(gm, om, rtol=RTOL, max_iteration_mismatch_frac=1e-4, envelope=None):
    """GPU interface fluxes against the oracle's: elementwise within rtol (and rtol of the field's
    largest value), or, with an envelope, within max(rtol·|b|, ENVELOPE_FACTOR·envelope)."""
    it_g = _np(gm.iterations)
    it_o = om.iterations
    mismatch = it_g != it_o
    frac = mismatch.mean()
    print(f"[flux] iteration mismatches: {mismatch.sum()} of {mismatch.size}", flush=True)
    assert frac <= max_iteration_mismatch_frac
    if mismatch.any():
        # A column whose stop test flipped (its drift met the tolerance within the last bits) stopped
        # one iteration before or after the oracle's: the two results are consecutive iterates, which
        # differ by that iteration's drift |Δu★| + |Δθ★| + |Δq★|, below the stop tolerance.
        assert np.all(np.abs(it_g[mismatch] - it_o[mismatch]) == 1), "a stop iteration differs by more than one"
        drift = sum(np.abs(_np(gm.fluxes[k])[mismatch] - om.flux_arrays[k][mismatch])
                    for k in ("friction_velocity", "temperature_scale", "water_vapor_scale"))
        assert np.all(drift <= om.params.tolerance * (1.0 + 1e-6)), f"flipped columns differ by {drift.max():.3e}"
    ok = ~mismatch
    worst = {}
    for k in O.FLUX_NAMES:
        a = _np(gm.fluxes[k])[ok]
        b = om.flux_arrays[k][ok]
        if envelope is None:
            worst[k] = assert_rel(a, b, rtol, k)
        else:
            worst[k] = assert_rel_envelope(a, b, ENVELOPE_FACTOR * envelope[k][ok], rtol, k)
    print("[flux] worst relative errors:", {k: f"{v[0]:.2e}" for k, v in worst.items()}, flush=True)
    if envelope is not None:
        print("[flux] largest use of the envelope (|Δ| / one-ulp response, beyond 1e-12 relative):",
              {k: f"{v[2] * ENVELOPE_FACTOR:.2f}" for k, v in worst.items()}, flush=True)


def _oracle_on_state(om, T, S, time=None, nudge=None):
    """A fresh oracle model with om's case and parameters holding the state (T, S): atmosphere and
    radiation interpolated at `time` (om's by default), one input moved by one ulp if `nudge` =
    (name, ±1) names one (an interpolated atmosphere field, or "sst", the top level of T), and the
    fluxes solved."""
    o = O.OracleModel(om.case, params=abi.Params.from_buffer_copy(om.params), radiation=bool(om.params.radiation),
                      time_indexing=om.time_indexing, time_period=om.time_period)
    o.T[...] = T
    o.S[...] = S
    if nudge and nudge[0] == "sst":
        o.T[-1] = np.nextafter(o.T[-1], np.inf * nudge[1])
    o.time = om.time if time is None else time
    o.interp()
    if nudge and nudge[0] != "sst":
        x = o.exchange_arrays[nudge[0]]
        x[...] = np.nextafter(x, np.inf * nudge[1])
    o.flux_solve()
    return o


def _flux_envelope(om, T=None, S=None):
    """Per-element response of the oracle's column solve to a one-ulp change of each of its main inputs
    (air temperature, humidity, wind, SST), up and down. Where θₐ − Tₛ nearly cancels (a column with
    almost no air-sea temperature difference), or the skin balance's denominator κΔT − Ωᵀδ does, or
    an iteration stops short of its fixed point, any one-ulp difference inside the column (a
    transcendental's last bit) is amplified by the same factor; this measures it."""
    T = om.T if T is None else T
    S = om.S if S is None else S
    base = _oracle_on_state(om, T, S)
    env = {k: np.zeros_like(v) for k, v in base.flux_arrays.items()}
    for name in ("T", "q", "u", "sst"):
        for sign in (1, -1):
            o = _oracle_on_state(om, T, S, nudge=(name, sign))
            for k in env:
                env[k] = np.maximum(env[k], np.abs(o.flux_arrays[k] - base.flux_arrays[k]))
    # θ★ and the sensible heat are proportional to Δθ = θₐ − Tₛ: a one-ulp change of a computed
    # interface temperature Tₛ (the skin balance's result) changes them by |Tₛ/Δθ| ulps, a factor
    # the input nudges above need not excite
    p = base.params
    cp_d = p.gas_constant / p.dry_air_molar_mass / p.dry_air_adiabatic_exponent
    qa, Ta = base.exchange_arrays["q"], base.exchange_arrays["T"]
    theta = Ta + p.gravitational_acceleration * p.surface_layer_height / (cp_d + (p.water_vapor_heat_capacity - cp_d) * qa)
    Ts = base.flux_arrays["interface_temperature"] + 273.15
    with np.errstate(divide="ignore", invalid="ignore"):
        cond = np.where(theta != Ts, np.abs(Ts) / np.abs(theta - Ts), 0.0)
    for k in ("temperature_scale", "sensible_heat"):
        env[k] = np.maximum(env[k], np.abs(base.flux_arrays[k]) * cond * 2.0 ** -52)
    return env


def _compare_fluxes_on_gpu_state(gm, om, envelope=False):
    """The GPU's fluxes after a step against the oracle's flux solve of the GPU's own stepped state:
    the two steps' states differ in the last bits, and a column with almost no air-sea temperature
    difference turns that into a large relative difference of its sensible heat, which is not a
    difference of the flux solve."""
    T, S = _np(gm.T).astype(np.float64), _np(gm.S).astype(np.float64)
    ref = _oracle_on_state(om, T, S)
    _compare_fluxes(gm, ref, envelope=_flux_envelope(om, T, S) if envelope else None)


def test_flux_solve_matches_oracle(require_gpu):
    case, om, gm = _pair()
    _solve_both(case, om, gm)
    _compare_fluxes(gm, om)


def test_flux_solve_column_order_changes_nothing(require_gpu):
    """The solve takes each block's columns in order of the iteration counts in the iterations
    view (so that wavefronts hold columns of like cost); without the view it keeps the natural
    order. Scrambled counts, the real ones and none must give the same fluxes bit for bit."""
    case, om, gm = _pair(seed=5)
    ph = _solve_both(case, om, gm)
    ref = {k: v.clone() for k, v in gm.fluxes.items()}
    iters = gm.iterations.clone()
    for trial in ("scrambled", "none"):
        if trial == "scrambled":
            gm.iterations.copy_(torch.randint(0, 40, gm.iterations.shape, device=gm.iterations.device,
                                              dtype=gm.iterations.dtype, generator=torch.Generator(
                                                  device=gm.iterations.device).manual_seed(3)))
            fl = gm.desc.fluxes
        else:
            fl = abi.InterfaceFluxes.from_buffer_copy(gm.desc.fluxes)
            fl.iterations = abi.View(None, 0, 0, 0)
        for v in gm.fluxes.values():
            v.zero_()
        from numericalearth_hip._lib import check
        check(gm.lib.ne_ao_flux_solve_f64(gm.ctx.handle, C.byref(gm.grid.struct), C.byref(gm.params),
                                          C.byref(ph._ex()), C.byref(gm.desc.ocean), C.byref(fl), None),
              "ne_ao_flux_solve_f64")
        torch.cuda.synchronize()
        for k in ref:
            assert torch.equal(gm.fluxes[k], ref[k]), f"{k} changed with {trial} column order"
    assert torch.equal(gm.iterations, iters)   # the scrambled counts were overwritten by the real ones


def test_flux_solve_across_the_stability_range(require_gpu):
    """The solve over winds of 0.5–30 m/s and air 25 K colder to 4 K warmer than the sea, with
    relative humidities of 30–100 % (the same exchange state given to both): ζ = h/L then runs
    from strongly unstable (free convection, where the stability functions' arctangents, cube
    roots and logarithms take arguments far from 1) to stable, and every column must match the
    oracle as at the workload's own conditions."""
    case, om, gm = _pair(seed=9)
    ph = Phases(gm)
    om.interp()
    ph.interp()
    torch.cuda.synchronize()
    Nx, Ny, _ = case.shape
    U = 10.0 ** np.linspace(np.log10(0.5), np.log10(30.0), Nx)[None, :] * np.ones((Ny, 1))
    dT = np.linspace(-25.0, 4.0, Ny)[:, None] * np.ones((1, Nx))
    rh = 0.3 + 0.7 * ((np.arange(Nx)[None, :] * 7 + np.arange(Ny)[:, None] * 3) % 11) / 10.0
    sst_k = om.T[-1] + 273.15
    Ta = sst_k + dT
    es = 611.2 * np.exp(17.67 * (Ta - 273.15) / (Ta - 29.65))        # Bolton's saturation pressure
    qa = rh * 0.622 * es / om.exchange_arrays["p"]
    fields = {"u": U * 0.8, "v": U * 0.6, "T": Ta, "q": qa}
    for k, v in fields.items():
        om.exchange_arrays[k][...] = v
        ph.exchange[k].copy_(torch.from_numpy(np.ascontiguousarray(v)))
    om.flux_solve()
    ph.flux_solve()
    torch.cuda.synchronize()
    us, ts, qs = (om.flux_arrays[k] for k in ("friction_velocity", "temperature_scale", "water_vapor_scale"))
    wet = om.iterations > 0
    with np.errstate(divide="ignore", invalid="ignore"):   # land columns: u★ = 0
        zeta = 10.0 * 0.4 * 9.80665 * (ts + 0.608 * Ta * qs) / (Ta * us ** 2)
    print(f"[flux] ζ from {zeta[wet].min():.2f} to {zeta[wet].max():.2f}; iterations up to {om.iterations.max()}",
          flush=True)
    assert zeta[wet].min() < -5.0 and zeta[wet].max() > 0.1
    _compare_fluxes(gm, om, max_iteration_mismatch_frac=0.01)


def test_flux_solve_fixed_iterations(require_gpu):
    p = to_params(SimilarityTheoryFluxes(solver_stop_criteria=FixedIterations(12)))
    case, om, gm = _pair(params=p, seed=3)
    _solve_both(case, om, gm)
    assert np.all(_np(gm.iterations) == 12)
    _compare_fluxes(gm, om, max_iteration_mismatch_frac=0.0)


def test_flux_solve_neutral_and_fixed_humidity(require_gpu):
    fl = SimilarityTheoryFluxes(momentum_roughness_length=1e-4, temperature_roughness_length=1e-4,
                                water_vapor_roughness_length=1e-4, subgrid_velocities=None, stability_functions=None)
    p = to_params(fl, ComponentInterfaces(atmosphere_ocean_interface_specific_humidity=FixedSpecificHumidity(0.003)))
    case, om, gm = _pair(params=p, seed=4)
    _solve_both(case, om, gm)
    _compare_fluxes(gm, om, max_iteration_mismatch_frac=0.0)


COEFFICIENT_FORMS = {
    "constant": CoefficientBasedFluxes(transfer_coefficients=(1.3e-3, 1.1e-3, 1.2e-3)),
    "polynomial": CoefficientBasedFluxes(
        transfer_coefficients=SimilarityScales(PolynomialNeutralDragCoefficient(), 1e-3, 1.2e-3)),
    "large_yeager": CoefficientBasedFluxes(transfer_coefficients=LargeYeagerTransferCoefficients()),
    "large_yeager_fixed5": CoefficientBasedFluxes(transfer_coefficients=LargeYeagerTransferCoefficients(),
                                                  solver_stop_criteria=FixedIterations(5)),
}


@pytest.mark.parametrize("form", sorted(COEFFICIENT_FORMS))
def test_flux_solve_coefficient_based(require_gpu, form):
    p = to_params(COEFFICIENT_FORMS[form])
    case, om, gm = _pair(params=p, seed=21)
    _solve_both(case, om, gm)
    # Large-Yeager columns that do not converge within maxiter amplify last-bit differences
    _compare_fluxes(gm, om, envelope=_flux_envelope(om) if form.startswith("large_yeager") else None)


SKIN = ComponentInterfaces(atmosphere_ocean_interface_temperature=SkinTemperature(DiffusiveFlux(1e-4, 1.0)))


@pytest.mark.parametrize("form", ["similarity", "large_yeager"])
@pytest.mark.parametrize("radiation", [True, False])
def test_flux_solve_skin_temperature(require_gpu, form, radiation):
    fl = SimilarityTheoryFluxes() if form == "similarity" else COEFFICIENT_FORMS["large_yeager"]
    p = to_params(fl, SKIN)
    case, om, gm = _pair(params=p, seed=22, radiation=radiation)
    _solve_both(case, om, gm)
    _compare_fluxes(gm, om, envelope=_flux_envelope(om))
    Ts, To = _np(gm.fluxes["interface_temperature"]), case.T[-1]
    wet = case.grid.kmin < case.grid.spec.Nz
    assert np.any(Ts[wet] != To[wet]) and np.all(np.abs(Ts[wet] - To[wet]) <= 5.0)


@pytest.mark.parametrize("form", ["large_yeager_skin", "constant"])
def test_model_coefficient_steps_match_oracle(require_gpu, form):
    """The fused step (solve with the radiation interpolated in K1 for the skin balance) over 3 steps."""
    p = (to_params(COEFFICIENT_FORMS["large_yeager"], SKIN) if form == "large_yeager_skin"
         else to_params(COEFFICIENT_FORMS["constant"]))
    case, om, gm = _pair(seed=23, params=p)
    skin = form == "large_yeager_skin"
    om.reconcile()
    torch.cuda.synchronize()
    _compare_fluxes(gm, om, envelope=_flux_envelope(om, case.T, case.S) if skin else None)
    for step in range(3):
        om.time_step(1200.0)
        gm.time_step(1200.0)
        torch.cuda.synchronize()
        assert_rel(_np(gm.T), om.T, RTOL, "T")
        assert_rel(_np(gm.S), om.S, RTOL, "S")
        _compare_fluxes_on_gpu_state(gm, om, envelope=skin)
        if skin:
            # the net fluxes assembled from the GPU's own interface fluxes (whose rounding-level
            # differences the envelope above bounds) on the GPU's state
            ref = _oracle_on_state(om, _np(gm.T), _np(gm.S))
            for k in O.FLUX_NAMES:
                ref.flux_arrays[k][...] = _np(gm.fluxes[k])
            ref.freeze(1200.0)
            ref.assemble()
            ref.radiation_apply()
            net_ref = ref
        else:
            net_ref = om
        scales = _net_term_scales(case, net_ref)
        for k in O.NET_NAMES:
            assert_rel_terms(_np(gm.net[k]), net_ref.net_arrays[k], scales.get(k, 0.0), RTOL, k)


def test_assembly_radiation_freeze_bit_exact(require_gpu):
    case, om, gm = _pair(seed=5)
    # force some freezing cells
    cold = case.T.copy()
    cold[-1, ::7, ::5] = -5.0
    om.T[...] = cold
    gm.T.copy_(torch.from_numpy(cold))
    ph = _solve_both(case, om, gm)
    om.freeze(1200.0)
    ph.freeze(1200.0)
    torch.cuda.synchronize()
    assert np.array_equal(_np(gm.T), om.T)
    assert np.array_equal(_np(gm.frazil_heat), om.frazil)
    assert (om.frazil < 0).sum() > 0
    # feed identical interface fluxes to both assemblies
    for k in O.FLUX_NAMES:
        gm.fluxes[k].copy_(torch.from_numpy(om.flux_arrays[k]))
    om.assemble()
    ph.assemble()
    om.radiation_apply()
    ph.radiation()
    torch.cuda.synchronize()
    for k in O.NET_NAMES:
        assert np.array_equal(_np(gm.net[k]), om.net_arrays[k]), k
    for k in O.RAD_NAMES:
        assert np.array_equal(_np(gm.radiative[k]), om.rad_arrays[k]), k


def _sync_net(gm, om):
    for k in O.NET_NAMES:
        gm.net[k].copy_(torch.from_numpy(om.net_arrays[k]))


TENDENCY_SHAPES = [(SMALL, 0), ((37, 23, 7), 0), ((48, 30, 8), 1), ((130, 17, 12), 0), ((5, 4, 3), 0)]


@pytest.mark.parametrize("shape,topology_x", TENDENCY_SHAPES)
def test_tendency_matches_oracle(require_gpu, shape, topology_x):
    """The WENO5 tendency (order reduction at immersed / bounded faces, top flux BC). The GPU kernel
    uses fused multiply-adds and a one-division WENO-Z normalisation, so parity is to 1e-12 relative
    to max(|G|, Σ|flux terms|/V) per cell: G is a sum of face fluxes that can cancel."""
    case, om, gm = _pair(shape, seed=6, topology_x=topology_x, longitude=(0.0, 360.0) if topology_x == 0 else (20.0, 80.0))
    om.reconcile()
    _sync_net(gm, om)
    GT_o, GS_o, sT, sS = om.tendency(scales=True)
    GT_g, GS_g = Phases(gm).tendency()
    torch.cuda.synchronize()
    eT = assert_rel_terms(_np(GT_g), GT_o, sT, RTOL, "G_T")
    eS = assert_rel_terms(_np(GS_g), GS_o, sS, RTOL, "G_S")
    fT = rel_err(_np(GT_g), GT_o)[1]
    print(f"[tendency] {shape}: G_T {eT:.2e} G_S {eS:.2e} (field-scaled G_T {fT:.2e})", flush=True)
    active = np.arange(case.grid.spec.Nz)[:, None, None] >= case.grid.kmin[None]
    assert np.all(_np(GT_g)[~active] == 0) and np.all(_np(GS_g)[~active] == 0)
    assert np.abs(GT_o).max() > 0


def test_full_size_snapshot_matches_oracle(require_gpu):
    """The quarter-degree workload (1440×560×10): the construction-time flux solve, the WENO5
    tendency of the resulting state and one coupled step match the oracle to 1e-12 (fluxes
    elementwise, the tendency relative to its terms, the stepped state elementwise)."""
    shape = (1440, 560, 10)
    case = make_case(*shape, seed=0)
    print(f"[full] land={case.meta['land_fraction']:.4f}", flush=True)
    om = O.OracleModel(case).reconcile()
    gm = OceanOnlyModel(case)
    torch.cuda.synchronize()
    _compare_fluxes(gm, om)
    GT_o, GS_o, sT, sS = om.tendency(scales=True)
    GT_g, GS_g = Phases(gm).tendency()
    torch.cuda.synchronize()
    eT = assert_rel_terms(_np(GT_g), GT_o, sT, RTOL, "G_T")
    eS = assert_rel_terms(_np(GS_g), GS_o, sS, RTOL, "G_S")
    print(f"[full] tendency G_T {eT:.2e} G_S {eS:.2e}", flush=True)
    om.time_step(1200.0)
    gm.time_step(1200.0)
    torch.cuda.synchronize()
    assert_rel(_np(gm.T), om.T, RTOL, "T")
    assert_rel(_np(gm.S), om.S, RTOL, "S")
    _compare_fluxes_on_gpu_state(gm, om)
    scales = _net_term_scales(case, om)
    for k in O.NET_NAMES:
        assert_rel_terms(_np(gm.net[k]), om.net_arrays[k], scales.get(k, 0.0), RTOL, k)


def test_tendency_two_color_matches_oracle(require_gpu):
    """WENO5 tendency with the TwoColorRadiation forcing of T (20 m levels, so the shortwave reaches
    several levels and the bottom of shallow columns)."""
    from numericalearth_hip.params import TwoColorRadiation
    p = to_params(radiative_forcing=TwoColorRadiation())
    case, om, gm = _pair(SMALL, seed=14, params=p, z=(-200.0, 0.0))
    om.reconcile()
    _sync_net(gm, om)
    gm.shortwave_surface_flux.copy_(torch.from_numpy(om.shortwave_surface_flux))
    assert np.abs(om.shortwave_surface_flux).max() > 0
    GT_o, GS_o, sT, sS = om.tendency(scales=True)
    GT_g, GS_g = Phases(gm).tendency()
    torch.cuda.synchronize()
    eT = assert_rel_terms(_np(GT_g), GT_o, sT, RTOL, "G_T")
    eS = assert_rel_terms(_np(GS_g), GS_o, sS, RTOL, "G_S")
    print(f"[two-colour] tendency G_T {eT:.2e} G_S {eS:.2e}", flush=True)


@pytest.mark.parametrize("timestepper", ["SplitRungeKutta3", "QuasiAdamsBashforth2"])
def test_model_two_color_steps_match_oracle(require_gpu, timestepper):
    from numericalearth_hip.params import TwoColorRadiation
    p = to_params(radiative_forcing=TwoColorRadiation())
    case = make_case(*SMALL, seed=15, z=(-200.0, 0.0))
    print(f"[case] shape={SMALL} seed=15 two-colour {timestepper}", flush=True)
    om = O.OracleModel(case, params=abi.Params.from_buffer_copy(p), timestepper=timestepper)
    gm = OceanOnlyModel(case, params=abi.Params.from_buffer_copy(p), timestepper=timestepper)
    om.reconcile()
    torch.cuda.synchronize()
    assert_rel(_np(gm.shortwave_surface_flux), om.shortwave_surface_flux, RTOL, "J0")
    for step in range(3):
        om.time_step(1200.0)
        gm.time_step(1200.0)
        torch.cuda.synchronize()
        assert_rel(_np(gm.T), om.T, RTOL, "T")
        assert_rel(_np(gm.S), om.S, RTOL, "S")
        _compare_fluxes_on_gpu_state(gm, om)
        scales = _net_term_scales(case, om)
        for k in O.NET_NAMES:
            assert_rel_terms(_np(gm.net[k]), om.net_arrays[k], scales.get(k, 0.0), RTOL, k)
        assert_rel(_np(gm.shortwave_surface_flux), om.shortwave_surface_flux, RTOL, "J0")


def test_rk_stage_matches_oracle(require_gpu):
    case, om, gm = _pair(seed=7)
    om.reconcile()
    _sync_net(gm, om)
    T1 = torch.zeros_like(gm.T)
    S1 = torch.zeros_like(gm.S)
    Phases(gm).rk_stage(gm.T, gm.S, gm.T, gm.S, T1, S1, 400.0)
    torch.cuda.synchronize()
    GT, GS = om.tendency()
    active = np.arange(case.grid.spec.Nz)[:, None, None] >= case.grid.kmin[None]
    Texp = np.where(active, case.T + 400.0 * GT, 0.0)
    Sexp = np.where(active, case.S + 400.0 * GS, 0.0)
    assert_rel(np.where(active, _np(T1), 0.0), Texp, RTOL, "T1")
    assert_rel(np.where(active, _np(S1), 0.0), Sexp, RTOL, "S1")
    # inactive cells of the output are not written
    assert np.all(_np(T1)[~active] == 0)


@pytest.mark.parametrize("precision", ["fp64"])
def test_model_steps_match_oracle(require_gpu, precision):
    case, om, gm = _pair(seed=11)
    om.reconcile()
    torch.cuda.synchronize()
    _compare_fluxes(gm, om)
    for step in range(5):
        om.time_step(1200.0)
        gm.time_step(1200.0)
        torch.cuda.synchronize()
        print(f"[model] step {step + 1}", flush=True)
        assert_rel(_np(gm.T), om.T, RTOL, "T")
        assert_rel(_np(gm.S), om.S, RTOL, "S")
        _compare_fluxes_on_gpu_state(gm, om)
        scales = _net_term_scales(case, om)
        for k in O.NET_NAMES:
            assert_rel_terms(_np(gm.net[k]), om.net_arrays[k], scales.get(k, 0.0), RTOL, k)
    assert gm.clock == (6000.0, 5)


def test_model_ab2_steps_match_oracle(require_gpu):
    """The reference's alternative timestepper, timestepper = :QuasiAdamsBashforth2
    (ocean_simulation.jl:323): one tendency and one update per step, forward Euler on the first
    step and on the step whose Δt changes (1200 s, then 900 s from the eighth step). T and S and
    the fluxes after each step match the oracle's QAB2 model to 1e-12, and differ from the RK3
    step's (the option is not ignored)."""
    case = make_case(*SMALL, seed=23)
    print(f"[case] shape={SMALL} seed=23 QAB2", flush=True)
    om = O.OracleModel(case, timestepper="QuasiAdamsBashforth2").reconcile()
    gm = OceanOnlyModel(case, timestepper="QuasiAdamsBashforth2")
    rk = OceanOnlyModel(case)
    torch.cuda.synchronize()
    for step in range(12):
        dt = 1200.0 if step < 7 else 900.0
        om.time_step(dt)
        gm.time_step(dt)
        rk.time_step(dt)
        torch.cuda.synchronize()
        print(f"[ab2] step {step + 1} dt {dt}", flush=True)
        assert_rel(_np(gm.T), om.T, RTOL, "T")
        assert_rel(_np(gm.S), om.S, RTOL, "S")
        _compare_fluxes_on_gpu_state(gm, om)
    assert gm.clock == (7 * 1200.0 + 5 * 900.0, 12)
    assert rel_err(_np(gm.T), _np(rk.T))[1] > 1e-9
    for m in (gm, rk):
        m.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_freezing_steps_clamp_every_flagged_column(require_gpu, precision):
    """Cold, windy air over water 2 mK above its freezing point on 10 m levels: each step cools
    surface cells below T_m and advection carries saltier (colder-freezing) water into fresher cells.
    The step's last RK stage flags those columns and the freeze clamp after it visits only them; every
    active cell must end at or above T_m, and in FP64 T, S and the frazil heat must equal the oracle's
    (which clamps every column)."""
    case = make_case(*SMALL, seed=21, z=(-100.0, 0.0))
    p = to_params()
    melt0, slope = p.freshwater_melting_temperature, p.liquidus_slope
    case.T[...] = melt0 - slope * case.S + 0.002
    case.atmos["T"][...] = 250.0
    case.atmos["u"][...] = 15.0
    print(f"[freeze] precision={precision} melt0={melt0} slope={slope}", flush=True)
    gm = OceanOnlyModel(case, params=abi.Params.from_buffer_copy(p), precision=precision)
    om = O.OracleModel(case, params=abi.Params.from_buffer_copy(p)) if precision == "fp64" else None
    if om is not None:
        om.reconcile()
    active = np.arange(case.grid.spec.Nz)[:, None, None] >= case.grid.kmin[None]
    frozen = 0
    for step in range(3):
        gm.time_step(1200.0)
        torch.cuda.synchronize()
        T, S = _np(gm.T).astype(np.float64), _np(gm.S).astype(np.float64)
        # the clamp stores T_m (formed in FP64) rounded to the storage type
        Tm = (melt0 - slope * S).astype(_np(gm.T).dtype).astype(np.float64)
        assert np.all(T[active] >= Tm[active]), f"a cell below freezing after step {step + 1}"
        fz = _np(gm.frazil_heat)
        frozen += int(np.count_nonzero(fz))
        if om is not None:
            om.time_step(1200.0)
            assert_rel(T, om.T, RTOL, "T")
            assert_rel(S, om.S, RTOL, "S")
            # The frazil heat is a column sum of ρc(T_m − T)Δz/Δt over cells a few mK below
            # T_m: each term is the difference of two values of size |T_m|, so it is compared
            # against the size of those terms (a 1-ulp difference in T, which the T comparison
            # allows, moves it by ulp(T)/(T_m − T) relative to itself).
            rho_c = p.ocean_reference_density * p.ocean_heat_capacity
            dz = case.grid.dz[:, None, None]
            terms = rho_c / 1200.0 * np.sum(np.where(active, 2.0 * np.abs(Tm) * dz, 0.0), axis=0)
            assert_rel_terms(fz, om.frazil, terms, RTOL, "frazil heat")
            assert np.array_equal(fz == 0.0, om.frazil == 0.0)
    assert frozen > 100, "the case froze too few columns to exercise the clamp"
    gm.close()


def _net_term_scales(case, om):
    """Magnitude of the terms summed into each net flux: τ is a face average of neighbouring
    ρτ/ρₒ (which may cancel), Jᵀ a sum of the turbulent and radiative heat fluxes over ρₒcₒ."""
    f, r = om.flux_arrays, om.rad_arrays
    rho, c = om.params.ocean_reference_density, om.params.ocean_heat_capacity
    spec = case.grid.spec
    rtx, rty = np.abs(f["x_momentum"]), np.abs(f["y_momentum"])
    west, south = np.roll(rtx, 1, axis=1), np.roll(rty, 1, axis=0)
    if spec.topology_x != abi.NE_PERIODIC:
        west[:, 0] = 0.0
    if spec.topology_y != abi.NE_PERIODIC:
        south[0, :] = 0.0
    heat = (np.abs(f["sensible_heat"]) + np.abs(f["latent_heat"]) + np.abs(r["upwelling_longwave"]) +
            np.abs(r["downwelling_longwave"]) + np.abs(r["downwelling_shortwave"])) / (rho * c)
    return {"tau_x": (rtx + west) / 2.0 / rho, "tau_y": (rty + south) / 2.0 / rho, "J_T": heat}


def test_budget_closure_on_device(require_gpu):
    case, om, gm = _pair(seed=12)
    rho_c = gm.params.ocean_reference_density * gm.params.ocean_heat_capacity
    H0 = gm.heat_content()
    Ho = om.budget(om.T, rho_c)
    assert abs(H0 - (Ho[0] + Ho[1])) <= 1e-15 * abs(H0)
    JT = gm.net["J_T"].clone()
    gm.time_step(1200.0)
    torch.cuda.synchronize()
    H1 = gm.heat_content()
    Az = case.grid.Az
    expected = -1200.0 * rho_c * float(np.sum(_np(JT) * Az[:, None]))
    # Flux-form advection telescopes, so the heat content changes by the surface flux alone
    # (the freeze clamp is inert at these temperatures). The only error left is the rounding
    # of the stored state: at most half an ulp of T per cell, weighted by ρc·V.
    T1 = _np(gm.T)
    active = np.arange(case.grid.spec.Nz)[:, None, None] >= case.grid.kmin[None]
    V = case.grid.Az[None, :, None] * case.grid.dz[:, None, None]
    bound = rho_c * float(np.sum(np.where(active, V * np.spacing(np.abs(T1)), 0.0)))
    err = abs((H1 - H0) - expected)
    print(f"[budget] ΔH={H1 - H0:.6e} expected={expected:.6e} err={err:.3e} bound={bound:.3e}", flush=True)
    assert err <= bound


def test_mixed_form_1000_steps_vs_fp64_oracle(require_gpu):
    """Mixed form (FP32 WENO5, interpolation inputs and state with stochastically rounded
    updates; FP32-then-FP64 flux iteration; FP64 fluxes and budgets) against the FP64 oracle
    after 1000 steps of 20 min on the 90×40×10 workload: SST within 1e-5 relative (largest
    |ΔSST| over the largest |SST|, and in the L2 norm), surface heat fluxes within 1e-3 of
    their scale, and heat content within 1e-3 of the heat gained over the run."""
    case = make_case(*SMALL, seed=21)
    om = O.OracleModel(case).reconcile()
    g32 = OceanOnlyModel(case, precision="mixed")
    rho_c = g32.params.ocean_reference_density * g32.params.ocean_heat_capacity
    H0 = sum(om.budget(om.T, rho_c))
    for n in range(1000):
        om.time_step(1200.0)
        g32.time_step(1200.0)
        if n % 200 == 199:
            torch.cuda.synchronize()
            print(f"[mixed] step {n + 1}", flush=True)
    torch.cuda.synchronize()
    wet = case.grid.kmin < case.grid.spec.Nz
    sst = _np(g32.T)[-1].astype(np.float64)
    ref = om.T[-1][wet]
    diff = np.abs(sst[wet] - ref)
    sst_err = float(diff.max() / np.max(np.abs(ref)))
    sst_l2 = float(np.sqrt(np.sum(diff ** 2) / np.sum(ref ** 2)))
    print(f"[mixed] SST max |Δ| {diff.max():.2e} K, rms {np.sqrt(np.mean(diff ** 2)):.2e} K", flush=True)
    qt = _np(g32.fluxes["sensible_heat"])
    qv = _np(g32.fluxes["latent_heat"])
    qt_err = float(np.max(np.abs(qt - om.flux_arrays["sensible_heat"])) / np.max(np.abs(om.flux_arrays["sensible_heat"])))
    qv_err = float(np.max(np.abs(qv - om.flux_arrays["latent_heat"])) / np.max(np.abs(om.flux_arrays["latent_heat"])))
    H32 = g32.budget(g32.T, rho_c)
    Ho = sum(om.budget(om.T, rho_c))
    h_err = abs(H32 - Ho) / abs(Ho - H0)
    print(f"[mixed] 1000 steps: SST {sst_err:.2e} (L2 {sst_l2:.2e}), sensible {qt_err:.2e}, latent {qv_err:.2e}, "
          f"heat {h_err:.2e} of the change", flush=True)
    assert sst_err <= 1e-5 and sst_l2 <= 1e-5
    assert qt_err <= 1e-3 and qv_err <= 1e-3
    assert h_err <= 1e-3


def test_mixed_form_budget_closure(require_gpu):
    """The mixed form conserves heat to FP64 round-off (test_tracer_budget.jl:66-76 is the
    reference's closure check): over 5 steps the heat content of the FP32 state plus ρ₀c₀ times its
    rounding residual changes by the surface heat flux alone, within the FP64 rounding of the
    stage update (half an FP64 ulp of T per cell and step, the FP64 model's bound). The FP32 state's
    own heat content misses by the FP32 rounding, which the residual accounts for; its budget sum
    equals the oracle's FP64 compensated sum of the same FP32 field."""
    case = make_case(*SMALL, seed=22)
    g32 = OceanOnlyModel(case, precision="mixed")
    om = O.OracleModel(case)
    rho_c = g32.params.ocean_reference_density * g32.params.ocean_heat_capacity
    torch.cuda.synchronize()
    H0 = g32.heat_content()
    Ho = om.budget(_np(g32.T).astype(np.float64), rho_c)
    assert abs(H0 - (Ho[0] + Ho[1])) <= 1e-15 * abs(H0)
    assert g32.rounding_residual() == (0.0, 0.0)
    C0 = g32.conserved_heat()
    Az = case.grid.Az
    active = np.arange(case.grid.spec.Nz)[:, None, None] >= case.grid.kmin[None]
    V = case.grid.Az[None, :, None] * case.grid.dz[:, None, None]
    expected, bound64, bound32 = 0.0, 0.0, 0.0
    for _ in range(5):
        JT = _np(g32.net["J_T"])
        expected += -1200.0 * rho_c * float(np.sum(JT * Az[:, None]))
        g32.time_step(1200.0)
        torch.cuda.synchronize()
        T1 = np.abs(_np(g32.T)).astype(np.float64)
        bound64 += rho_c * float(np.sum(np.where(active, V * np.spacing(T1), 0.0)))
        bound32 += rho_c * float(np.sum(np.where(active, V * np.spacing(T1.astype(np.float32)), 0.0)))
    H1 = g32.heat_content()
    C1 = g32.conserved_heat()
    err = abs((C1 - C0) - expected)
    err32 = abs((H1 - H0) - expected)
    print(f"[mixed budget] ΔH={H1 - H0:.6e} expected={expected:.6e} state alone misses by {err32:.3e} "
          f"(FP32 bound {bound32:.3e}); with the residual {err:.3e} (FP64 bound {bound64:.3e}, "
          f"{err / abs(expected):.2e} of the change)", flush=True)
    assert err32 <= bound32
    assert err <= bound64
    assert err32 > bound64          # the FP32 state alone does not close to FP64 round-off


def test_mixed_precision_model_tracks_fp64(require_gpu):
    case = make_case(*SMALL, seed=13)
    g64 = OceanOnlyModel(case, precision="fp64")
    g32 = OceanOnlyModel(case, precision="mixed")
    for _ in range(10):
        g64.time_step(1200.0)
        g32.time_step(1200.0)
    torch.cuda.synchronize()
    sst64 = _np(g64.T)[-1]
    sst32 = _np(g32.T)[-1].astype(np.float64)
    wet = case.grid.kmin < case.grid.spec.Nz
    err = np.max(np.abs(sst32 - sst64)[wet] / np.abs(sst64[wet]))
    print(f"[mixed] SST relative error after 10 steps: {err:.2e}", flush=True)
    assert err < 1e-5


def test_ab2_update_entry_point_bit_exact(require_gpu):
    """ne_tracer_ab2_update_f64, the per-phase Quasi-Adams-Bashforth 2 update the Julia binding
    calls: on active cells T + Δt((3/2 + χ)G − (1/2 + χ)G⁻) evaluated as the oracle's step does
    (bit for bit: the library is built without contraction), inactive cells untouched; forward
    Euler (χ = −1/2) does not read G⁻; a tendency view that shares T's storage is refused."""
    import ctypes as C
    from numericalearth_hip._lib import check
    from numericalearth_hip.model import stream_handle, torch_view
    case = make_case(37, 19, 5, seed=31)
    gm = OceanOnlyModel(case)
    Nz, Ny, Nx = gm.T.shape
    g = torch.Generator(device="cuda").manual_seed(7)
    G = [torch.randn((Nz, Ny, Nx), dtype=torch.float64, device="cuda", generator=g) * 1e-6 for _ in range(4)]
    active = np.arange(Nz)[:, None, None] >= case.grid.kmin[None]
    dt, chi = 1200.0, abi.NE_AB2_CHI
    for euler in (0, 1):
        T0, S0 = _np(gm.T).copy(), _np(gm.S).copy()
        gmT = G[2] if not euler else torch.full_like(G[2], float("nan"))   # never read on an Euler step
        check(gm.lib.ne_tracer_ab2_update_f64(gm.ctx.handle, C.byref(gm.grid.struct), torch_view(gm.T),
                                              torch_view(gm.S), torch_view(G[0]), torch_view(G[1]), torch_view(gmT),
                                              torch_view(G[3]), dt, chi, euler, stream_handle()),
              "ne_tracer_ab2_update_f64")
        torch.cuda.synchronize()
        c = -0.5 if euler else chi
        a, b = 1.5 + c, 0.5 + c
        Gm = [np.zeros_like(T0), np.zeros_like(S0)] if euler else [_np(G[2]), _np(G[3])]
        Tref = np.where(active, T0 + dt * (a * _np(G[0]) - b * Gm[0]), T0)
        Sref = np.where(active, S0 + dt * (a * _np(G[1]) - b * Gm[1]), S0)
        assert np.array_equal(_np(gm.T), Tref) and np.array_equal(_np(gm.S), Sref), f"euler={euler}"
    assert gm.lib.ne_tracer_ab2_update_f64(gm.ctx.handle, C.byref(gm.grid.struct), torch_view(gm.T), torch_view(gm.S),
                                           torch_view(gm.T), torch_view(G[1]), torch_view(G[2]), torch_view(G[3]),
                                           dt, chi, 0, stream_handle()) == abi.NE_ERR_SHAPE
    gm.close()
