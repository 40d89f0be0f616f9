"""Pins of the oracle's net ocean flux assembly, radiative contribution and split
Runge-Kutta step — no GPU.

The functions pinned here are restatements of
  * _assemble_net_ocean_fluxes! (src/Oceans/assemble_net_ocean_fluxes.jl:74-153):
    Jᵀ = (𝒬ᵀ + 𝒬ᵛ)/(ρₒcₒ), Jʷ = (rain + snow)/ρₒ − Jᵛ/ρₒ, Jˢ = 0, Jᴴ = SST·Jʷ,
    τˣ = ℑx(ρτˣ)/ρₒ, τʸ = ℑy(ρτʸ)/ρₒ, zero on inactive columns;
  * _apply_air_sea_radiative_fluxes! (src/Radiations/apply_air_sea_radiative_fluxes.jl:62-111,
    radiation_kernels.jl:3-5): Jᵀ += (σεTₛ⁴ − εLW↓ − (1 − α)SW↓)/(ρₒcₒ);
  * the split Runge-Kutta step (src/Oceans/ocean_simulation.jl:347): c⁽ᵐ⁾ = c⁰ + βₘΔt G(c⁽ᵐ⁻¹⁾)
    with β = (1/3, 1/2, 1).
Each test checks them against what physics or mathematics fixes independently of the
restatement: SI definitions of the constants, radiative equilibrium, energy and volume
budgets in SI units with hand-chosen fluxes, the physical direction of salinity change
under rain and evaporation, face averages of uniform and linear fields, and the Taylor
polynomial that a third-order Runge-Kutta step must reproduce on a linear equation.
"""
import ctypes as C
import math

import mpmath as mp
import numpy as np
import pytest

import oracle as O
from numericalearth_hip.generators import make_case
from numericalearth_hip.params import to_params

P = to_params()
RHO_O = 1020.0                         # ocean reference density (component_interfaces.jl:433)
C_O = 3991.86795711963                 # TEOS-10 cₚ⁰, J kg⁻¹ K⁻¹ (published constant)
EMISSIVITY, ALBEDO = 0.97, 0.05        # PrescribedRadiation ocean surface (prescribed_radiation.jl:68)


def test_radiation_constants_are_the_published_ones():
    """σ from its SI definition 2π⁵k⁴/(15h³c²) with the exact 2019 SI values of k, h and c, and the
    ocean surface properties and TEOS-10 heat capacity the reference uses."""
    mp.mp.dps = 40
    k, h, c = mp.mpf("1.380649e-23"), mp.mpf("6.62607015e-34"), mp.mpf(299792458)
    sigma = 2 * mp.pi ** 5 * k ** 4 / (15 * h ** 3 * c ** 2)
    assert abs(P.stefan_boltzmann_constant / float(sigma) - 1) < 1e-9
    assert (P.ocean_emissivity, P.ocean_albedo) == (EMISSIVITY, ALBEDO)
    assert (P.ocean_reference_density, P.ocean_heat_capacity) == (RHO_O, C_O)


@pytest.fixture()
def column_model():
    """A small oracle model whose surface inputs the tests overwrite by hand."""
    case = make_case(8, 6, 3, seed=11)
    m = O.OracleModel(case).reconcile()
    wet = case.grid.kmin < case.grid.spec.Nz
    assert wet.any() and (~wet).any()
    return m, wet


def _surface(m, *, Qt=0.0, Qv=0.0, Jv=0.0, rain=0.0, snow=0.0, sst=15.0, Ts=15.0, SW=0.0, LW=0.0,
             rtx=0.0, rty=0.0):
    f = m.flux_arrays
    f["sensible_heat"][...] = Qt
    f["latent_heat"][...] = Qv
    f["water_vapor"][...] = Jv
    f["x_momentum"][...] = rtx
    f["y_momentum"][...] = rty
    f["interface_temperature"][...] = Ts
    m.exchange_arrays["rain"][...] = rain
    m.exchange_arrays["snow"][...] = snow
    m.T[-1][...] = sst
    m.sw[...] = SW
    m.lw[...] = LW
    m.assemble()
    m.radiation_apply()
    return m.net_dict()


def test_surface_heat_flux_in_si_units(column_model):
    """Hand-chosen fluxes in W m⁻²: the ocean loses 𝒬ᵀ + 𝒬ᵛ + εσTₛ⁴ − εLW↓ − (1 − α)SW↓, and Jᵀ is that
    power divided by ρₒcₒ (K m s⁻¹, positive upward). Evaluated in 40-digit arithmetic."""
    m, wet = column_model
    Qt, Qv, SW, LW, TsC = 25.0, 110.0, 240.0, 350.0, 19.0
    net = _surface(m, Qt=Qt, Qv=Qv, SW=SW, LW=LW, Ts=TsC)
    mp.mp.dps = 40
    Ts = mp.mpf(TsC) + mp.mpf("273.15")
    power = (Qt + Qv + mp.mpf("5.670374419e-8") * mp.mpf("0.97") * Ts ** 4 - mp.mpf("0.97") * LW
             - (1 - mp.mpf("0.05")) * SW)
    expect = float(power / (mp.mpf(RHO_O) * mp.mpf(C_O)))
    terms = float((abs(Qt) + abs(Qv) + 5.670374419e-8 * 0.97 * Ts ** 4 + 0.97 * LW + SW) / (RHO_O * C_O))
    np.testing.assert_allclose(net["J_T"][wet], expect, rtol=0, atol=2e-15 * terms)   # a sum that cancels
    assert abs(expect) > 0.02 * terms
    assert np.all(net["J_T"][~wet] == 0.0)
    # the radiative diagnostics: emitted, absorbed (downwelling) longwave and transmitted shortwave
    up = float(mp.mpf("5.670374419e-8") * mp.mpf("0.97") * Ts ** 4)
    np.testing.assert_allclose(m.rad_arrays["upwelling_longwave"], up, rtol=1e-15)
    np.testing.assert_allclose(m.rad_arrays["downwelling_longwave"], EMISSIVITY * LW, rtol=1e-15)
    np.testing.assert_allclose(m.rad_arrays["downwelling_shortwave"], (1 - ALBEDO) * SW, rtol=1e-15)


def test_longwave_vanishes_at_radiative_equilibrium(column_model):
    """Kirchhoff: a grey surface of emissivity ε absorbs ε of the downwelling longwave, so under a sky
    radiating σTₛ⁴ the net longwave is zero; a surface warmer than the sky's brightness temperature loses
    heat (Jᵀ > 0), a colder one gains it."""
    m, wet = column_model
    Ts = 288.0
    LW = 5.670374419e-8 * Ts ** 4
    net = _surface(m, Ts=Ts - 273.15, LW=LW)
    assert np.max(np.abs(net["J_T"][wet])) <= 1e-14 * LW / (RHO_O * C_O)
    warm = _surface(m, Ts=Ts - 273.15 + 5.0, LW=LW)
    cold = _surface(m, Ts=Ts - 273.15 - 5.0, LW=LW)
    assert np.all(warm["J_T"][wet] > 0) and np.all(cold["J_T"][wet] < 0)


def test_shortwave_heats_the_ocean(column_model):
    """Sunlight warms the ocean: Jᵀ = −(1 − α)SW↓/(ρₒcₒ) < 0 with no other flux (and a surface at 0 K
    would emit nothing)."""
    m, wet = column_model
    net = _surface(m, SW=300.0, Ts=-273.15)
    np.testing.assert_allclose(net["J_T"][wet], -(1 - ALBEDO) * 300.0 / (RHO_O * C_O), rtol=1e-15)


@pytest.mark.parametrize("rain,snow,Jv", [(3e-5, 0.0, 0.0), (0.0, 2e-5, 0.0), (0.0, 0.0, 4e-5),
                                          (3e-5, 1e-5, 4e-5)])
def test_freshwater_flux_signs_and_units(column_model, rain, snow, Jv):
    """Rain and snow (kg m⁻² s⁻¹, positive down) add volume at (rain + snow)/ρₒ m s⁻¹; evaporation Jᵛ
    (positive up) removes Jᵛ/ρₒ. Jᴴ is that volume flux carried at the SST, Jˢ is zero without sea ice."""
    m, wet = column_model
    sst = 17.5
    net = _surface(m, rain=rain, snow=snow, Jv=Jv, sst=sst)
    expect = (rain + snow - Jv) / RHO_O
    np.testing.assert_allclose(net["J_w"][wet], expect, rtol=1e-15, atol=0)
    np.testing.assert_allclose(net["J_H"][wet], sst * expect, rtol=1e-15, atol=0)
    assert np.all(net["J_S"] == 0.0)
    for name in ("J_w", "J_H"):
        assert np.all(net[name][~wet] == 0.0)


@pytest.mark.parametrize("rain,Jv,sign", [(3e-5, 0.0, -1), (0.0, 4e-5, +1)])
def test_rain_freshens_and_evaporation_salts_the_surface(column_model, rain, Jv, sign):
    """Through the salinity top boundary condition Fˢ = Jˢ + S·Jʷ: rain lowers the surface salinity,
    evaporation raises it, at the rate S·|Jʷ|/Δz with no flow."""
    m, wet = column_model
    net = _surface(m, rain=rain, Jv=Jv)
    for name in ("u", "v", "w"):
        getattr(m, name)[...] = 0.0
    _, GS = m.tendency()
    dz = m.case.grid.dz[-1]
    rate = m.S[-1] * np.abs(net["J_w"]) / dz
    np.testing.assert_allclose(GS[-1][wet], sign * rate[wet], rtol=1e-14)
    assert np.all(GS[:-1] == 0.0)


def test_momentum_flux_face_averages(column_model):
    """τ = ℑ(ρτ)/ρₒ at the west (x) and south (y) faces: a uniform stress arrives unchanged at every face
    of an active column (x is periodic, so column 0 averages with column Nx−1), except the southern edge
    row of the bounded y direction, whose outer neighbour lies outside the domain and carries the zero
    state; a stress linear in i arrives at the face midway between the two centres."""
    m, wet = column_model
    Nx, Ny = m.case.grid.spec.Nx, m.case.grid.spec.Ny
    net = _surface(m, rtx=0.08, rty=-0.05)
    np.testing.assert_allclose(net["tau_x"][wet], 0.08 / RHO_O, rtol=1e-15)
    inner = wet.copy()
    inner[0] = False
    np.testing.assert_allclose(net["tau_y"][inner], -0.05 / RHO_O, rtol=1e-15)
    np.testing.assert_allclose(net["tau_y"][0][wet[0]], 0.5 * -0.05 / RHO_O, rtol=1e-15)
    assert np.all(net["tau_x"][~wet] == 0.0) and np.all(net["tau_y"][~wet] == 0.0)
    # linear in i (west face of column i is at i − ½) and in j (south face of row j at j − ½)
    i = np.arange(Nx)[None, :] * np.ones((Ny, 1))
    j = np.arange(Ny)[:, None] * np.ones((1, Nx))
    m.flux_arrays["x_momentum"][...] = 0.01 + 0.002 * i
    m.flux_arrays["y_momentum"][...] = 0.03 - 0.004 * j
    m.assemble()
    tx, ty = m.net_arrays["tau_x"], m.net_arrays["tau_y"]
    west = wet.copy()
    west[:, 0] = False                                   # column 0's west neighbour is the wrapped Nx−1
    np.testing.assert_allclose(tx[west], ((0.01 + 0.002 * (i - 0.5)) / RHO_O)[west], rtol=1e-14)
    np.testing.assert_allclose(tx[:, 0][wet[:, 0]],
                               (0.5 * (0.01 + 0.01 + 0.002 * (Nx - 1)) / RHO_O) * np.ones(int(wet[:, 0].sum())),
                               rtol=1e-14)
    np.testing.assert_allclose(ty[inner], ((0.03 - 0.004 * (j - 0.5)) / RHO_O)[inner], rtol=1e-14)


# ---------------------------------------------------------------------------
# Split Runge-Kutta step
# ---------------------------------------------------------------------------

def _still_model(Nz=3, seed=4):
    """An oracle model with no flow, so the only tendency is the surface boundary flux."""
    case = make_case(6, 5, Nz, seed=seed, z=(-30.0, 0.0))
    m = O.OracleModel(case).reconcile()
    for name in ("u", "v", "w"):
        getattr(m, name)[...] = 0.0
    for name in m.net_arrays:
        m.net_arrays[name][...] = 0.0
    wet = case.grid.kmin < case.grid.spec.Nz
    return m, wet


def _ocean_step(m, dt):
    O.lib().neo_model_ocean_step(C.byref(m._model), dt)


@pytest.mark.parametrize("Js", [0.0, 2.5e-4])
def test_rk3_reproduces_the_cubic_taylor_polynomial(Js):
    """With no flow the surface salinity obeys dS/dt = −(Jˢ + S·Jʷ)/Δz, a linear equation with rate
    λ = −Jʷ/Δz and fixed point S* = −Jˢ/Jʷ. A third-order three-stage Runge-Kutta step maps the departure
    from S* by the cubic Taylor polynomial of e^{λΔt}: 1 + z + z²/2 + z³/6 with z = λΔt. Any other
    stage coefficients β₁, β₂ give another cubic (β₂ = 1/3 gives 1 + z + z²/3 + z³/9)."""
    m, wet = _still_model()
    dt = 1200.0
    dz = m.case.grid.dz[-1]
    Jw = 0.35 * dz / dt                 # z = −0.35: the cubic term is 7e-3 of the step
    m.net_arrays["J_w"][...] = np.where(wet, Jw, 0.0)
    m.net_arrays["J_S"][...] = np.where(wet, Js, 0.0)
    S0 = m.S.copy()
    T0 = m.T.copy()
    _ocean_step(m, dt)
    z = -Jw * dt / dz
    poly = 1 + z + z * z / 2 + z ** 3 / 6
    Sstar = -Js / Jw
    expect = Sstar + (S0[-1] - Sstar) * poly
    np.testing.assert_allclose(m.S[-1][wet], expect[wet], rtol=4e-15)
    assert abs(poly - math.exp(z)) > 1e-4                  # the step is not the exact exponential
    assert np.array_equal(m.S[:-1], S0[:-1])               # no flow: the interior does not change
    assert np.array_equal(m.T, T0)                         # Jᵀ = 0


def test_rk3_advances_a_constant_forcing_by_one_full_step():
    """A constant surface heat flux changes the surface temperature by −Δt·Jᵀ/Δz in one step whatever β₁,
    β₂ are (the stages see the same tendency); only β₃ = 1 enters."""
    m, wet = _still_model()
    dt = 1200.0
    dz = m.case.grid.dz[-1]
    JT = 3e-5
    m.net_arrays["J_T"][...] = np.where(wet, JT, 0.0)
    T0 = m.T.copy()
    _ocean_step(m, dt)
    np.testing.assert_allclose(m.T[-1][wet], (T0[-1] - dt * JT / dz)[wet], rtol=1e-15)
    assert np.array_equal(m.T[:-1], T0[:-1])

# ---------------------------------------------------------------------------
# Quasi-Adams-Bashforth 2 step (the reference's alternative timestepper)
# ---------------------------------------------------------------------------

def _salinity_decay(m, dt, Js=2.5e-4, z=-0.35):
    """Fluxes for the linear surface-salinity equation of rate λ = z/Δt and fixed point S* = −Jˢ/Jʷ."""
    wet = m.case.grid.kmin < m.case.grid.spec.Nz
    dz = m.case.grid.dz[-1]
    Jw = -z * dz / dt
    m.net_arrays["J_w"][...] = np.where(wet, Jw, 0.0)
    m.net_arrays["J_S"][...] = np.where(wet, Js, 0.0)
    return -Js / Jw


def test_ab2_first_step_and_a_changed_step_are_forward_euler():
    """The first QAB2 step, and a step whose Δt differs from the last one, are forward Euler: the
    departure from S* is multiplied by 1 + z (z = λΔt), not by the AB2 factor; a step with the last
    Δt is not (its factor involves the previous tendency)."""
    m, wet = _still_model()
    m.timestepper = "QuasiAdamsBashforth2"
    dt = 1200.0
    Sstar = _salinity_decay(m, dt, z=-0.35)
    S0 = m.S[-1].copy()
    m.ab2_step(dt)
    m.iteration, m.last_dt = 1, dt
    np.testing.assert_allclose(m.S[-1][wet], (Sstar + (S0 - Sstar) * (1 - 0.35))[wet], rtol=1e-15)
    S1 = m.S[-1].copy()
    m.ab2_step(dt)                                   # same Δt: an AB2 step
    m.iteration, m.last_dt = 2, dt
    assert not np.allclose(m.S[-1][wet], (Sstar + (S1 - Sstar) * (1 - 0.35))[wet], rtol=1e-6)
    S2 = m.S[-1].copy()
    dt2 = 600.0
    Sstar2 = _salinity_decay(m, dt2, z=-0.2)
    m.ab2_step(dt2)                                  # Δt changed: Euler again
    np.testing.assert_allclose(m.S[-1][wet], (Sstar2 + (S2 - Sstar2) * (1 - 0.2))[wet], rtol=1e-15)


@pytest.mark.parametrize("z", [-1e-3, -2e-3])
def test_ab2_local_error_is_chi_z_squared(z):
    """From the exact history of y' = λy (y = S − S*: the state yₙ and the previous tendency
    λyₙe^{−z}), one QAB2 step gives yₙ(1 + z((3/2 + χ) − (1/2 + χ)e^{−z})), whose departure from the
    exact yₙe^{z} is χz²yₙ + O(z³): the scheme is consistent (its two weights sum to one) with
    Oceananigans' χ = 0.1. Coefficients that do not sum to one leave an O(z) error; χ = 0 an O(z³)
    one."""
    m, wet = _still_model()
    m.timestepper = "QuasiAdamsBashforth2"
    dt = 1200.0
    Sstar = _salinity_decay(m, dt, z=z)
    y = (m.S[-1] - Sstar)
    Sprev = m.S.copy()
    Sprev[-1] = Sstar + y * np.exp(-z)
    m.GmT, m.GmS = m.tendency(m.T, Sprev)
    m.iteration, m.last_dt = 5, dt
    m.ab2_step(dt)
    ratio = ((m.S[-1] - Sstar) - y * np.exp(z))[wet] / (y[wet] * z * z)
    assert np.all(np.abs(ratio - 0.1) < 1e-3), ratio
