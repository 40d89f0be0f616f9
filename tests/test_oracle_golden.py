"""Pins of the CPU oracle (oracle/ne_oracle.c) — no GPU.

Each test checks the oracle against something it was not derived from: a known answer
of the reference's own test suite, a closed form or limit fixed by the mathematics, an
integral identity evaluated by mpmath quadrature, a physical constant, or an invariant of
the discretisation (telescoping, exactness, convergence order, symmetry).
"""
import ctypes as C
import math

import mpmath as mp
import numpy as np
import pytest

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import make_case
from numericalearth_hip.params import (ComponentInterfaces, ConvectiveGustiness, FixedSpecificHumidity,
                                       SimilarityTheoryFluxes, SubgridVelocityCorrection,
                                       mahrt_sun_subgrid_velocity, to_params)

L = O.lib()
P = to_params()
PP = C.byref(P)


# ---------------------------------------------------------------------------
# Known answers from the reference test suite
# ---------------------------------------------------------------------------

def test_gustiness_bit_identity_and_floor():
    """test_subgrid_velocity_corrections.jl:17-42: vsgs² == Uᴳ² exactly, the stable floor, zero gustiness."""
    for ustar, bstar, hbl in ((0.3, -0.005, 600.0), (0.3, 0.005, 600.0), (1e-4, 0.0, 512.0)):
        Jb = -ustar * bstar
        Ug = max(0.01, 1.2 * float(np.cbrt(max(0.0, Jb) * hbl)))
        assert L.neo_gustiness_squared(PP, ustar, bstar, hbl) == Ug ** 2
    assert L.neo_gustiness_squared(PP, 0.3, -0.005, 600.0) > 0.01 ** 2
    assert L.neo_gustiness_squared(PP, 0.3, 0.005, 600.0) == 0.01 ** 2
    p0 = to_params(SimilarityTheoryFluxes(subgrid_velocities=ConvectiveGustiness(0.0, 0.0)))
    assert L.neo_gustiness_squared(C.byref(p0), 0.3, -0.005, 600.0) == 0.0
    # SubgridVelocityCorrection composes in quadrature
    vsg = mahrt_sun_subgrid_velocity(25e3)
    pm = to_params(SimilarityTheoryFluxes(subgrid_velocities=SubgridVelocityCorrection(ConvectiveGustiness(), vsg)))
    assert L.neo_gustiness_squared(C.byref(pm), 0.3, -0.005, 600.0) == \
        L.neo_gustiness_squared(PP, 0.3, -0.005, 600.0) + vsg ** 2


def test_mahrt_sun():
    """test_subgrid_velocity_corrections.jl:44-50."""
    assert mahrt_sun_subgrid_velocity(1e3) == 0
    assert mahrt_sun_subgrid_velocity(5e3) == 0
    assert math.isclose(mahrt_sun_subgrid_velocity(25e3), 0.32 * 4 ** 0.33, rel_tol=1e-15)
    assert mahrt_sun_subgrid_velocity(25e3) > mahrt_sun_subgrid_velocity(10e3) > 0


def _column(p, ua, va, Ta, pa, qa, ToC, So=30.0, uo=0.0, vo=0.0, inactive=0, SW=0.0, LW=0.0):
    inp = (C.c_double * 11)(ua, va, Ta, pa, qa, ToC, So, uo, vo, SW, LW)
    out = (C.c_double * 9)()
    it = L.neo_flux_column(C.byref(p), inp, inactive, out)
    return list(out), it


def test_neutral_closed_form():
    """test_surface_fluxes.jl:148-211: ψ ≡ 0, ℓ = 1e-4, no gustiness ⇒ u★ = ϰΔU/log(h/ℓ), etc."""
    ell = 1e-4
    p = to_params(SimilarityTheoryFluxes(momentum_roughness_length=ell, temperature_roughness_length=ell,
                                         water_vapor_roughness_length=ell, subgrid_velocities=None,
                                         stability_functions=None))
    h, g = 10.0, 9.80665
    Ta, qa, pa, ua, va = 15 + 273.15, 0.003, 101325.0, 7.3, -2.1
    To = 15.0
    out, it = _column(p, ua, va, Ta, pa, qa, To, 30.0)
    cp = L.neo_cp_m(C.byref(p), qa)
    rho = L.neo_air_density(C.byref(p), Ta, pa, qa)
    Lv = L.neo_latent_heat_vapor(C.byref(p), Ta)
    qo = L.neo_surface_specific_humidity(C.byref(p), pa, To + 273.15)
    dU = math.hypot(ua, va)
    dtheta = Ta - (To + 273.15) + h / cp * g
    dq = qa - qo
    k = 0.4 / math.log(h / ell)
    us, ts, qs = k * dU, k * dtheta, k * dq
    expect = [-rho * us * qs * Lv, -rho * cp * us * ts, -rho * us * qs,
              -rho * us ** 2 * ua / dU, -rho * us ** 2 * va / dU, us, ts, qs]
    for got, exp in zip(out[:8], expect):
        assert math.isclose(got, exp, rel_tol=1e-13), (got, exp)
    assert it == 2  # the second iterate reproduces the first: converged


def test_zero_flux_state():
    """test_surface_fluxes.jl:81-110: Δu = 0, Δq = 0 (fixed qₛ), Tₒ = Tᵃ − 273.15 + h g/cₚ ⇒ fluxes ≈ 0."""
    qa = 0.003
    p = to_params(interfaces=ComponentInterfaces(atmosphere_ocean_interface_specific_humidity=FixedSpecificHumidity(qa)))
    Ta, pa, ua, va = 15 + 273.15, 101325.0, 4.0, 1.5
    cp = L.neo_cp_m(C.byref(p), qa)
    To = Ta - 273.15 + 10.0 / cp * 9.80665
    out, it = _column(p, ua, va, Ta, pa, qa, To, 35.0, uo=ua, vo=va)
    for v in out[:5]:
        assert abs(v) < np.finfo(np.float32).eps


def test_inactive_column_is_the_zero_state():
    out, it = _column(P, 5.0, 1.0, 290.0, 101325.0, 0.01, 18.0, 35.0, 0.1, 0.1, inactive=1)
    assert it == 0 and out[5:8] == [0.0, 0.0, 0.0]
    assert out[8] == 0.0  # 273.15 K − 273.15


# ---------------------------------------------------------------------------
# Stability functions: limits and integral identities (Edson et al. 2013)
# ---------------------------------------------------------------------------

def test_stability_function_limits():
    psi_m = lambda z: L.neo_psi_momentum(PP, z)  # noqa: E731
    psi_s = lambda z: L.neo_psi_scalar(PP, z)  # noqa: E731
    assert psi_m(0.0) == 0.0
    # ψ_s(0⁺) = −1 + B⁺D⁺ − E⁺: the branch jump at neutral stratification
    assert math.isclose(psi_s(0.0), -1 + (2 / 3) * 14.28 - 8.525, rel_tol=1e-12, abs_tol=1e-15)
    assert abs(psi_m(-1e-14)) < 1e-12 and abs(psi_s(-1e-14)) < 1e-12
    # stable branch: the exponential term saturates once A⁺ζ > ζmax (= 50)
    z = 200.0
    assert math.isclose(psi_m(z), -0.7 * z - 0.75 * (z - 5 / 0.35) * math.exp(-50.0) - 0.75 * 5 / 0.35, rel_tol=1e-14)


def _phi_integral(phi, z):
    """ψ(ζ) = ∫₀^ζ (1 − φ(x))/x dx (Monin–Obukhov integral form)."""
    return mp.quad(lambda x: (1 - phi(x)) / x, [0, z])


@pytest.mark.parametrize("z", [-0.01, -0.3, -2.0, -15.0])
def test_unstable_branches_match_flux_profile_integrals(z):
    """Unstable ψ = (1 − f)ψ_K + fψ_C with f = ζ²/(1+ζ²), where ψ_K is the Paulson/Businger–Dyer integral
    of φ = (1 − γζ)^(−1/4) (momentum) or (1 − γζ)^(−1/2) (scalar) and ψ_C the free-convection integral of
    φ = (1 − cζ)^(−1/3) — evaluated by quadrature, independent of the closed forms."""
    mp.mp.dps = 30
    zz = mp.mpf(z)
    f = zz ** 2 / (1 + zz ** 2)
    psiK_m = _phi_integral(lambda x: (1 - 15 * x) ** mp.mpf(-0.25), zz)
    psiC_m = _phi_integral(lambda x: (1 - mp.mpf("10.15") * x) ** (-mp.mpf(1) / 3), zz)
    psiK_s = _phi_integral(lambda x: (1 - 15 * x) ** mp.mpf(-0.5), zz)
    psiC_s = _phi_integral(lambda x: (1 - mp.mpf("34.15") * x) ** (-mp.mpf(1) / 3), zz)
    exp_m = float((1 - f) * psiK_m + f * psiC_m)
    exp_s = float((1 - f) * psiK_s + f * psiC_s)
    assert math.isclose(L.neo_psi_momentum(PP, z), exp_m, rel_tol=1e-12, abs_tol=1e-14)
    assert math.isclose(L.neo_psi_scalar(PP, z), exp_s, rel_tol=1e-12, abs_tol=1e-14)


@pytest.mark.parametrize("z", [0.01, 0.5, 3.0, 40.0])
def test_stable_branches_match_beljaars_holtslag(z):
    """Beljaars & Holtslag (1991) forms with Edson (2013) constants, written in the published arrangement:
    ψ_m = −[aζ + b(ζ − c/d)e^{−dζ} + bc/d], ψ_h = −[(1 + 2aζ/3)^{3/2} + b(ζ − c/d)e^{−dζ} + E⁺]."""
    mp.mp.dps = 30
    zz = mp.mpf(z)
    a, b, cd, d = mp.mpf("0.7"), mp.mpf("0.75"), mp.mpf(5) / mp.mpf("0.35"), mp.mpf("0.35")
    psim = -(a * zz + b * (zz - cd) * mp.e ** (-d * zz) + b * cd)
    bs, ds = mp.mpf(2) / 3, mp.mpf("14.28")
    psis = -((1 + bs * zz) ** mp.mpf(1.5) + bs * (zz - ds) * mp.e ** (-d * zz) + mp.mpf("8.525"))
    assert math.isclose(L.neo_psi_momentum(PP, z), float(psim), rel_tol=1e-13)
    assert math.isclose(L.neo_psi_scalar(PP, z), float(psis), rel_tol=1e-13)


# ---------------------------------------------------------------------------
# Roughness lengths and thermodynamics
# ---------------------------------------------------------------------------

def test_roughness_lengths():
    # Charnock + smooth-flow limits (Edson 2013 eq. 13 / 28 forms) and the caps
    us = 0.3
    lu = L.neo_momentum_roughness(PP, us)
    assert math.isclose(lu, 0.02 * us * us / 9.80665 + 0.11 * 1.5e-5 / us, rel_tol=1e-15)
    assert L.neo_momentum_roughness(PP, 0.0) == 1.0          # ν/u★ → ∞ is capped at ℓmax
    assert L.neo_scalar_roughness(PP, lu, 0.0, 0) == 0.0     # R★ = 0 ⇒ ℓₛ = 0 (roughness_lengths.jl:231)
    Rs = lu * us / 1.5e-5
    assert math.isclose(L.neo_scalar_roughness(PP, lu, us, 1), min(5.85e-5 * Rs ** -0.72, 1.6e-4), rel_tol=1e-14)
    assert L.neo_scalar_roughness(PP, 1e-9, 1e-3, 0) == 1.6e-4  # small R★ hits the cap


def test_saturation_vapor_pressure_pins():
    ptr = 611.657
    assert L.neo_saturation_vapor_pressure(PP, 273.16) == ptr          # exact at the triple point
    # published saturation vapour pressure of water at 20 °C: 2.3392 kPa (the opposite sign of
    # Δcₚ in the docs' Clausius–Clapeyron form would give ≈ 2.39 kPa)
    assert abs(L.neo_saturation_vapor_pressure(PP, 293.15) / 2339.2 - 1) < 5e-3
    # Clausius–Clapeyron: d ln pₛₐₜ/dT = ℒᵥ(T)/(RᵥT²), by central differences
    Rv = 8.3144598 / 0.018015
    for T in (270.0, 290.0, 305.0):
        h = 1e-3
        dlog = (math.log(L.neo_saturation_vapor_pressure(PP, T + h)) -
                math.log(L.neo_saturation_vapor_pressure(PP, T - h))) / (2 * h)
        assert math.isclose(dlog, L.neo_latent_heat_vapor(PP, T) / (Rv * T * T), rel_tol=1e-7)
    # ℒᵥ decreases with temperature (cₚᵥ < cₚₗ)
    assert L.neo_latent_heat_vapor(PP, 300.0) < L.neo_latent_heat_vapor(PP, 280.0)


def test_air_properties_special_cases():
    Rd = 8.3144598 / 0.02897
    assert L.neo_air_density(PP, 290.0, 101325.0, 0.0) == 101325.0 / (Rd * 290.0)
    assert L.neo_cp_m(PP, 0.0) == Rd / (2 / 7)
    assert L.neo_virtual_temperature(PP, 300.0, 0.0) == 300.0
    # moist air is lighter: ρ(q) < ρ(0) and 𝒯ᵥ > T
    assert L.neo_air_density(PP, 290.0, 101325.0, 0.01) < L.neo_air_density(PP, 290.0, 101325.0, 0.0)
    assert L.neo_virtual_temperature(PP, 300.0, 0.01) > 300.0
    # qₛ = 0.98 ε⁻¹pᵥ/(p − (1−ε⁻¹)pᵥ) is below qᵥ⁺ and saturates below 1
    qs = L.neo_surface_specific_humidity(PP, 101325.0, 300.0)
    qsat = L.neo_saturation_specific_humidity(PP, 300.0, 101325.0)
    assert 0.97 * qsat < qs < qsat
    assert L.neo_saturation_specific_humidity(PP, 400.0, 101325.0) < 1.0


# ---------------------------------------------------------------------------
# WENO reconstructions
# ---------------------------------------------------------------------------

def test_weno_reconstructs_quadratics_exactly():
    """Every candidate stencil reproduces a quadratic, so WENO5 returns the exact face value whatever its
    nonlinear weights; WENO3 does the same for linear data. Cells are [n, n+1], the face is x = 0."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        a, b, c = rng.standard_normal(3)
        F = lambda x: a * x ** 3 / 3 + b * x ** 2 / 2 + c * x  # noqa: E731  antiderivative of a x² + b x + c
        avg = [F(n + 1) - F(n) for n in range(-3, 2)]             # cells −3 … 1, upwind cell −1
        assert math.isclose(L.neo_weno5z(*avg), c, rel_tol=1e-12, abs_tol=1e-12)
        s, t = rng.standard_normal(2)
        G = lambda x: s * x ** 2 / 2 + t * x  # noqa: E731  antiderivative of s x + t
        avg1 = [G(n + 1) - G(n) for n in range(-2, 1)]            # cells −2 … 0, upwind cell −1
        assert math.isclose(L.neo_weno3z(*avg1), t, rel_tol=1e-12, abs_tol=1e-13)


def test_weno5_fifth_order_on_smooth_data():
    """Error of the reconstructed face value of sin at a smooth point falls as h⁵ (WENO-Z keeps the
    optimal order at smooth extrema-free points) — pins the candidate stencils and linear weights."""
    errs = []
    for n in (10, 20, 40, 80):
        h = 1.0 / n
        x0 = 0.3
        avg = [(math.cos(x0 + k * h) - math.cos(x0 + (k + 1) * h)) / h for k in range(-3, 2)]
        errs.append(abs(L.neo_weno5z(*avg) - math.sin(x0)))
    orders = [math.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]
    assert min(orders) > 4.5, orders


def test_weno_picks_the_smooth_stencil_at_a_jump():
    # discontinuity between c2 and c3: the reconstruction stays on the upwind side (no new extremum)
    v = L.neo_weno5z(0.0, 0.0, 0.0, 1.0, 1.0)
    assert 0.0 <= v < 1e-6
    v = L.neo_weno5z(1.0, 1.0, 1.0, 1.0, 0.0)
    assert abs(v - 1.0) < 1e-6


# ---------------------------------------------------------------------------
# Tendency invariants on a grid
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def small_model():
    case = make_case(24, 18, 6, seed=5)
    m = O.OracleModel(case).reconcile()
    return case, m


def test_flux_form_telescopes(small_model):
    """Σ V·G = −Σ_top Az·F_top: interior face fluxes cancel pairwise and boundary faces carry none."""
    case, m = small_model
    GT, GS = m.tendency()
    g = case.grid
    V = g.Az[None, :, None] * g.dz[:, None, None]
    total_T = np.sum(V * GT)
    expect_T = -np.sum(g.Az[:, None] * m.net_arrays["J_T"])
    scale = np.sum(np.abs(V * GT))
    assert abs(total_T - expect_T) <= 1e-12 * scale
    total_S = np.sum(V * GS)
    expect_S = -np.sum(g.Az[:, None] * (m.net_arrays["J_S"] + m.S[-1] * m.net_arrays["J_w"]))
    assert abs(total_S - expect_S) <= 1e-12 * np.sum(np.abs(V * GS))


def test_constant_tracer_is_steady_below_the_surface(small_model):
    """w is diagnosed from discrete continuity, so a uniform tracer has zero advective tendency except in
    the surface cells (whose top face carries no advective flux)."""
    case, m = small_model
    c = np.full_like(m.T, 12.5)
    GT, _ = m.tendency(c, c.copy())
    assert np.max(np.abs(GT[:-1])) <= 1e-15 * 12.5 * np.max(np.abs(case.u)) / np.min(case.grid.dz)
    assert np.max(np.abs(GT[-1])) > 0


def test_tendency_periodic_translation_and_mirror():
    case = make_case(32, 12, 5, seed=9)
    m = O.OracleModel(case).reconcile()
    GT, _ = m.tendency()
    # translate by 5 columns (periodic λ): the tendency translates bit for bit
    shift = 5
    c2 = make_case(32, 12, 5, seed=9)
    for name in ("T", "S", "v", "w"):
        setattr(c2, name, np.roll(getattr(case, name), shift, axis=-1))
    u = case.u[..., :32]
    c2.u = np.concatenate([np.roll(u, shift, axis=-1), np.roll(u, shift, axis=-1)[..., :1]], axis=-1)
    c2.grid.kmin = np.roll(case.grid.kmin, shift, axis=-1)
    m2 = O.OracleModel(c2)
    for k in O.NET_NAMES:
        m2.net_arrays[k][...] = np.roll(m.net_arrays[k], shift, axis=-1)
    GT2, _ = m2.tendency()
    assert np.array_equal(GT2, np.roll(GT, shift, axis=-1))
    # mirror in λ with u → −u: the tendency mirrors bit for bit
    c3 = make_case(32, 12, 5, seed=9)
    for name in ("T", "S", "v", "w"):
        setattr(c3, name, getattr(case, name)[..., ::-1].copy())
    uu = case.u.copy()
    uu[..., 32] = uu[..., 0]
    c3.u = -uu[..., ::-1].copy()
    c3.grid.kmin = case.grid.kmin[..., ::-1].copy()
    c3.grid.dx_yface = case.grid.dx_yface
    m3 = O.OracleModel(c3)
    for k in O.NET_NAMES:
        m3.net_arrays[k][...] = m.net_arrays[k][..., ::-1]
    GT3, _ = m3.tendency()
    assert np.array_equal(GT3, GT[..., ::-1])


def test_rk3_budget_closes_to_round_off():
    case = make_case(20, 14, 6, seed=3)
    m = O.OracleModel(case).reconcile()
    g = case.grid
    V = g.Az[None, :, None] * g.dz[:, None, None]
    active = np.arange(g.spec.Nz)[:, None, None] >= g.kmin[None]
    T0 = m.T.copy()
    JT = m.net_arrays["J_T"].copy()
    import ctypes as Cc
    Cc  # noqa
    O.lib().neo_model_ocean_step(Cc.byref(m._model), 1200.0)
    dH = np.sum(np.where(active, V * (m.T - T0), 0.0))
    expect = -1200.0 * np.sum(g.Az[:, None] * JT)
    bound = np.sum(np.where(active, V * np.spacing(np.abs(m.T)), 0.0))
    assert abs(dH - expect) <= bound


def test_freeze_clamp_and_frazil_heat():
    case = make_case(10, 8, 4, seed=2)
    m = O.OracleModel(case)
    m.T[...] = 5.0
    m.T[:, 3, 4] = -4.0
    S = m.S[:, 3, 4].copy()
    m.freeze(1200.0)
    Tm = -0.054 * S
    assert np.array_equal(m.T[:, 3, 4], Tm)
    expect = -np.sum(1020.0 * 3991.86795711963 * (Tm + 4.0) * case.grid.dz / 1200.0)
    assert math.isclose(m.frazil[3, 4], expect, rel_tol=1e-14)
    assert np.all(m.frazil[np.arange(8) != 3] == 0)
    m.T[:, 3, 4] = -4.0
    m.freeze(float("inf"))                    # first call: Δt = ∞, no frazil heat
    assert m.frazil[3, 4] == 0.0


def test_interpolation_invariants():
    case = make_case(40, 20, 4, seed=1)
    m = O.OracleModel(case)
    m.interp()
    # constant pressure (101325 Pa) is reproduced bit for bit
    assert np.all(m.exchange_arrays["p"] == 101325.0)
    # a field linear in the source indices is reproduced exactly away from the bounded edges
    a = case.atmos_grid
    ii, jj = np.meshgrid(np.arange(a.Nx), np.arange(a.Ny))
    lin = (3.0 + 0.5 * jj + 0.0 * ii)[None].astype(float)
    m.atm["T"][...] = lin
    m.interp()
    fj = m.afj
    inner = (fj >= 0) & (fj <= a.Ny - 1)
    expect = 3.0 + 0.5 * fj[:, None] * np.ones((1, case.grid.spec.Nx))
    assert np.allclose(m.exchange_arrays["T"][inner], expect[inner], rtol=0, atol=1e-13)
    # Neumann (edge-value) extension at the bounded latitude edges
    assert np.all(m.exchange_arrays["T"][~inner] == np.where(fj[~inner] < 0, 3.0, 3.0 + 0.5 * (a.Ny - 1))[:, None])


# ---------------------------------------------------------------------------
# TwoColorRadiation (Oceans/radiative_forcing.jl:5-106)
# ---------------------------------------------------------------------------

def _two_color_model(Nx, Ny, Nz, seed, z=(-3000.0, 0.0), tcr=None):
    from numericalearth_hip.params import TwoColorRadiation, to_params
    case = make_case(Nx, Ny, Nz, seed=seed, z=z)
    p = to_params(radiative_forcing=tcr or TwoColorRadiation())
    return case, O.OracleModel(case, params=p).reconcile()


def test_two_color_top_cell_absorbs_the_two_band_fraction():
    """Paulson & Simpson (1977) two-band law: a layer of thickness Δz at the surface absorbs
    1 − ε₁e^{−κ₁Δz} − (1 − ε₁)e^{−κ₂Δz} of the penetrating shortwave. On 1 m levels the top cell's
    heating V·F equals that fraction of Az·J₀, and J₀ = (1 − α)SW↓/(ρₒcₒ) > 0."""
    from numericalearth_hip.params import default_radiative_forcing
    case, m = _two_color_model(12, 8, 40, seed=3, z=(-40.0, 0.0))
    p = m.params
    tcr = default_radiative_forcing()
    J0 = m.shortwave_surface_flux
    SW = m.sw
    np.testing.assert_allclose(J0, (1 - p.ocean_albedo) * SW / (p.ocean_reference_density * p.ocean_heat_capacity),
                               rtol=1e-15)
    zero = {k: np.zeros_like(v) for k, v in m.net_arrays.items()}
    for k, v in zero.items():       # isolate the forcing: no surface fluxes, uniform tracer (no advection)
        m.net_arrays[k][...] = v
    c = np.full_like(m.T, 10.0)
    GT, _ = m.tendency(c, c.copy())
    g = case.grid
    dz = g.dz[-1]
    frac = 1 - tcr.first_color_fraction * math.exp(-tcr.first_absorption_coefficient * dz) \
        - (1 - tcr.first_color_fraction) * math.exp(-tcr.second_absorption_coefficient * dz)
    wet = g.kmin < g.spec.Nz
    heat_top = GT[-1] * dz          # V·F / Az
    np.testing.assert_allclose(heat_top[wet], frac * J0[wet], rtol=1e-12)
    assert 0.5 < frac < 0.6


def test_two_color_column_absorbs_everything_down_to_the_bottom():
    """Σ_k V·F over a column is Az·J₀: all the penetrating shortwave is absorbed, the remainder by the
    lowest active cell (the flux through the face below it is zero) — over immersed bottoms, and on a
    10 m deep column where a quarter of the shortwave reaches the bottom face."""
    for z, Nz in (((-3000.0, 0.0), 30), ((-10.0, 0.0), 5)):
        case, m = _two_color_model(16, 10, Nz, seed=4, z=z)
        for k in m.net_arrays:
            m.net_arrays[k][...] = 0.0
        c = np.full_like(m.T, 7.0)
        GT, _ = m.tendency(c, c.copy())
        g = case.grid
        V = g.Az[None, :, None] * g.dz[:, None, None]
        col = np.sum(V * GT, axis=0)
        wet = g.kmin < g.spec.Nz
        expect = g.Az[:, None] * m.shortwave_surface_flux
        np.testing.assert_allclose(col[wet], expect[wet], rtol=1e-12)
        if z[0] == -3000.0:
            assert np.any(wet & (g.kmin > 0))
        else:
            p = m.params
            reach = p.first_color_fraction * math.exp(-p.first_absorption_coefficient * 10.0) + \
                (1 - p.first_color_fraction) * math.exp(-p.second_absorption_coefficient * 10.0)
            assert reach > 0.25
            # the bottom cell absorbs what reaches its top face, including the part that would leave
            zt = g.z_face[1]
            top = p.first_color_fraction * math.exp(p.first_absorption_coefficient * zt) + \
                (1 - p.first_color_fraction) * math.exp(p.second_absorption_coefficient * zt)
            np.testing.assert_allclose((V[0] * GT[0])[wet], (g.Az[:, None] * m.shortwave_surface_flux * top)[wet],
                                       rtol=1e-12)


def test_two_color_moves_shortwave_out_of_the_surface_flux():
    """With TwoColorRadiation the transmitted shortwave leaves Jᵀ and becomes J₀:
    Jᵀ(none) − Jᵀ(two-colour) = −J₀ on active columns (shortwave_radiative_forcing)."""
    from numericalearth_hip.params import to_params
    case, m2 = _two_color_model(20, 12, 6, seed=6)
    m1 = O.OracleModel(case, params=to_params()).reconcile()
    wet = case.grid.kmin < case.grid.spec.Nz
    diff = m1.net_arrays["J_T"] - m2.net_arrays["J_T"]
    np.testing.assert_allclose(diff[wet], -m2.shortwave_surface_flux[wet], rtol=1e-12, atol=1e-22)
    assert np.all(m1.shortwave_surface_flux == 0.0)
