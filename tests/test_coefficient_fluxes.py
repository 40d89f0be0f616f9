"""Pins of the oracle's coefficient-based fluxes, Large–Yeager transfer coefficients and
skin-temperature balance — no GPU.

Each test checks the oracle against something it was not derived from: the known answers
of the reference's own tests (test_coefficient_based_fluxes.jl, test_surface_fluxes.jl),
the published Large & Yeager (2004) drag polynomial and drag stability correction, the
Businger–Dyer flux-profile integrals behind Paulson's (1970) ψ, the classic bulk formulae
the constant-coefficient fluxes reduce to, and the surface heat balance the converged
skin temperature satisfies.
"""
import ctypes as C
import math

import mpmath as mp
import pytest

import oracle as O
from numericalearth_hip.params import (BulkTemperature, CoefficientBasedFluxes, ComponentInterfaces,
                                       ConvergenceStopCriteria, DiffusiveFlux, FixedIterations,
                                       FixedSpecificHumidity, InteriorDiffusivity, LargeYeagerTransferCoefficients,
                                       PolynomialNeutralDragCoefficient, SimilarityScales, SimilarityTheoryFluxes,
                                       SkinTemperature, SurfaceRadiationProperties, to_params)

L = O.lib()
G = 9.80665
K0 = 273.15


def _column(p, ua, va, Ta, pa, qa, ToC, uo=0.0, vo=0.0, SW=0.0, LW=0.0):
    inp = (C.c_double * 11)(ua, va, Ta, pa, qa, ToC, 35.0, uo, vo, SW, LW)
    out = (C.c_double * 9)()
    it = L.neo_flux_column(C.byref(p), inp, 0, out)
    names = ("Qv", "Qt", "Jv", "rtx", "rty", "us", "ts", "qs", "Ts")
    return dict(zip(names, out)), it


def _poly_mp(U, a=2.7, b=0.142, c=1 / 13.09, d=3.14807e-10):
    """Large & Yeager (2004) eq. 11a with the U⁶ term of Large & Yeager (2009), in mpmath."""
    U = mp.mpf(U)
    return float((mp.mpf(a) / U + mp.mpf(b) + mp.mpf(c) * U - mp.mpf(d) * U ** 6) / 1000)


LY = CoefficientBasedFluxes(transfer_coefficients=LargeYeagerTransferCoefficients())


# ---------------------------------------------------------------------------
# PolynomialNeutralDragCoefficient and the Large-Yeager stability functions
# ---------------------------------------------------------------------------

def test_polynomial_neutral_drag_known_answers():
    """test_coefficient_based_fluxes.jl:16-37, and the published polynomial at 1-ulp level."""
    p = C.byref(to_params(LY))
    d = lambda U: L.neo_neutral_drag(p, U)  # noqa: E731
    assert 0 < d(3.0) < 5e-3
    assert d(40.0) == 2.34e-3                 # high-wind cap
    assert d(33.0) == 2.34e-3 and d(32.999) != 2.34e-3
    assert d(0.0) == d(0.5)                   # wind floor
    assert d(20.0) > d(5.0)
    for U in (0.5, 1.0, 3.0, 7.5, 10.0, 18.0, 32.0):
        assert math.isclose(d(U), _poly_mp(U), rel_tol=4e-16), U
    assert math.isclose(d(10.0), 1.1756e-3, rel_tol=1e-4)    # C_dN(10 m/s) of Large & Yeager (2004), fig. 1


def test_linear_stable_and_split_profiles_known_answers():
    """test_coefficient_based_fluxes.jl:39-57."""
    p = C.byref(to_params(LY))
    for scalar in (0, 1):
        psi = lambda z: L.neo_psi_large_yeager(p, scalar, z)  # noqa: E731
        assert psi(-1.0) > 0
        assert psi(1.0) == -5.0
        assert psi(20.0) == -50.0              # bounded at ζmax = 10
        assert abs(psi(0.0)) < 1e-10
    assert L.neo_psi_large_yeager(p, 0, 0.0) == 0.0 or abs(L.neo_psi_large_yeager(p, 0, 0.0)) < 1e-15


def test_paulson_profiles_are_businger_dyer_integrals():
    """ψ(ζ) = ∫₀^ζ (1 − φ(x))/x dx with φₘ = (1 − 16x)^(−1/4), φₕ = (1 − 16x)^(−1/2) (Paulson 1970)."""
    p = C.byref(to_params(LY))
    mp.mp.dps = 30
    for scalar, expo in ((0, mp.mpf(-1) / 4), (1, mp.mpf(-1) / 2)):
        for z in (-1e-3, -0.05, -0.3, -1.0, -4.0, -25.0):
            ref = mp.quad(lambda x: (1 - (1 - 16 * x) ** expo) / x, [0, z / 2, z])
            got = L.neo_psi_large_yeager(p, scalar, z)
            # the closed form sums O(1) logs and π/2 into a small ψ near ζ = 0: a few ulp of 1 absolute
            assert math.isclose(got, float(ref), rel_tol=1e-13, abs_tol=2e-15), (scalar, z, got, ref)


# ---------------------------------------------------------------------------
# Constant and polynomial transfer coefficients
# ---------------------------------------------------------------------------

ATM = dict(ua=7.3, va=-2.1, Ta=288.4, pa=101325.0, qa=0.0085)


def _atm_derived(p, Ta, pa, qa, h=10.0):
    pp = C.byref(p)
    cp = L.neo_cp_m(pp, qa)
    return dict(cp=cp, rho=L.neo_air_density(pp, Ta, pa, qa), Lv=L.neo_latent_heat_vapor(pp, Ta), theta=Ta + G * h / cp)


@pytest.mark.parametrize("form", ["tuple", "dict", "scales"])
def test_constant_coefficients_are_the_bulk_formulae(form):
    Cd, Ch, Cq = 1.3e-3, 1.1e-3, 1.2e-3
    tc = {"tuple": (Cd, Ch, Cq), "dict": {"momentum": Cd, "temperature": Ch, "water_vapor": Cq},
          "scales": SimilarityScales(Cd, Ch, Cq)}[form]
    p = to_params(CoefficientBasedFluxes(transfer_coefficients=tc))
    uo, vo, ToC = 0.4, -0.2, 17.0
    out, it = _column(p, ToC=ToC, uo=uo, vo=vo, **ATM)
    d = _atm_derived(p, ATM["Ta"], ATM["pa"], ATM["qa"])
    du, dv = ATM["ua"] - uo, ATM["va"] - vo
    U = math.hypot(du, dv)
    dtheta = d["theta"] - (ToC + K0)
    qs = L.neo_surface_specific_humidity(C.byref(p), ATM["pa"], ToC + K0)
    dq = ATM["qa"] - qs
    assert it == 2                                    # the second iterate repeats the first
    assert math.isclose(out["us"], math.sqrt(Cd) * U, rel_tol=1e-15)
    assert math.isclose(out["ts"], Ch / math.sqrt(Cd) * dtheta, rel_tol=1e-14)
    assert math.isclose(out["qs"], Cq / math.sqrt(Cq * 0 + Cd) * dq, rel_tol=1e-14)
    # sensible heat ρ cₚ Cₕ U (Tₛ − θₐ), latent ρ ℒᵥ C_q U (qₛ − qₐ), stress ρ C_d U Δu
    assert math.isclose(out["Qt"], d["rho"] * d["cp"] * Ch * U * -dtheta, rel_tol=1e-13)
    assert math.isclose(out["Qv"], d["rho"] * d["Lv"] * Cq * U * -dq, rel_tol=1e-13)
    assert math.isclose(out["rtx"], -d["rho"] * Cd * U * du, rel_tol=1e-13)
    assert math.isclose(out["rty"], -d["rho"] * Cd * U * dv, rel_tol=1e-13)
    assert out["Ts"] == ToC
    # FixedIterations(1) reaches the same state
    p1 = to_params(CoefficientBasedFluxes(transfer_coefficients=tc, solver_stop_criteria=FixedIterations(1)))
    out1, it1 = _column(p1, ToC=ToC, uo=uo, vo=vo, **ATM)
    assert it1 == 1 and out1 == out


def test_coefficient_validation_and_defaults():
    """validate_coefficients (coefficient_based_turbulent_fluxes.jl:231-258) and the constructor defaults."""
    with pytest.raises(ValueError):
        CoefficientBasedFluxes(transfer_coefficients=(1e-3, 1e-3))
    with pytest.raises(ValueError):
        CoefficientBasedFluxes(transfer_coefficients={"momentum": 1e-3, "temperature": 1e-3})
    f = CoefficientBasedFluxes()
    assert f.transfer_coefficients == (1e-3, 1e-3, 1e-3)
    assert isinstance(f.solver_stop_criteria, ConvergenceStopCriteria)
    assert (f.solver_stop_criteria.tolerance, f.solver_stop_criteria.maxiter) == (1e-8, 20)
    ly = LargeYeagerTransferCoefficients()
    assert (ly.reference_height, ly.stable_heat_transfer_coefficient, ly.unstable_heat_transfer_coefficient,
            ly.moisture_transfer_coefficient) == (10.0, 18.0, 32.7, 34.6)
    with pytest.raises(NotImplementedError):
        to_params(CoefficientBasedFluxes(transfer_coefficients=(lambda *a: 1e-3, 1e-3, 1e-3)))
    with pytest.raises(NotImplementedError):
        to_params(interfaces=ComponentInterfaces(
            atmosphere_ocean_interface_temperature=SkinTemperature(DiffusiveFlux(InteriorDiffusivity(), 1.0))))


def test_polynomial_momentum_coefficient():
    """test_coefficient_based_fluxes.jl:105-140: SimilarityScales(polynomial, 1e-3, 1e-3) ⇒ u★ = √C_dN(ΔU) ΔU > 0."""
    p = to_params(CoefficientBasedFluxes(
        transfer_coefficients=SimilarityScales(PolynomialNeutralDragCoefficient(), 1e-3, 1e-3)))
    out, it = _column(p, ToC=15.0, **ATM)
    U = math.hypot(ATM["ua"], ATM["va"])
    assert math.isclose(out["us"], math.sqrt(_poly_mp(U)) * U, rel_tol=1e-15)
    assert out["us"] > 0 and all(math.isfinite(v) for v in out.values())
    # zero wind: no floor outside Large-Yeager, so u★ = 0 and the scalar scales vanish with it
    out0, _ = _column(p, ToC=15.0, ua=0.0, va=0.0, Ta=ATM["Ta"], pa=ATM["pa"], qa=ATM["qa"])
    assert out0["us"] == 0.0 and out0["Qt"] == 0.0 and out0["rtx"] == 0.0


# ---------------------------------------------------------------------------
# Large-Yeager transfer coefficients
# ---------------------------------------------------------------------------

def test_large_yeager_known_answer_fixed_iterations():
    """test_coefficient_based_fluxes.jl:142-178: FixedIterations(5), T = 15 °C, S = 35: finite fluxes, u★ > 0."""
    p = to_params(CoefficientBasedFluxes(transfer_coefficients=LargeYeagerTransferCoefficients(),
                                         solver_stop_criteria=FixedIterations(5)))
    out, it = _column(p, ToC=15.0, **ATM)
    assert it == 5 and out["us"] > 0 and all(math.isfinite(v) for v in out.values())


@pytest.mark.parametrize("sign,number", [(-1.0, 32.7), (1.0, 18.0)])
def test_large_yeager_neutral_stanton_and_dalton(sign, number):
    """At the reference height (ln(z/h₀) = 0) and near-neutral stratification, C_d = C_dN(ΔU),
    θ★/Δθ = C_hN/√C_dN = number·10⁻³ (unstable 32.7, stable 18; Large & Yeager 2004 eq. 6c) and
    q★/Δq = 34.6·10⁻³ (eq. 6d)."""
    q0 = 0.0085
    fixed = ComponentInterfaces(atmosphere_ocean_interface_specific_humidity=FixedSpecificHumidity(q0))
    p = to_params(LY, fixed)
    d = _atm_derived(p, 288.4, 101325.0, q0)
    ToC = d["theta"] - K0 - sign * 1e-6          # Δθ = ±1e-6 K: |ζ| ~ 1e-7, ψ corrections ~ 1e-8
    out, it = _column(p, ua=7.3, va=-2.1, Ta=288.4, pa=101325.0, qa=q0, ToC=ToC)
    U = math.hypot(7.3, -2.1)
    dtheta = d["theta"] - (ToC + K0)
    assert 2 <= it < 20
    assert math.isclose(out["us"], math.sqrt(_poly_mp(U)) * U, rel_tol=1e-6)
    assert math.isclose(out["ts"] / dtheta, number * 1e-3, rel_tol=1e-6)
    assert out["qs"] == 0.0
    # a tiny humidity difference instead, neutral temperature: the Dalton number
    pq = to_params(LY, ComponentInterfaces(atmosphere_ocean_interface_specific_humidity=FixedSpecificHumidity(q0 - 1e-7)))
    outq, _ = _column(pq, ua=7.3, va=-2.1, Ta=288.4, pa=101325.0, qa=q0, ToC=d["theta"] - K0)
    assert math.isclose(outq["qs"] / 1e-7, 34.6e-3, rel_tol=1e-4)


def test_large_yeager_drag_height_correction():
    """With neutral scales (u★ = 0 ⇒ C_dp = C_dN(ΔU), θ★ = q★ = 0 ⇒ ζ = 0) the coefficients are Large &
    Yeager (2004) eqs. 10a/11a: U_N10 = ΔU/(1 + √C_dN(ΔU)/κ ln(z/10)), C_d = C_dN(U_N10)/(1 + √C_dN(U_N10)/κ ln(z/10))²."""
    p = C.byref(to_params(LY))
    mp.mp.dps = 30
    for dU, z in ((8.0, 2.0), (3.0, 25.0), (15.0, 10.0), (0.2, 2.0)):
        out = (C.c_double * 3)()
        L.neo_large_yeager_coefficients(p, dU, 290.0, 0.01, z, 0.0, 0.0, 0.0, out)
        U = max(mp.mpf(dU), mp.mpf("0.5"))
        lz = mp.log(mp.mpf(z) / 10) / mp.mpf("0.4")
        UN = max(U / (1 + mp.sqrt(_poly_mp(U)) * lz), mp.mpf("0.5"))
        CdN = mp.mpf(_poly_mp(UN))
        Cd = CdN / (1 + mp.sqrt(CdN) * lz) ** 2
        assert math.isclose(out[0], float(Cd), rel_tol=1e-14), (dU, z)
        if z == 10.0:   # at the reference height the neutral numbers are the coefficients
            assert math.isclose(out[1], math.sqrt(float(CdN)) * 32.7e-3, rel_tol=1e-14)
            assert math.isclose(out[2], math.sqrt(float(CdN)) * 34.6e-3, rel_tol=1e-14)


def _psi_businger_dyer(zeta, scalar):
    """ψ from the flux-profile integral (unstable) or −5 min(ζ, 10) (stable), in mpmath."""
    if zeta > 0:
        return -5 * min(mp.mpf(zeta), 10)
    expo = mp.mpf(-1) / (2 if scalar else 4)
    return mp.quad(lambda x: (1 - (1 - 16 * x) ** expo) / x, [0, mp.mpf(zeta) / 2, mp.mpf(zeta)])


@pytest.mark.parametrize("ts", [-0.05, 0.004, 0.05])
def test_large_yeager_stratified_corrections(ts):
    """Given previous-iterate scales, the coefficients are the stability-corrected forms with the
    Monin–Obukhov ζ = z κ b★/u★² and ψ from the flux-profile integrals: U_N10 = ΔU/(1 + √(u★²/ΔU²)/κ
    (ln(z/10) − ψₘ)), C_d = C_dN(U_N10)/(1 + √C_dN/κ (ln(z/10) − ψₘ))² (Large & Yeager 2004 eq. 10a), and
    C_h = C_hN √(C_d/C_dN)/(1 + C_hN √C_dN/κ (ln(z/10) − ψₕ)) as the reference writes it
    (coefficient_based_turbulent_fluxes.jl:333-337), C_q likewise with 34.6."""
    p = to_params(LY)
    pp = C.byref(p)
    mp.mp.dps = 30
    us, qs, z, dU, Ts, qsurf = 0.3, -1e-4, 2.0, 8.0, 290.0, 0.012
    out = (C.c_double * 3)()
    L.neo_large_yeager_coefficients(pp, dU, Ts, qsurf, z, us, ts, qs, out)
    eps = mp.mpf(0.02897) / mp.mpf(0.018015)
    Tv = (1 + (eps - 1) * qsurf) * mp.mpf(Ts)
    bstar = mp.mpf(G) / Tv * (ts * (1 + (eps - 1) * qsurf) + (eps - 1) * Tv * qs)
    zeta = z * mp.mpf("0.4") * bstar / mp.mpf(us) ** 2
    psim, psih = _psi_businger_dyer(zeta, False), _psi_businger_dyer(zeta, True)
    lz = mp.log(mp.mpf(z) / 10)
    UN = max(dU / (1 + (mp.mpf(us) / dU) / mp.mpf("0.4") * (lz - psim)), mp.mpf("0.5"))
    CdN = mp.mpf(_poly_mp(UN))
    Cd = CdN / (1 + mp.sqrt(CdN) / mp.mpf("0.4") * (lz - psim)) ** 2
    number = 18 if zeta > 0 else mp.mpf("32.7")
    ChN, CqN = mp.sqrt(CdN) / 1000 * number, mp.sqrt(CdN) / 1000 * mp.mpf("34.6")
    ratio = mp.sqrt(Cd / CdN)
    Ch = ChN * ratio / (1 + ChN * mp.sqrt(CdN) / mp.mpf("0.4") * (lz - psih))
    Cq = CqN * ratio / (1 + CqN * mp.sqrt(CdN) / mp.mpf("0.4") * (lz - psih))
    for got, want, name in ((out[0], Cd, "Cd"), (out[1], Ch, "Ch"), (out[2], Cq, "Cq")):
        assert math.isclose(got, float(want), rel_tol=1e-12), (ts, name, got, float(want))


def test_large_yeager_converged_state_is_a_fixed_point():
    """The converged scales reproduce themselves: u★ = √C_d ΔU, θ★ = C_h/√C_d Δθ, q★ = C_q/√C_d Δq
    with the coefficients evaluated at the converged scales, in stable and unstable air."""
    p = to_params(CoefficientBasedFluxes(transfer_coefficients=LargeYeagerTransferCoefficients(),
                                         solver_tolerance=1e-15, solver_maxiter=200))
    pp = C.byref(p)
    for ToC in (10.0, 22.0):                 # θₐ ≈ 15.4 °C: stable, then unstable
        out, it = _column(p, ToC=ToC, **ATM)
        assert it < 200
        d = _atm_derived(p, ATM["Ta"], ATM["pa"], ATM["qa"])
        Ts = ToC + K0
        qs = L.neo_surface_specific_humidity(pp, ATM["pa"], Ts)
        U = math.hypot(ATM["ua"], ATM["va"])
        c = (C.c_double * 3)()
        L.neo_large_yeager_coefficients(pp, U, Ts, qs, 10.0, out["us"], out["ts"], out["qs"], c)
        assert math.isclose(out["us"], math.sqrt(c[0]) * U, rel_tol=1e-12)
        assert math.isclose(out["ts"], c[1] / math.sqrt(c[0]) * (d["theta"] - Ts), rel_tol=1e-12)
        assert math.isclose(out["qs"], c[2] / math.sqrt(c[0]) * (ATM["qa"] - qs), rel_tol=1e-12)
        assert (out["ts"] > 0) == (ToC < 15.0)


# ---------------------------------------------------------------------------
# SkinTemperature
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("temperature", [BulkTemperature(), SkinTemperature(DiffusiveFlux(1e-2, 1.0))])
def test_skin_zero_flux_known_answer(temperature):
    """test_surface_fluxes.jl:78-109: Δθ = 0, qₛ = qₐ, ocean moving with the wind, no radiation ⇒ zero fluxes."""
    qa = 0.003
    ifc = ComponentInterfaces(atmosphere_ocean_interface_specific_humidity=FixedSpecificHumidity(qa),
                              atmosphere_ocean_interface_temperature=temperature)
    p = to_params(SimilarityTheoryFluxes(), ifc, radiation=None)
    Ta, pa, ua, va = 15 + K0, 101325.0, 5.2, -3.1
    ToC = Ta - K0 + 10.0 / L.neo_cp_m(C.byref(p), qa) * G
    out, _ = _column(p, ua=ua, va=va, Ta=Ta, pa=pa, qa=qa, ToC=ToC, uo=ua, vo=va)
    eps = 2.0 ** -52
    for k in ("rtx", "rty", "Qt", "Qv", "Jv"):
        assert abs(out[k]) < eps, (k, out[k])


def _balance_residual(p, out, ToC, Ta, pa, qa, SW, LW):
    """κ(Tᵢ − Tₛ)/δ − (𝒬ᵀ + 𝒬ᵛ(ℒₛ) + σεTₛ⁴ − (1−α)SW − εLW)/(ρₒcₒ), and the magnitude of its terms."""
    pp = C.byref(p)
    Ts = out["Ts"] + K0
    rho, cp = L.neo_air_density(pp, Ta, pa, qa), L.neo_cp_m(pp, qa)
    Ls = 2834400.0 + (1859.0 - 2100.0) * (Ta - 273.16)       # latent_heat_sublim, Thermodynamics.jl
    QT = -rho * cp * out["us"] * out["ts"]
    Qv = -rho * Ls * out["us"] * out["qs"]
    up = 5.670374419e-8 * 0.97 * Ts ** 4 if p.radiation else 0.0
    down = (-(1 - 0.05) * SW - 0.97 * LW) if p.radiation else 0.0
    lam = 1.0 / (p.ocean_reference_density * p.ocean_heat_capacity)
    cond = p.skin_diffusivity * ((ToC + K0) - Ts) / p.skin_depth
    terms = (QT, Qv, up, down)
    return cond - sum(terms) * lam, (abs(cond) + sum(abs(t) for t in terms) * lam)


@pytest.mark.parametrize("fluxes", ["similarity", "large_yeager", "constant"])
@pytest.mark.parametrize("radiation", [True, False])
def test_skin_temperature_balances_the_surface_heat_flux(fluxes, radiation):
    """At the fixed point the conductive flux κ(Tᵢ − Tₛ)/δ through the skin equals the net surface
    heat flux (turbulent with the sublimation latent heat, and radiative) over ρₒcₒ."""
    fl = {"similarity": SimilarityTheoryFluxes(solver_stop_criteria=FixedIterations(150)),
          "large_yeager": CoefficientBasedFluxes(transfer_coefficients=LargeYeagerTransferCoefficients(),
                                                 solver_stop_criteria=FixedIterations(150)),
          "constant": CoefficientBasedFluxes(solver_stop_criteria=FixedIterations(150))}[fluxes]
    ifc = ComponentInterfaces(atmosphere_ocean_interface_temperature=SkinTemperature(DiffusiveFlux(1e-4, 1.0)))
    p = to_params(fl, ifc, radiation=SurfaceRadiationProperties() if radiation else None)
    SW, LW = (420.0, 330.0) if radiation else (0.0, 0.0)
    for ToC in (12.0, 19.0):
        out, it = _column(p, ToC=ToC, SW=SW, LW=LW, **ATM)
        assert it == 150
        r, scale = _balance_residual(p, out, ToC, ATM["Ta"], ATM["pa"], ATM["qa"], SW, LW)
        assert abs(r) <= 1e-10 * scale, (fluxes, radiation, ToC, r, scale)
        assert abs(out["Ts"] - ToC) < 5.0           # the clamp is not active at this coupling
        assert out["Ts"] != ToC


def test_skin_temperature_clamp():
    """max_ΔT = 0 pins Tₛ to the top-cell temperature: the iterates are those of BulkTemperature,
    bit for bit; a weak coupling (tiny κ) saturates the step at ±max_ΔT."""
    for fl in (SimilarityTheoryFluxes(), CoefficientBasedFluxes(transfer_coefficients=LargeYeagerTransferCoefficients())):
        bulk = to_params(fl)
        skin0 = to_params(fl, ComponentInterfaces(
            atmosphere_ocean_interface_temperature=SkinTemperature(DiffusiveFlux(1e-4, 1.0), max_ΔT=0.0)))
        a, ita = _column(bulk, ToC=19.0, SW=400.0, LW=330.0, **ATM)
        b, itb = _column(skin0, ToC=19.0, SW=400.0, LW=330.0, **ATM)
        assert a == b and ita == itb
    weak = to_params(SimilarityTheoryFluxes(solver_stop_criteria=FixedIterations(3)), ComponentInterfaces(
        atmosphere_ocean_interface_temperature=SkinTemperature(DiffusiveFlux(1e-9, 1.0), max_ΔT=0.75)))
    out, _ = _column(weak, ToC=19.0, SW=0.0, LW=300.0, **ATM)
    assert abs(abs(out["Ts"] - 19.0) - 0.75) < 1e-12


def test_latent_heat_of_sublimation():
    """ℒₛ(T₀) = ℒₛ₀, slope cₚᵥ − cₚᵢ, and ℒₛ − ℒᵥ = the latent heat of fusion ≈ 333.6 kJ/kg at T₀."""
    p = C.byref(to_params())
    assert L.neo_latent_heat_sublim(p, 273.16) == 2834400.0
    assert math.isclose(L.neo_latent_heat_sublim(p, 283.16) - L.neo_latent_heat_sublim(p, 273.16),
                        (1859.0 - 2100.0) * 10.0, rel_tol=1e-9)
    assert abs(L.neo_latent_heat_sublim(p, 273.16) - L.neo_latent_heat_vapor(p, 273.16) - 333.6e3) < 0.5e3


def test_pow_int_is_correctly_rounded():
    mp.mp.dps = 40
    for x in (1.0000001, 3.3, 17.25, 288.15, 301.7, 0.731, 32.99):
        for n in (4, 6):
            assert L.neo_pow_int(x, n) == float(mp.mpf(x) ** n), (x, n)
