"""Mutation check of the oracle's pins — no GPU.

Each mutant plants one plausible mistake in a copy of the oracle (the C restatement
oracle/ne_oracle.c, or the Python time indexing in oracle/oracle.py), builds the copy in a
temporary directory, and runs the pins that cover the mutated function against it in a child
process (NE_ORACLE_LIB / NE_ORACLE_DIR select the copy). A mutant survives when every one of
those pins still passes; the test fails if any mutant survives, so a pin that no longer
distinguishes the restatement from a wrong one is noticed. The product is not involved.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

REPO = Path(__file__).resolve().parent.parent
ORACLE_C = REPO / "oracle" / "ne_oracle.c"
ORACLE_PY = REPO / "oracle" / "oracle.py"

GOLDEN = "tests/test_oracle_golden.py"
COEF = "tests/test_coefficient_fluxes.py"
NET = "tests/test_oracle_net_fluxes.py"
TSER = "tests/test_time_series.py"
# the oracle's own pins in that file (its other tests compare the library with the oracle)
TSER_PINS = [TSER, "-k", "not library"]

# (name, file, text of the restatement, planted mistake, pins that must catch it)
MUTANTS = [
    # thermodynamics, similarity theory, WENO, tendency
    ("dcp sign in the saturation vapour pressure", "c",
     "double dcp = p->water_vapor_heat_capacity - p->liquid_water_heat_capacity;\n    double Rv = R_v(p);",
     "double dcp = p->liquid_water_heat_capacity - p->water_vapor_heat_capacity;\n    double Rv = R_v(p);",
     [GOLDEN]),
    ("WENO5 linear weight 0.6 -> 0.5", "c",
     "double a1 = 0.6 * (1.0 + r1 * r1);", "double a1 = 0.5 * (1.0 + r1 * r1);", [GOLDEN]),
    ("sign of the stable momentum constant", "c",
     "double psi_p = -Bp * zp - Cp * (zp - Dp) * exp(-dz) - Cp * Dp;",
     "double psi_p = -Bp * zp - Cp * (zp - Dp) * exp(-dz) + Cp * Dp;", [GOLDEN]),
    ("y-face metric of the neighbouring row", "c",
     "double A = g->dx_yface[jf] * g->dz[k];", "double A = g->dx_yface[jf > 0 ? jf - 1 : jf] * g->dz[k];",
     [GOLDEN]),
    ("sign of the top boundary flux", "c",
     "G = G - top_flux * g->Az[j] / V;", "G = G + top_flux * g->Az[j] / V;", [GOLDEN, NET]),
    ("gustiness floor as a cap", "c",
     "double Ug = jmax(p->minimum_gustiness,", "double Ug = jmin(p->minimum_gustiness,", [GOLDEN]),
    ("candidate stencil q1", "c",
     "double q1 = (-c1 + 5.0 * c2 + 2.0 * c3) / 6.0;", "double q1 = (-c1 + 5.0 * c2 + 2.0 * c4) / 6.0;",
     [GOLDEN]),
    ("WENO3 candidate value", "c",
     "double q0 = (-c0 + 3.0 * c1) / 2.0;", "double q0 = (-c0 + 3.0 * c1) / 3.0;", [GOLDEN]),
    ("momentum's fourth root in the unstable scalar profile", "c",
     "double f1 = sqrt(1.0 - Am * zm);", "double f1 = sqrt(sqrt(1.0 - Am * zm));", [GOLDEN]),
    # coefficient-based fluxes and the skin temperature
    ("stable / unstable Stanton numbers swapped", "c",
     "(stable ? p->stable_heat_transfer_coefficient : p->unstable_heat_transfer_coefficient)",
     "(stable ? p->unstable_heat_transfer_coefficient : p->stable_heat_transfer_coefficient)", [COEF]),
    ("sign of Paulson's arctangent term", "c",
     "double psi3 = -2.0 * atan(z);", "double psi3 = 2.0 * atan(z);", [COEF]),
    ("linear-stable profile without its zeta cap", "c",
     "double stable = -ls[0] * jmin(zp, ls[1]);", "double stable = -ls[0] * zp;", [COEF]),
    ("vaporisation for sublimation in the skin balance", "c",
     "double Li = neo_latent_heat_sublim(p, A->T);", "double Li = neo_latent_heat_vapor(p, A->T);", [COEF]),
    ("sign of the skin balance denominator", "c",
     "double D = kappa * dT - OmT * delta;", "double D = kappa * dT + OmT * delta;", [COEF]),
    ("C_h times sqrt(C_d) for C_h / sqrt(C_d)", "c",
     "out[1] = Cd == 0.0 ? 0.0 : C[1] / sqrt(Cd) * dtheta;",
     "out[1] = Cd == 0.0 ? 0.0 : C[1] * sqrt(Cd) * dtheta;", [COEF]),
    ("drag polynomial without its linear term", "c",
     "(c[0] / U + c[1] + c[2] * U - c[3] * neo_pow_int(U, 6))", "(c[0] / U + c[1] + c[2] - c[3] * neo_pow_int(U, 6))",
     [COEF]),
    ("scalar profile in the drag height correction", "c",
     "double xim = sqrt(CdN) / kappa * (log(dh / h0) - psim);",
     "double xim = sqrt(CdN) / kappa * (log(dh / h0) - psih);", [COEF]),
    ("neutral drag for the previous drag in U_N10", "c",
     "double UN10 = dU / (1.0 + sqrt(Cdp) / kappa", "double UN10 = dU / (1.0 + sqrt(neo_neutral_drag(p, dU)) / kappa",
     [COEF]),
    ("sign of the absorbed longwave in the skin balance", "c",
     "double Qd = -(1.0 - R->alpha) * R->SW - R->eps * R->LW;",
     "double Qd = -(1.0 - R->alpha) * R->SW + R->eps * R->LW;", [COEF]),
    ("wind floor outside Large-Yeager", "c",
     "double Umin = ly ? p->minimum_wind_speed : 0.0;", "double Umin = p->minimum_wind_speed;", [COEF]),
    ("skin temperature change not clamped", "c",
     "double cl = x > maxdT ? maxdT : (x < -maxdT ? -maxdT : x);", "double cl = x;", [COEF]),
    ("Paulson scalar profile with the momentum factor", "c",
     "unstable = 2.0 * log((1.0 + z * z) / 2.0);", "unstable = 2.0 * log((1.0 + z) / 2.0);", [COEF]),
    ("T^3 for T^4 in the skin balance", "c",
     "double Iup = R->sigma * R->eps * neo_pow_int(Tm, 4);", "double Iup = R->sigma * R->eps * neo_pow_int(Tm, 3);",
     [COEF]),
    # net ocean fluxes, radiation, split Runge-Kutta
    ("sign of the freshwater flux", "c",
     "double Jwao = -SFao;", "double Jwao = SFao;", [NET]),
    ("momentum flux without 1/rho_o", "c",
     "double tx_ao = (rinv * (1.0 - aice) * V2(f->x_momentum, i, j) + rinv * (1.0 - aice) * rtx_w) / 2.0;",
     "double tx_ao = ((1.0 - aice) * V2(f->x_momentum, i, j) + (1.0 - aice) * rtx_w) / 2.0;", [NET]),
    ("face average with the east neighbour", "c",
     "int64_t cw = column(g, i - 1, j), cs = column(g, i, j - 1);\n"
     "            double rtx_w = cw >= 0 ? V2(f->x_momentum, wx(g, i - 1), j) : 0.0;",
     "int64_t cw = column(g, i + 1, j), cs = column(g, i, j - 1);\n"
     "            double rtx_w = cw >= 0 ? V2(f->x_momentum, wx(g, i + 1), j) : 0.0;", [NET]),
    ("freshwater density for the ocean reference density", "c",
     "double rinv = 1.0 / p->ocean_reference_density;\n    double cinv = 1.0 / p->ocean_heat_capacity;\n"
     "#pragma omp parallel for NEO_THREADS\n    for (int j = 0; j < g->Ny; j++)\n"
     "        for (int i = 0; i < g->Nx; i++) {\n            double aice = 0.0;\n            double To",
     "double rinv = 1.0 / p->freshwater_density;\n    double cinv = 1.0 / p->ocean_heat_capacity;\n"
     "#pragma omp parallel for NEO_THREADS\n    for (int j = 0; j < g->Ny; j++)\n"
     "        for (int i = 0; i < g->Nx; i++) {\n            double aice = 0.0;\n            double To", [NET]),
    ("freshwater heat content at the SST in kelvin", "c",
     "V2(n->J_H, i, j) = inactive ? 0.0 : To * Jwao;", "V2(n->J_H, i, j) = inactive ? 0.0 : (To + 273.15) * Jwao;",
     [NET]),
    ("sign of the emitted longwave", "c",
     "double SQ = Iup_ocean + Ila + Qss;", "double SQ = -Iup_ocean + Ila + Qss;", [NET]),
    ("albedo left out of the transmitted shortwave", "c",
     "double Itsw = -(1.0 - alpha) * SW;", "double Itsw = -SW;", [NET]),
    ("T^3 for T^4 in the emitted longwave", "c",
     "double Iup = sigma * eps * neo_pow_int(Ts, 4);", "double Iup = sigma * eps * neo_pow_int(Ts, 3);", [NET]),
    ("RK3 beta_2 = 1/3", "c",
     "const double beta[3] = {1.0 / 3.0, 0.5, 1.0};", "const double beta[3] = {1.0 / 3.0, 1.0 / 3.0, 1.0};", [NET]),
    ("RK3 beta_1 = 1/2", "c",
     "const double beta[3] = {1.0 / 3.0, 0.5, 1.0};", "const double beta[3] = {0.5, 0.5, 1.0};", [NET]),
    ("RK3 stage reads S at the start of the step in the surface salt flux", "c",
     "FS = V2(n->J_S, i, j) + ((V3(Sin, i, j, kN) * V2(n->J_w, i, j) - 0.0) + 0.0);",
     "FS = V2(n->J_S, i, j) + ((V3(S0, i, j, kN) * V2(n->J_w, i, j) - 0.0) + 0.0);", [NET]),
    # snapshot time indexing (oracle.py)
    ("cyclical time reduced from t = 0 instead of the first snapshot", "py",
     "tm = (t - t1) % period + t1", "tm = t % period", TSER_PINS),
    ("truncated instead of floored modulo", "py",
     "tm = (t - t1) % period + t1", "tm = float(np.fmod(t - t1, period)) + t1", TSER_PINS),
    ("wrap weight over the whole period", "py",
     "return n - 1, 0, (tm - tN) / (period - (tN - t1))", "return n - 1, 0, (tm - tN) / period", TSER_PINS),
    ("wrap snapshots in the wrong order", "py",
     "return n - 1, 0, (tm - tN) / (period - (tN - t1))", "return 0, n - 1, (tm - tN) / (period - (tN - t1))",
     TSER_PINS),
    ("next-cycle position of the wrapped snapshot", "py",
     "(cycle + (1 if n2 < n1 else 0)) * len(times) + n2", "cycle * len(times) + n2", TSER_PINS),
    ("cycle counted from the end of the first period", "py",
     "cycle = round((x - x % period) / period)", "cycle = round((x - x % period) / period) + (1 if x < 0 else 0)",
     TSER_PINS),
    ("QAB2 with chi = 0 (plain AB2)", "py",
     "chi = -0.5 if euler else abi.NE_AB2_CHI", "chi = -0.5 if euler else 0.0", [NET, "-k", "ab2"]),
    ("QAB2 without forward Euler on a changed step", "py",
     "euler = self.iteration == 0 or dt != self.last_dt", "euler = self.iteration == 0",
     [NET, "-k", "ab2"]),
    ("QAB2 weights that do not sum to one", "py",
     "a, b = 1.5 + chi, 0.5 + chi", "a, b = 1.5 + chi, 0.5 - chi", [NET, "-k", "ab2"]),
]


def _cc() -> str:
    return shutil.which(os.environ.get("CC", "gcc")) or "gcc"


def _build_c(src: str, out: Path) -> None:
    c = out.with_suffix(".c")
    c.write_text(src)
    cmd = [_cc(), "-O1", "-std=gnu11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           f"-I{REPO / 'include'}", f"-I{REPO / 'oracle'}", "-shared", "-o", str(out), str(c), "-lm"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def _run_mutant(idx: int, mutant, tmp: Path) -> tuple[str, bool, str]:
    name, kind, old, new, pins = mutant
    d = tmp / f"m{idx:02d}"
    d.mkdir()
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = "1"
    real_lib = REPO / "oracle" / "_build" / "libne_oracle.so"
    if kind == "c":
        src = ORACLE_C.read_text()
        assert src.count(old) == 1, f"{name}: the restatement's text is not found exactly once"
        lib = d / "libne_oracle.so"
        _build_c(src.replace(old, new), lib)
        env["NE_ORACLE_LIB"] = str(lib)
    else:
        src = ORACLE_PY.read_text()
        assert src.count(old) == 1, f"{name}: the restatement's text is not found exactly once"
        (d / "oracle.py").write_text(src.replace(old, new))
        env["NE_ORACLE_DIR"] = str(d)
        env["NE_ORACLE_LIB"] = str(real_lib)
    res = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                          *pins], cwd=REPO, env=env, capture_output=True, text=True, timeout=600)
    caught = res.returncode == 1          # a pin failed (2+ would be an error of the harness itself)
    return name, caught, (res.stdout + res.stderr)[-600:]


def test_every_planted_mistake_is_caught(tmp_path):
    import oracle as O
    O.lib()                                # the unmodified oracle is built
    workers = max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(lambda a: _run_mutant(a[0], a[1], tmp_path), enumerate(MUTANTS)))
    survivors = [(n, tail) for n, caught, tail in results if not caught]
    assert not survivors, "mutants not caught by any pin:\n" + "\n".join(f"- {n}\n{t}" for n, t in survivors)
    assert len(results) == len(MUTANTS) >= 40


def test_unmodified_oracle_passes_the_same_pins(tmp_path):
    """The harness itself: the unmodified restatement, built the way the mutants are, passes every pin
    the mutants are run against (so a caught mutant is caught by its mistake, not by the build)."""
    lib = tmp_path / "libne_oracle.so"
    _build_c(ORACLE_C.read_text(), lib)
    env = dict(os.environ, NE_ORACLE_LIB=str(lib), OMP_NUM_THREADS="1")
    for pins in ([GOLDEN, COEF, NET], TSER_PINS):
        res = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                              *pins], cwd=REPO, env=env, capture_output=True, text=True, timeout=600)
        assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
