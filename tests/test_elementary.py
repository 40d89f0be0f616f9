"""The polynomial of the flux solve's arctangent (ne_atan_poly.h, written by gen_atan_poly.py):
the header is the generator's output, and with its coefficients as written,
u + u·w·P(w) (w = u²) is within 0.07 ulp of mpmath's arctangent, at 60 digits, on random points
of |u| ≤ tan(π/8), the interval the reduction leaves, and at its ends. The check parses the
header's literals, so a coefficient edited or mistyped in it fails here."""
import pathlib
import re
import subprocess
import sys

import mpmath as mp
import numpy as np

CSRC = pathlib.Path(__file__).resolve().parents[1] / "numericalearth.jl_amd" / "csrc"


def _coefficients():
    text = (CSRC / "ne_atan_poly.h").read_text()
    lits = [float(x) for x in re.findall(r"scalar_const\(([-+0-9.e]+)\)", text)]
    return lits[::-1]        # Horner order in the header: highest degree first


def test_header_is_the_generator_output():
    out = subprocess.run([sys.executable, str(CSRC / "gen_atan_poly.py")], capture_output=True, text=True,
                         check=True).stdout
    assert out == (CSRC / "ne_atan_poly.h").read_text()


def test_polynomial_is_within_bound_of_atan():
    c = _coefficients()
    assert len(c) == 11 and c[0] == float(mp.mpf(-1) / 3)
    mp.mp.dps = 60
    edge = mp.tan(mp.pi / 8)
    rng = np.random.default_rng(5)
    us = [mp.mpf(float(x)) for x in rng.uniform(-float(edge), float(edge), 3000)] + [edge, -edge, mp.mpf(1e-9)]
    worst = mp.mpf(0)
    for u in us:
        w = u * u
        p = mp.mpf(0)
        for x in reversed(c):
            p = p * w + x
        worst = max(worst, abs((u + u * w * p) / mp.atan(u) - 1))
    assert worst <= 0.07 * mp.mpf(2) ** -53, float(worst / mp.mpf(2) ** -53)
