"""The oracle's ring of the surface phases (the reference launches its interpolation and flux
kernels over 0:N+1 in every non-Flat direction, InterfaceComputations.jl:99-116), pinned by what
the grid fixes: in a periodic direction the ring node is the wrapped interior node (its fractional
index differs by exactly the source's period, so the interpolation and the flux solve reproduce the
wrapped column); beyond a bounded edge the ring cell is inactive (zero interface state) while the
atmosphere is still interpolated there (the bounded source holds its edge value); the views'
second halo layer is not written; and with no halo nothing outside the interior is."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from helpers import assert_rel
from numericalearth_hip import _ctypes_abi as abi
from oracle import EXCHANGE_NAMES, FLUX_NAMES
from ring_case import H, masks, oracle_ring, padded_np, ring_case


@pytest.mark.parametrize("topology", ["periodic_x", "bounded_x"])
def test_oracle_ring_conventions(topology):
    case, periodic = ring_case(topology)
    Nx, Ny, _ = case.shape
    om, out = oracle_ring(case, periodic)
    ring, inner, outside = masks(Ny, Nx)
    for name, a in out.items():
        if name == "iterations":
            assert np.all(a[outside] == -7)
            continue
        assert np.all(np.isnan(a[outside])), f"{name}: written beyond the ring"
        assert np.all(np.isfinite(a[ring | inner])), name
        # the interior is what the oracle computes without a halo
    om.interp()
    om.flux_solve()
    for k in EXCHANGE_NAMES:
        assert np.array_equal(out[f"exchange {k}"][inner].reshape(Ny, Nx), om.exchange_arrays[k]), k
    for k in FLUX_NAMES:
        assert np.array_equal(out[k][inner].reshape(Ny, Nx), om.flux_arrays[k]), k
    us = out["friction_velocity"]
    # rows j = −1 and Ny lie beyond the bounded latitude edges: zero interface state
    for j in (H - 1, H + Ny):
        assert np.all(us[j, H - 1:H + Nx + 1] == 0.0)
        assert np.all(out["sensible_heat"][j, H - 1:H + Nx + 1] == 0.0)
        assert np.all(out["interface_temperature"][j, H - 1:H + Nx + 1] == 0.0)
        # the atmosphere is interpolated there: the source holds its edge row
        assert np.all(np.isfinite(out["exchange T"][j, H - 1:H + Nx + 1]))
    west, east = (slice(H, H + Ny), H - 1), (slice(H, H + Ny), H + Nx)
    last, first = (slice(H, H + Ny), H + Nx - 1), (slice(H, H + Ny), H)
    if periodic:
        for name in [f"exchange {k}" for k in ("u", "v", "T", "p", "q")] + ["shortwave", "longwave"] + list(FLUX_NAMES):
            a = out[name]
            assert_rel(a[west], a[last], 1e-12, f"{name} west ring")
            assert_rel(a[east], a[first], 1e-12, f"{name} east ring")
        it = out["iterations"]
        assert np.array_equal(it[west], it[last]) and np.array_equal(it[east], it[first])
    else:
        for name in FLUX_NAMES:
            if name != "interface_temperature":
                assert np.all(out[name][west] == 0.0) and np.all(out[name][east] == 0.0), name


def test_oracle_without_halo_writes_the_interior_only():
    case, periodic = ring_case("periodic_x")
    Nx, Ny, _ = case.shape
    om = O.OracleModel(case)
    ex = {k: padded_np(Ny, Nx) for k in EXCHANGE_NAMES}
    exs = abi.ExchangeState(*(ex[k][1] for k in EXCHANGE_NAMES))
    g = abi.Grid.from_buffer_copy(om.grid)
    assert g.halo == 0
    O.lib().neo_interp_atmos(C.byref(g), C.byref(om.atmosphere), O._dp(om.afi), O._dp(om.afj), C.byref(exs))
    _, inner, _ = masks(Ny, Nx)
    for k in EXCHANGE_NAMES:
        a = ex[k][0]
        assert np.all(np.isfinite(a[inner]))
        assert np.all(np.isnan(a[~inner])), k
