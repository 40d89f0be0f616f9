"""Fractional indices in source grids given by node coordinates (unequally spaced latitudes
such as those of JRA55's grid), on the CPU.

The oracle's node-based indices are pinned three ways: on equally spaced nodes they are the
regular grid's indices (prescribed_atmosphere_regridder.jl:51-71); the bilinear value they
give is numpy's own piecewise-linear interpolation (np.interp, with ``period=360`` across the
longitude wrap; clamped at bounded edges as the reference's zero-gradient halo holds the edge
value); and at a node the index is that node's integer.
"""
import numpy as np
import pytest

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import (NE_BOUNDED, NE_FLAT, NE_PERIODIC, SourceSpec, gaussian_latitude_bounds,
                                           gaussian_latitudes, jra55_source_grid, source_grid_from_interfaces)


def _grid(lam, phi):
    lam = np.ascontiguousarray(lam, dtype=np.float64)
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    g = abi.Grid()
    g.Nx, g.Ny, g.Nz = len(lam), len(phi), 1
    g.lambda_c, g.phi_c = lam.ctypes.data, phi.ctypes.data
    return g, (lam, phi)                 # keep the arrays alive beside the struct


def _indices(spec, lam, phi):
    g, keep = _grid(lam, phi)
    fi, fj = np.empty(len(lam)), np.empty(len(phi))
    O.fractional_indices(g, spec, fi, fj)
    return fi, fj


def _value(f, frac, topo):
    """The interpolation's 1-D weights applied to f (neighbours: periodic wrap, or the edge
    value held beyond a bounded edge)."""
    n = len(f)
    lo = np.floor(frac).astype(int)
    w = frac - lo
    hi = lo + 1
    if topo == NE_PERIODIC:
        lo, hi = lo % n, hi % n
    else:
        lo, hi = np.clip(lo, 0, n - 1), np.clip(hi, 0, n - 1)
    return (1.0 - w) * f[lo] + w * f[hi]


def _nodes(rng, n, lo, hi):
    d = rng.uniform(0.3, 1.7, n)
    x = np.cumsum(d)
    return lo + (x - x[0]) * (hi - lo) / (x[-1] - x[0])


@pytest.mark.parametrize("topo_x", [NE_PERIODIC, NE_BOUNDED])
def test_equally_spaced_nodes_give_the_regular_indices(topo_x):
    rng = np.random.default_rng(3)
    reg = SourceSpec(24, 18, topo_x, NE_BOUNDED, -20.0, 15.0, -72.0, 8.0)
    lam_n, phi_n = reg.centres()
    nodes = SourceSpec(24, 18, topo_x, NE_BOUNDED, 0.0, 0.0, 0.0, 0.0, lambda_nodes=lam_n, phi_nodes=phi_n)
    lo = -20.0 if topo_x == NE_BOUNDED else -300.0          # periodic: ocean longitudes outside the record's range
    lam = np.concatenate([rng.uniform(lo, 400.0, 500), lam_n])
    phi = np.concatenate([rng.uniform(-85.0, 85.0, 300), phi_n])
    fi_r, fj_r = _indices(reg, lam, phi)
    fi_n, fj_n = _indices(nodes, lam, phi)
    assert np.all(np.abs(fj_n - fj_r) <= 1e-12 * 18)
    if topo_x == NE_PERIODIC:                                # the same point on the circle: indices agree mod Nx
        d = (fi_n - fi_r) % 24
        d = np.minimum(d, 24 - d)
        assert np.all(d <= 1e-11)
    else:
        inside = (lam >= lam_n[0]) & (lam <= lam_n[-1])
        assert np.all(np.abs(fi_n - fi_r)[inside] <= 1e-12 * 24)
    # the values read through either set of indices agree everywhere
    f = rng.standard_normal(24)
    assert np.allclose(_value(f, fi_n, topo_x), _value(f, fi_r, topo_x), atol=1e-12)


def test_stretched_nodes_interpolate_like_numpy():
    rng = np.random.default_rng(4)
    lam_n = np.sort(rng.uniform(0.0, 360.0, 30))
    phi_n = _nodes(rng, 21, -80.0, 84.0)
    spec = SourceSpec(30, 21, NE_PERIODIC, NE_BOUNDED, 0.0, 0.0, 0.0, 0.0, lambda_nodes=lam_n, phi_nodes=phi_n)
    lam = np.concatenate([rng.uniform(-540.0, 720.0, 2000), lam_n, lam_n + 360.0, lam_n - 720.0])
    phi = np.concatenate([rng.uniform(-90.0, 90.0, 2000), phi_n, [phi_n[0] - 3.0, phi_n[-1] + 3.0]])
    fi, fj = _indices(spec, lam, phi)
    f, g = rng.standard_normal(30), rng.standard_normal(21)
    assert np.allclose(_value(f, fi, NE_PERIODIC), np.interp(lam, lam_n, f, period=360.0), atol=1e-12)
    assert np.allclose(_value(g, fj, NE_BOUNDED), np.interp(phi, phi_n, g), atol=1e-12)
    # at a node: its integer index exactly; a turn away, the same up to the rounding of λ ± 360
    assert np.array_equal(fj[2000:2021], np.arange(21.0))
    assert np.array_equal(fi[2000:2030], np.arange(30.0))
    for shifted in (fi[2030:2060], fi[2060:2090]):
        d = (shifted - np.arange(30.0)) % 30
        assert np.all(np.minimum(d, 30 - d) <= 1e-11)
    # bounded edges continue the first / last interval
    assert fj[-2] == pytest.approx(-3.0 / (phi_n[1] - phi_n[0]))
    assert fj[-1] == pytest.approx(20.0 + 3.0 / (phi_n[-1] - phi_n[-2]))


def test_flat_and_single_node_axes():
    spec = SourceSpec(1, 1, NE_FLAT, NE_FLAT, 0.0, 0.0, 0.0, 0.0, lambda_nodes=np.array([5.0]),
                      phi_nodes=np.array([1.0]))
    fi, fj = _indices(spec, [0.0, 200.0], [-30.0, 60.0])
    assert np.array_equal(fi, [0.0, 0.0]) and np.array_equal(fj, [0.0, 0.0])


def test_jra55_gaussian_grid():
    lat = gaussian_latitudes(320)
    assert len(lat) == 320 and np.all(np.diff(lat) > 0)
    assert np.allclose(lat, -lat[::-1], atol=1e-12)            # symmetric about the equator
    # Gauss–Legendre: the sines of the latitudes integrate polynomials of degree < 640 exactly
    x, w = np.polynomial.legendre.leggauss(320)
    assert np.allclose(np.sin(np.radians(lat)), x, atol=1e-14)
    assert np.sum(w * x ** 2) == pytest.approx(2.0 / 3.0, rel=1e-13)
    # the Gauss-weight cells: each brackets its latitude and holds its weight's area
    b = gaussian_latitude_bounds(320)
    assert b[0] == -90.0 and b[-1] == 90.0 and np.all((b[:-1] < lat) & (lat < b[1:]))
    assert np.allclose(np.diff(np.sin(np.radians(b))), w, atol=1e-13)
    # the grid the reference builds: interfaces closed by +360 and 90°, centres their midpoints
    spec = jra55_source_grid()
    lam_c, phi_c = spec.centres()
    assert spec.irregular and (spec.Nx, spec.Ny) == (640, 320)
    assert np.allclose(lam_c, 0.5625 * np.arange(640), atol=1e-12)
    assert np.allclose(phi_c, 0.5 * (b[:-1] + b[1:]), atol=1e-13)
    assert np.abs(phi_c - lat).max() < 0.1 and np.ptp(np.diff(phi_c)) > 1e-4    # near, not on, the Gaussian latitudes
    rng = np.random.default_rng(8)
    lam = np.arange(1440) * 0.25 + 0.125 - 180.0                # a quarter-degree ocean grid from 180°W
    phi = np.linspace(-79.9, 79.9, 560)
    fi, fj = _indices(spec, lam, phi)
    f = rng.standard_normal(320)
    assert np.allclose(_value(f, fj, NE_BOUNDED), np.interp(phi, phi_c, f), atol=1e-12)
    assert np.allclose(fi % 640, ((lam % 360.0) / 0.5625), atol=1e-10)


def test_source_grid_from_interfaces():
    lf = np.array([10.0, 11.0, 13.0, 16.0])
    pf = np.array([-30.0, -10.0, 0.0, 5.0, 50.0])
    s = source_grid_from_interfaces(lf, pf, NE_BOUNDED, NE_BOUNDED)
    assert (s.Nx, s.Ny) == (3, 4)
    lam_c, phi_c = s.centres()
    assert np.array_equal(lam_c, [10.5, 12.0, 14.5]) and np.array_equal(phi_c, [-20.0, -5.0, 2.5, 27.5])
    # the user's own lower bounds go through jra55_source_grid's closure
    g = jra55_source_grid(lat_bounds=np.linspace(-90.0, 89.4375, 320), lon_bounds=np.arange(640) * 0.5625)
    assert g.centres()[0][-1] == 359.71875 and g.centres()[1][-1] == pytest.approx(89.71875)
    for bad in ([1.0], [2.0, 1.0]):
        with pytest.raises(ValueError):
            source_grid_from_interfaces(bad, pf)
