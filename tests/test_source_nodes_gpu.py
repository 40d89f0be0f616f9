"""Source grids given by node coordinates on the GPU: ne_fractional_indices_nodes_f64 must match
the oracle bit for bit (a binary search there, a linear scan in the oracle, the same correctly
rounded arithmetic), and a model whose prescribed atmosphere sits on stretched nodes must step
like the oracle's to the FP64 tolerance of the other model tests."""
import ctypes as C
import dataclasses

import numpy as np
import pytest
import torch

from helpers import assert_rel, assert_rel_terms

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import NE_BOUNDED, NE_FLAT, NE_PERIODIC, jra55_source_grid, make_case
from numericalearth_hip.model import OceanOnlyModel, fractional_indices, stream_handle
from numericalearth_hip.params import to_params
from test_gpu_parity import RTOL, _compare_fluxes, _compare_fluxes_on_gpu_state, _net_term_scales, _np

pytestmark = pytest.mark.gpu


def _host_indices(case, spec):
    lam = np.ascontiguousarray(case.grid.lambda_c, dtype=np.float64)
    phi = np.ascontiguousarray(case.grid.phi_c, dtype=np.float64)
    g = abi.Grid()
    g.Nx, g.Ny, g.Nz = len(lam), len(phi), 1
    g.lambda_c, g.phi_c = lam.ctypes.data, phi.ctypes.data
    fi, fj = np.empty(len(lam)), np.empty(len(phi))
    O.fractional_indices(g, spec, fi, fj)
    return fi, fj


def _stretched(spec, rng):
    """The same source domain on nodes moved by up to a third of a cell."""
    lam, phi = spec.centres()
    lam = lam + rng.uniform(-1 / 3, 1 / 3, len(lam)) * spec.dlambda
    phi = phi + rng.uniform(-1 / 3, 1 / 3, len(phi)) * spec.dphi
    return dataclasses.replace(spec, lambda_nodes=lam, phi_nodes=phi)


@pytest.mark.parametrize("longitude", [(-180.0, 180.0), (0.0, 360.0), (20.0, 140.0)])
def test_node_indices_bit_exact_on_the_jra55_grid(require_gpu, longitude):
    case = make_case(1440, 560, 2, seed=3, longitude=longitude, latitude=(-79.5, 79.5))
    print(f"[nodes] jra55 grid, ocean longitude {longitude}", flush=True)
    m = OceanOnlyModel(case)
    spec = jra55_source_grid()
    fi = torch.empty(1440, dtype=torch.float64, device=m.device)
    fj = torch.empty(560, dtype=torch.float64, device=m.device)
    fractional_indices(m.ctx, m.grid, spec, fi, fj)
    torch.cuda.synchronize()
    hfi, hfj = _host_indices(case, spec)
    assert np.array_equal(_np(fi), hfi)
    assert np.array_equal(_np(fj), hfj)
    assert np.all((hfj > 0) & (hfj < 319))


def test_node_indices_reject_bad_arguments(require_gpu):
    case = make_case(16, 8, 2, seed=3)
    m = OceanOnlyModel(case)
    lam = torch.zeros(4, dtype=torch.float64, device=m.device)
    fi = torch.empty(16, dtype=torch.float64, device=m.device)
    fj = torch.empty(8, dtype=torch.float64, device=m.device)
    L, ctx, g = m.lib, m.ctx.handle, C.byref(m.grid.struct)
    p = lam.data_ptr()
    assert L.ne_fractional_indices_nodes_f64(ctx, g, p, 0, NE_PERIODIC, p, 4, NE_BOUNDED, fi.data_ptr(),
                                             fj.data_ptr(), stream_handle()) == abi.NE_ERR_SHAPE
    assert L.ne_fractional_indices_nodes_f64(ctx, g, p, 4, 7, p, 4, NE_BOUNDED, fi.data_ptr(), fj.data_ptr(),
                                             stream_handle()) == abi.NE_ERR_UNSUPPORTED
    assert L.ne_fractional_indices_nodes_f64(ctx, g, None, 4, NE_FLAT, p, 4, NE_BOUNDED, fi.data_ptr(),
                                             fj.data_ptr(), stream_handle()) == abi.NE_ERR_NULL


def test_model_on_stretched_source_nodes_matches_oracle(require_gpu):
    rng = np.random.default_rng(21)
    case = make_case(90, 40, 10, seed=13, n_times=2)
    case.atmos_grid = _stretched(case.atmos_grid, rng)
    case.rad_grid = _stretched(case.rad_grid, rng)
    print("[nodes] model on stretched source nodes, shape (90, 40, 10) seed 13", flush=True)
    p = to_params()
    om = O.OracleModel(case, params=abi.Params.from_buffer_copy(p))
    gm = OceanOnlyModel(case, params=abi.Params.from_buffer_copy(p))
    torch.cuda.synchronize()
    for a, b in ((gm.afi, om.afi), (gm.afj, om.afj), (gm.rfi, om.rfi), (gm.rfj, om.rfj)):
        assert np.array_equal(_np(a), b)
    # the stretched nodes move the indices away from the regular grid's
    reg = dataclasses.replace(case.atmos_grid, lambda_nodes=None, phi_nodes=None)
    assert not np.array_equal(_host_indices(case, reg)[0], om.afi)
    om.reconcile()
    torch.cuda.synchronize()
    _compare_fluxes(gm, om)
    for step in range(3):
        om.time_step(1200.0)
        gm.time_step(1200.0)
        torch.cuda.synchronize()
        print(f"[nodes] step {step + 1}", flush=True)
        assert_rel(_np(gm.T), om.T, RTOL, "T")
        assert_rel(_np(gm.S), om.S, RTOL, "S")
        _compare_fluxes_on_gpu_state(gm, om)
        scales = _net_term_scales(case, om)
        for k in O.NET_NAMES:
            assert_rel_terms(_np(gm.net[k]), om.net_arrays[k], scales.get(k, 0.0), RTOL, k)
