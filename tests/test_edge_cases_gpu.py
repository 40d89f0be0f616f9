"""Edge shapes of the coupled step against the oracle on the GPU: a single column, a single
level, two levels, ragged sizes that fill no tile, bounded longitude, and a domain that is
land everywhere, with the split RK3 and the Quasi-Adams-Bashforth 2 timesteppers. Same FP64 bar
as the other model tests (1e-12 relative)."""
import numpy as np
import pytest
import torch

from helpers import assert_rel, assert_rel_terms

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import NE_BOUNDED, NE_PERIODIC, make_case
from numericalearth_hip.model import OceanOnlyModel
from numericalearth_hip.params import to_params
from test_gpu_parity import RTOL, _compare_fluxes, _compare_fluxes_on_gpu_state, _net_term_scales, _np

pytestmark = pytest.mark.gpu

EDGE_CASES = [
    ("single column", (1, 1, 12), NE_PERIODIC, (0.0, 360.0)),
    ("one level", (7, 5, 1), NE_PERIODIC, (0.0, 360.0)),
    ("two levels, bounded", (9, 6, 2), NE_BOUNDED, (-40.0, 40.0)),
    ("ragged, bounded", (33, 17, 3), NE_BOUNDED, (100.0, 180.0)),
    ("one row", (70, 1, 6), NE_PERIODIC, (0.0, 360.0)),
]


def _step_and_compare(case, steps=3, dt=1200.0, timestepper="SplitRungeKutta3"):
    p = to_params()
    om = O.OracleModel(case, params=abi.Params.from_buffer_copy(p), timestepper=timestepper)
    gm = OceanOnlyModel(case, params=abi.Params.from_buffer_copy(p), timestepper=timestepper)
    om.reconcile()
    torch.cuda.synchronize()
    _compare_fluxes(gm, om)
    for step in range(steps):
        om.time_step(dt)
        gm.time_step(dt)
        torch.cuda.synchronize()
        print(f"[edge] step {step + 1}", flush=True)
        assert_rel(_np(gm.T), om.T, RTOL, "T")
        assert_rel(_np(gm.S), om.S, RTOL, "S")
        _compare_fluxes_on_gpu_state(gm, om)
        scales = _net_term_scales(case, om)
        for k in O.NET_NAMES:
            assert_rel_terms(_np(gm.net[k]), om.net_arrays[k], scales.get(k, 0.0), RTOL, k)
    return gm, om


@pytest.mark.parametrize("timestepper", ["SplitRungeKutta3", "QuasiAdamsBashforth2"])
@pytest.mark.parametrize("name,shape,topology_x,longitude", EDGE_CASES, ids=[c[0] for c in EDGE_CASES])
def test_edge_shapes_step_like_oracle(require_gpu, name, shape, topology_x, longitude, timestepper):
    case = make_case(*shape, seed=23, topology_x=topology_x, longitude=longitude)
    print(f"[edge] {name}: shape={shape} topology_x={topology_x} longitude={longitude} {timestepper}", flush=True)
    _step_and_compare(case, timestepper=timestepper)


def test_all_land_domain(require_gpu):
    case = make_case(40, 12, 4, seed=29)
    case.grid.bottom_height[...] = 1.0           # above every level: no active cell anywhere
    case.grid.kmin[...] = 4
    print("[edge] all land: shape=(40, 12, 4)", flush=True)
    T0, S0 = case.T.copy(), case.S.copy()
    gm, om = _step_and_compare(case, steps=2)
    assert np.array_equal(_np(gm.grid.kmin), case.grid.kmin)
    assert np.array_equal(_np(gm.T), T0) and np.array_equal(_np(gm.S), S0)   # nothing moves on land
