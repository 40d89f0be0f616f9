"""The allocation-free budget entry point (ne_budget_f64_noalloc) against ne_budget_f64.

It runs the same kernels with the block partials in the library's fixed device array instead of
the context's workspace, so the two results are equal bit for bit. Two calls in a row, the
second on a smaller grid, show that the shared array's contents from a larger call do not leak
into a smaller one (every block writes its own partials; nothing is read past the grid's
blocks). A grid with more blocks than the array holds is refused on the host.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import make_case
from numericalearth_hip.model import OceanOnlyModel, stream_handle, torch_view
import oracle as O

pytestmark = pytest.mark.gpu


def _budgets(gm, field, weight):
    lib, ctx, g = gm.lib, gm.ctx.handle, C.byref(gm.grid.struct)
    ref = torch.full((2,), np.nan, dtype=torch.float64, device=field.device)
    out = torch.full((2,), np.nan, dtype=torch.float64, device=field.device)
    s = stream_handle()
    assert lib.ne_budget_f64(ctx, g, torch_view(field), weight, ref.data_ptr(), s) == abi.NE_OK
    assert lib.ne_budget_f64_noalloc(ctx, g, torch_view(field), weight, out.data_ptr(), s) == abi.NE_OK
    torch.cuda.synchronize()
    return ref.cpu().numpy(), out.cpu().numpy()


def test_noalloc_budget_equals_allocating_budget(require_gpu):
    # the larger grid first (ceil(300/64)·ceil(130/4) = 165 blocks), then a smaller one (6 blocks)
    for n, (Nx, Ny, Nz, seed) in enumerate([(300, 130, 10, 3), (70, 18, 6, 4)]):
        print(f"[noalloc] call {n}: grid {Nx}x{Ny}x{Nz} seed {seed}", flush=True)
        case = make_case(Nx, Ny, Nz, seed=seed)
        gm = OceanOnlyModel(case)
        rho_c = gm.params.ocean_reference_density * gm.params.ocean_heat_capacity
        for field, weight in ((gm.T, rho_c), (gm.S, 1.0)):
            ref, out = _budgets(gm, field, weight)
            assert np.array_equal(ref.view(np.uint64), out.view(np.uint64)), (ref, out)
            om = O.OracleModel(case)
            Ho = om.budget(field.cpu().numpy(), weight)
            assert abs((out[0] + out[1]) - (Ho[0] + Ho[1])) <= 1e-15 * abs(Ho[0])
        gm.close()


def test_noalloc_budget_refuses_grids_beyond_its_array(require_gpu):
    case = make_case(40, 16, 4, seed=1)
    gm = OceanOnlyModel(case)
    big = abi.Grid.from_buffer_copy(gm.grid.struct)
    big.Nx, big.Ny = 64 * 512, 4 * 129          # 66 048 blocks > NE_BUDGET_NOALLOC_MAX_BLOCKS
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    st = gm.lib.ne_budget_f64_noalloc(gm.ctx.handle, C.byref(big), torch_view(gm.T), 1.0, out.data_ptr(),
                                      stream_handle())
    assert st == abi.NE_ERR_CAPACITY
    gm.close()
