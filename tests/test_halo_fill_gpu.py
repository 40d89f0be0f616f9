"""ne_fill_halo_f64/_f32 (Oceananigans' fill_halo_regions! of a cell-centred tracer after a step,
ocean_simulation.jl:487-499) against numpy's own padding: a periodic direction padded with
mode="wrap", a bounded one and the vertical with mode="edge" (the zero-gradient halo of a flux
boundary condition), applied axis by axis as Oceananigans fills x, then y, then z; the interior
is left untouched."""
import ctypes as C

import numpy as np
import pytest
import torch

from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip._lib import check
from numericalearth_hip.generators import make_case
from numericalearth_hip.model import Context, DeviceGrid, stream_handle

pytestmark = pytest.mark.gpu


def _expected(interior, H, periodic_x, periodic_y):
    Hx, Hy, Hz = H
    a = np.pad(interior, ((0, 0), (0, 0), (Hx, Hx)), mode="wrap" if periodic_x else "edge")
    a = np.pad(a, ((0, 0), (Hy, Hy), (0, 0)), mode="wrap" if periodic_y else "edge")
    return np.pad(a, ((Hz, Hz), (0, 0), (0, 0)), mode="edge")


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("topology_x,H", [(abi.NE_PERIODIC, (7, 7, 7)), (abi.NE_BOUNDED, (3, 2, 1)),
                                          (abi.NE_PERIODIC, (12, 1, 0))])
def test_fill_halo_matches_numpy_padding(require_gpu, dtype, topology_x, H):
    lon = (0.0, 360.0) if topology_x == abi.NE_PERIODIC else (20.0, 80.0)
    case = make_case(10, 6, 4, seed=3, longitude=lon, topology_x=topology_x)
    Nx, Ny, Nz = case.shape
    Hx, Hy, Hz = H
    print(f"[halo] {Nx}x{Ny}x{Nz} topology_x {topology_x} halo {H} {dtype}", flush=True)
    ctx = Context.get(0)
    grid = DeviceGrid(case.grid, ctx, "cuda")
    g = torch.Generator(device="cpu").manual_seed(5)
    padded = torch.full((Nz + 2 * Hz, Ny + 2 * Hy, Nx + 2 * Hx), float("nan"), dtype=dtype)
    interior = torch.rand((Nz, Ny, Nx), generator=g, dtype=torch.float64).to(dtype)
    padded[Hz:Hz + Nz, Hy:Hy + Ny, Hx:Hx + Nx] = interior
    d = padded.cuda()
    es = d.element_size()
    origin = d.data_ptr() + (Hz * d.stride(0) + Hy * d.stride(1) + Hx) * es
    v = abi.View(origin, 1, d.stride(1), d.stride(0))
    fn = ctx.lib.ne_fill_halo_f64 if dtype == torch.float64 else ctx.lib.ne_fill_halo_f32
    check(fn(ctx.handle, C.byref(grid.struct), v, Hx, Hy, Hz, stream_handle()), "ne_fill_halo")
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    want = _expected(interior.numpy(), H, topology_x == abi.NE_PERIODIC, False)
    assert np.array_equal(got, want)


def test_fill_halo_refuses_bad_widths(require_gpu):
    case = make_case(8, 4, 3, seed=1)
    ctx = Context.get(0)
    grid = DeviceGrid(case.grid, ctx, "cuda")
    t = torch.zeros((3, 4, 8), dtype=torch.float64, device="cuda")
    v = abi.View(t.data_ptr(), 1, 8, 32)
    assert ctx.lib.ne_fill_halo_f64(ctx.handle, C.byref(grid.struct), v, -1, 0, 0, stream_handle()) == abi.NE_ERR_SHAPE
    assert ctx.lib.ne_fill_halo_f64(ctx.handle, C.byref(grid.struct), v, 65, 0, 0, stream_handle()) == abi.NE_ERR_SHAPE
    assert ctx.lib.ne_fill_halo_f64(ctx.handle, C.byref(grid.struct), v, 0, 0, 0, stream_handle()) == abi.NE_OK
