"""The per-phase surface entry points on halo-padded views: the interpolations and the flux solve
also write the ring i = -1, Nx and j = -1, Ny, as the reference launches them over 0:N+1 in every
non-Flat direction (InterfaceComputations.jl:99-116), so that its _assemble_net_ocean_fluxes!,
which averages ρτ with the west / south neighbour (assemble_net_ocean_fluxes.jl:136-137), reads
freshly computed values at i = 0 and j = 0.

The oracle computes the ring as the reference does (tests/ring_case.py): the fractional index of
a ring node from the halo node's coordinate, and the flux solve from the ocean surface in the
caller's halo, filled as Oceananigans fills it. The library extends the fractional indices from the
edge nodes and reads the wrapped interior column instead, so ring values agree up to the rounding
of the ring node's index; the interior is unchanged.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from helpers import assert_rel
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.model import OceanOnlyModel, stream_handle
from oracle import EXCHANGE_NAMES, FLUX_NAMES
from ring_case import H, masks, oracle_ring, ring_case

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _padded_torch(Ny, Nx, dtype=torch.float64, fill=float("nan")):
    t = torch.full((Ny + 2 * H, Nx + 2 * H), fill, dtype=dtype, device="cuda")
    v = abi.View(t.data_ptr() + (H * (Nx + 2 * H) + H) * t.element_size(), 1, Nx + 2 * H, 0)
    return t, v


@pytest.mark.parametrize("topology", ["periodic_x", "bounded_x"])
def test_ring_of_interpolation_and_flux_solve(require_gpu, topology):
    case, periodic = ring_case(topology)
    Nx, Ny, Nz = case.shape
    print(f"[ring] {topology} grid {Nx}x{Ny}x{Nz} halo {H}", flush=True)
    gm = OceanOnlyModel(case)
    lib, ctx, s = gm.lib, gm.ctx.handle, stream_handle()
    grid = abi.Grid.from_buffer_copy(gm.grid.struct)
    grid.halo = H
    dex = {k: _padded_torch(Ny, Nx) for k in EXCHANGE_NAMES}
    dsw, dlw = _padded_torch(Ny, Nx), _padded_torch(Ny, Nx)
    ex = abi.ExchangeState(*(dex[k][1] for k in EXCHANGE_NAMES))
    ex.downwelling_shortwave, ex.downwelling_longwave = dsw[1], dlw[1]
    dfl = {k: _padded_torch(Ny, Nx) for k in FLUX_NAMES}
    dit = _padded_torch(Ny, Nx, torch.int32, -7)
    fl = abi.InterfaceFluxes(*(dfl[k][1] for k in FLUX_NAMES), dit[1])
    atm = abi.PrescribedAtmosphereDesc.from_buffer_copy(gm.desc.atmosphere)
    rad = abi.PrescribedRadiationDesc.from_buffer_copy(gm.desc.radiation)
    assert lib.ne_interp_atmos_f64(ctx, C.byref(grid), C.byref(atm), gm.afi.data_ptr(), gm.afj.data_ptr(),
                                   C.byref(ex), s) == abi.NE_OK
    assert lib.ne_interp_radiation_f64(ctx, C.byref(grid), C.byref(rad), gm.rfi.data_ptr(), gm.rfj.data_ptr(),
                                       dsw[1], dlw[1], s) == abi.NE_OK
    assert lib.ne_ao_flux_solve_f64(ctx, C.byref(grid), C.byref(gm.params), C.byref(ex), C.byref(gm.desc.ocean),
                                    C.byref(fl), s) == abi.NE_OK
    torch.cuda.synchronize()

    _, ref = oracle_ring(case, periodic)
    ring, inner, outside = masks(Ny, Nx)
    dev = {f"exchange {k}": dex[k][0] for k in EXCHANGE_NAMES}
    dev["shortwave"], dev["longwave"] = dsw[0], dlw[0]
    dev.update({k: dfl[k][0] for k in FLUX_NAMES})
    for name, t in dev.items():
        d = t.cpu().numpy()
        assert np.all(np.isnan(d[outside])), f"{name}: written beyond the ring"
        assert np.all(np.isfinite(d[ring])), f"{name}: ring not written"
        if name in FLUX_NAMES:
            assert_rel(d[inner], ref[name][inner], RTOL, f"{name} interior")
        else:   # the interpolation: the oracle's arithmetic, bit for bit
            assert np.array_equal(d[inner], ref[name][inner]), f"{name}: interior differs"
        assert_rel(d[ring], ref[name][ring], RTOL, f"{name} ring")
    it = dit[0].cpu().numpy()
    assert np.all(it[outside] == -7)
    assert np.array_equal(it[ring | inner], ref["iterations"][ring | inner])
    gm.close()


def test_no_ring_without_halo(require_gpu):
    """halo = 0 (views of the interior only): nothing outside the interior is written."""
    case, _ = ring_case("periodic_x")
    gm = OceanOnlyModel(case)
    Nx, Ny, _ = case.shape
    lib, ctx, s = gm.lib, gm.ctx.handle, stream_handle()
    grid = abi.Grid.from_buffer_copy(gm.grid.struct)
    assert grid.halo == 0
    dex = {k: _padded_torch(Ny, Nx) for k in EXCHANGE_NAMES}
    ex = abi.ExchangeState(*(dex[k][1] for k in EXCHANGE_NAMES))
    dfl = {k: _padded_torch(Ny, Nx) for k in FLUX_NAMES}
    fl = abi.InterfaceFluxes(*(dfl[k][1] for k in FLUX_NAMES), abi.View(None, 0, 0, 0))
    atm = abi.PrescribedAtmosphereDesc.from_buffer_copy(gm.desc.atmosphere)
    assert lib.ne_interp_atmos_f64(ctx, C.byref(grid), C.byref(atm), gm.afi.data_ptr(), gm.afj.data_ptr(),
                                   C.byref(ex), s) == abi.NE_OK
    assert lib.ne_ao_flux_solve_f64(ctx, C.byref(grid), C.byref(gm.params), C.byref(ex), C.byref(gm.desc.ocean),
                                    C.byref(fl), s) == abi.NE_OK
    torch.cuda.synchronize()
    _, inner, _ = masks(Ny, Nx)
    for name, (t, _) in list(dex.items()) + list(dfl.items()):
        d = t.cpu().numpy()
        assert np.all(np.isfinite(d[inner])), name
        assert np.all(np.isnan(d[~inner])), f"{name}: written outside the interior"
    gm.close()
