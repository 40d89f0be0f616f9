"""The FP32 per-phase entry points (FP32 prescribed series and 3-D fields, FP64 exchange state and
2-D fluxes) against the oracle run in FP64 on the same FP32-rounded inputs.

* ne_interp_atmos_f32 / ne_interp_radiation_f32 form the interpolation in FP64 from the float
  series, with the oracle's operations: bit for bit.
* ne_tracer_tendency_weno5_f32 reconstructs in FP32: within 1e-5 of the terms summed into G.
* ne_tracer_rk_stage_f32 forms c0 + β·Δt·G in FP64 and rounds it to float to nearest: within one
  FP32 ulp of the oracle's FP64 stage plus β·Δt times the tendency bound above.
"""
import copy
import ctypes as C

import numpy as np
import pytest
import torch

import oracle as O
from helpers import assert_rel_terms
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import make_case
from numericalearth_hip.model import OceanOnlyModel, stream_handle, torch_view
from oracle import EXCHANGE_NAMES, np_view

pytestmark = pytest.mark.gpu
RTOL32 = 1e-5


def _rounded(case):
    """The case with every series and 3-D field rounded to FP32 (kept as FP64)."""
    r = copy.copy(case)
    f = lambda a: np.asarray(a, dtype=np.float32).astype(np.float64)   # noqa: E731
    r.T, r.S, r.u, r.v, r.w = f(case.T), f(case.S), f(case.u), f(case.v), f(case.w)
    r.atmos = {k: f(v) for k, v in case.atmos.items()}
    r.rad = {k: f(v) for k, v in case.rad.items()}
    return r


@pytest.mark.parametrize("topology_x", [abi.NE_PERIODIC, abi.NE_BOUNDED])
def test_f32_phases_match_oracle_on_rounded_inputs(require_gpu, topology_x):
    lon = (0.0, 360.0) if topology_x == abi.NE_PERIODIC else (20.0, 80.0)
    case = make_case(60, 32, 8, seed=17, longitude=lon, topology_x=topology_x)
    print(f"[f32] grid 60x32x8 seed 17 topology_x {topology_x}", flush=True)
    gm = OceanOnlyModel(case, precision="mixed")       # FP32 series and 3-D fields on the device
    om = O.OracleModel(_rounded(case)).reconcile()
    Nx, Ny, Nz = case.shape
    lib, ctx, s = gm.lib, gm.ctx.handle, stream_handle()
    g = C.byref(gm.grid.struct)

    # interpolation: FP64 exchange state from the float series, bit for bit
    ex = {k: torch.zeros((Ny, Nx), dtype=torch.float64, device="cuda") for k in EXCHANGE_NAMES}
    sw = torch.zeros((Ny, Nx), dtype=torch.float64, device="cuda")
    lw = torch.zeros_like(sw)
    exs = abi.ExchangeState(*(torch_view(ex[k]) for k in EXCHANGE_NAMES))
    assert lib.ne_interp_atmos_f32(ctx, g, C.byref(gm.desc.atmosphere), gm.afi.data_ptr(), gm.afj.data_ptr(),
                                   C.byref(exs), s) == abi.NE_OK
    assert lib.ne_interp_radiation_f32(ctx, g, C.byref(gm.desc.radiation), gm.rfi.data_ptr(), gm.rfj.data_ptr(),
                                       torch_view(sw), torch_view(lw), s) == abi.NE_OK
    torch.cuda.synchronize()
    om.interp()
    for k in ("u", "v", "T", "p", "q"):
        assert np.array_equal(ex[k].cpu().numpy(), om.exchange_arrays[k]), k
    assert np.array_equal(sw.cpu().numpy(), om.sw) and np.array_equal(lw.cpu().numpy(), om.lw)

    # tendency: the mixed model's state and the oracle's net fluxes
    for k in O.NET_NAMES:
        gm.net[k].copy_(torch.from_numpy(om.net_arrays[k]))
    GT = torch.zeros((Nz, Ny, Nx), dtype=torch.float32, device="cuda")
    GS = torch.zeros_like(GT)
    assert lib.ne_tracer_tendency_weno5_f32(ctx, g, C.byref(gm.params), C.byref(gm.desc.ocean),
                                            C.byref(gm.desc.net), torch_view(GT), torch_view(GS), s) == abi.NE_OK
    torch.cuda.synchronize()
    GT_o, GS_o, sT, sS = om.tendency(scales=True)
    eT = assert_rel_terms(GT.cpu().numpy(), GT_o, sT, RTOL32, "G_T")
    eS = assert_rel_terms(GS.cpu().numpy(), GS_o, sS, RTOL32, "G_S")
    print(f"[f32] tendency relative to its terms: G_T {eT:.2e} G_S {eS:.2e}", flush=True)
    assert max(eT, eS) > 0.0                               # FP32 arithmetic, not an FP64 path

    # one stage: c_out = c0 + β·Δt·G(c_in) with c0 = c_in = the state
    beta_dt = 400.0
    To = torch.empty_like(gm.T)
    So = torch.empty_like(gm.S)
    v = lambda t: torch_view(t)   # noqa: E731
    assert lib.ne_tracer_rk_stage_f32(ctx, g, C.byref(gm.params), v(gm.u), v(gm.v), v(gm.w), v(gm.T), v(gm.S),
                                      v(gm.T), v(gm.S), v(To), v(So), C.byref(gm.desc.net), beta_dt,
                                      s) == abi.NE_OK
    torch.cuda.synchronize()
    Tref = np.ascontiguousarray(om.T.copy())
    Sref = np.ascontiguousarray(om.S.copy())
    O.lib().neo_rk_stage(C.byref(om.grid), C.byref(om.params), np_view(om.u), np_view(om.v), np_view(om.w),
                         np_view(om.T), np_view(om.S), np_view(om.T), np_view(om.S), np_view(Tref), np_view(Sref),
                         C.byref(om.net), beta_dt)
    active = np.arange(Nz)[:, None, None] >= case.grid.kmin[None]
    for name, got, ref, scale in (("T", To, Tref, sT), ("S", So, Sref, sS)):
        d = np.abs(got.cpu().numpy().astype(np.float64) - ref)[active]
        bound = (np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64) + beta_dt * RTOL32 * scale)[active]
        assert np.all(d <= bound), f"{name}: stage differs by {np.max(d - bound):.3e} beyond its bound"
    gm.close()
