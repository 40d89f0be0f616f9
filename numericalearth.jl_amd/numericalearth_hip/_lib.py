"""Loader of libnumericalearth_hip.so and the ctypes signatures of its C ABI.

There is no fallback: if the library is missing or fails to load, every entry
point raises. Build it with ``python -m numericalearth_hip.build`` (or the
repository's ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

from . import _ctypes_abi as abi
from .build import library_path

P = C.POINTER
vp = C.c_void_p


class NEError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what} failed: {abi.STATUS_NAMES.get(status, status)}")
        self.status = status


def check(status: int, what: str) -> None:
    if status != abi.NE_OK:
        raise NEError(status, what)


# name → argtypes (restype is int for every entry point)
SIGNATURES = {
    "ne_version": [],
    "ne_params_default": [P(abi.Params)],
    "ne_ctx_create": [C.c_int, P(vp)],
    "ne_ctx_destroy": [vp],
    "ne_ctx_counters": [vp, P(abi.Counters), C.c_int, vp],
    "ne_grid_immersed_levels": [vp, P(abi.Grid), abi.View, vp, vp],
    "ne_fractional_indices_f64": [vp, P(abi.Grid), P(abi.SourceGrid), vp, vp, vp],
    "ne_log_f64": [vp, vp, vp, C.c_int64, vp],
    "ne_atan_ge1_f64": [vp, vp, vp, C.c_int64, vp],
    "ne_cbrt_ge1_f64": [vp, vp, vp, C.c_int64, vp],
    "ne_sqrt_f64": [vp, vp, vp, C.c_int64, vp],
    "ne_exp_f64": [vp, vp, vp, C.c_int64, vp],
    "ne_stream_probe_f64": [vp, vp, C.c_int32, vp, C.c_int32, C.c_int64, vp],
    "ne_fill_halo_f64": [vp, vp, abi.View, C.c_int32, C.c_int32, C.c_int32, vp],
    "ne_fill_halo_f32": [vp, vp, abi.View, C.c_int32, C.c_int32, C.c_int32, vp],
    "ne_fractional_indices_nodes_f64": [vp, P(abi.Grid), vp, C.c_int32, C.c_int32, vp, C.c_int32, C.c_int32, vp, vp,
                                        vp],
    "ne_interp_atmos_f64": [vp, P(abi.Grid), P(abi.PrescribedAtmosphereDesc), vp, vp, P(abi.ExchangeState), vp],
    "ne_interp_radiation_f64": [vp, P(abi.Grid), P(abi.PrescribedRadiationDesc), vp, vp, abi.View, abi.View, vp],
    "ne_ao_flux_solve_f64": [vp, P(abi.Grid), P(abi.Params), P(abi.ExchangeState), P(abi.OceanState),
                             P(abi.InterfaceFluxes), vp],
    "ne_freeze_clamp_f64": [vp, P(abi.Grid), P(abi.Params), abi.View, abi.View, abi.View, C.c_double, vp],
    "ne_assemble_net_ocean_fluxes_f64": [vp, P(abi.Grid), P(abi.Params), P(abi.InterfaceFluxes),
                                         P(abi.ExchangeState), abi.View, P(abi.NetOceanFluxes), vp],
    "ne_apply_air_sea_radiation_f64": [vp, P(abi.Grid), P(abi.Params), abi.View, abi.View, abi.View,
                                       P(abi.NetOceanFluxes), P(abi.RadiativeFluxes), vp],
    "ne_tracer_tendency_weno5_f64": [vp, P(abi.Grid), P(abi.Params), P(abi.OceanState), P(abi.NetOceanFluxes),
                                     abi.View, abi.View, vp],
    "ne_tracer_ab2_update_f64": [vp, P(abi.Grid)] + [abi.View] * 6 + [C.c_double, C.c_double, C.c_int32, vp],
    "ne_tracer_rk_stage_f64": [vp, P(abi.Grid), P(abi.Params)] + [abi.View] * 9 + [P(abi.NetOceanFluxes),
                                                                                 C.c_double, vp],
    "ne_budget_f64": [vp, P(abi.Grid), abi.View, C.c_double, vp, vp],
    "ne_budget_f32": [vp, P(abi.Grid), abi.View, C.c_double, vp, vp],
    "ne_budget_f64_noalloc": [vp, P(abi.Grid), abi.View, C.c_double, vp, vp],
    "ne_model_create": [vp, P(abi.ModelDesc), P(vp), vp],
    "ne_model_destroy": [vp],
    "ne_model_time_step": [vp, C.c_double, vp],
    "ne_model_update_state": [vp, vp],
    "ne_model_clock": [vp, P(C.c_double), P(C.c_int64)],
    "ne_model_rounding_residual": [vp, vp, vp],
    "ne_model_timing_begin": [vp, C.c_int],
    "ne_model_timing_end": [vp, P(C.c_double), P(C.c_int)],
    "ne_model_workspace_bytes": [P(abi.ModelDesc), P(C.c_int64)],
    "ne_model_update_vertical_velocity": [vp, vp],
    "ne_snapshot_time_indices": [P(C.c_double), C.c_int32, C.c_int32, C.c_double, C.c_double,
                                 P(abi.TimeInterp), P(C.c_int64)],
    "ne_model_snapshots_at": [vp, C.c_double, P(C.c_int64)],
    "ne_model_snapshot_resident": [vp, C.c_int32, C.c_int64],
    "ne_compute_w_from_continuity_f64": [vp, P(abi.Grid), abi.View, abi.View, abi.View, vp],
    "ne_interp_atmos_f32": [vp, P(abi.Grid), P(abi.PrescribedAtmosphereDesc), vp, vp, P(abi.ExchangeState), vp],
    "ne_interp_radiation_f32": [vp, P(abi.Grid), P(abi.PrescribedRadiationDesc), vp, vp, abi.View, abi.View, vp],
    "ne_tracer_tendency_weno5_f32": [vp, P(abi.Grid), P(abi.Params), P(abi.OceanState), P(abi.NetOceanFluxes),
                                     abi.View, abi.View, vp],
    "ne_tracer_rk_stage_f32": [vp, P(abi.Grid), P(abi.Params)] + [abi.View] * 9 + [P(abi.NetOceanFluxes),
                                                                                 C.c_double, vp],
    "ne_batch_create": [vp, C.c_int32, C.c_int, P(vp), P(vp), vp],
    "ne_batch_destroy": [vp],
    "ne_batch_add": [vp, vp],
    "ne_batch_remove": [vp, vp],
    "ne_batch_time_step": [vp, C.c_double, vp],
    "ne_batch_set_graph": [vp, C.c_int],
    "ne_batch_autotune": [vp, P(abi.BatchTuning), vp],
    "ne_batch_get_tuning": [vp, P(abi.BatchTuning)],
    "ne_batch_set_tuning": [vp, P(abi.BatchTuning)],
    "ne_batch_info": [vp, P(C.c_int), P(C.c_int64), P(C.c_int), P(C.c_int)],
}

TIMED_KERNELS = ("rk_stage_1", "rk_stage_2", "rk_stage_3", "surface_update", "surface_tau")

_lib = None


def load(path: str | Path | None = None):
    """Load the HIP library (raises if it is missing: there is no CPU fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    env = os.environ.get("NE_HIP_LIBRARY")
    p = Path(path) if path else (Path(env) if env else library_path())
    if not p.exists():
        raise RuntimeError(f"HIP library not found at {p}; build it with `python -m numericalearth_hip.build`")
    L = C.CDLL(str(p))
    for name, args in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    if L.ne_version() != 1:
        raise RuntimeError("libnumericalearth_hip ABI version mismatch")
    if path is None:
        _lib = L
    return L
