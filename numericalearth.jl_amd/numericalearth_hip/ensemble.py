"""Ensembles of coupled models stepped together on one MI355X (``ne_batch_*`` of the C ABI).

An ensemble binds existing :class:`~numericalearth_hip.model.OceanOnlyModel` instances of
one precision; ``Ensemble.time_step(dt)`` is ``time_step!(model, Δt)`` for every member,
issued as six persistent-grid launches for the whole ensemble (or one HIP graph launch),
with results identical to stepping each member on its own. The reference has no batched
form: calibration ensembles there are separate models stepped one after another.

``draw_ensemble`` is the seeded generator of uneven members: sizes log-uniform between a
single column (1×1×50, no advection) and a 360×140×20 regional grid (Bounded λ at 1/4°),
each with its own bathymetry, atmosphere and flux parameters (Charnock parameter
U(0.01, 0.02), roughness ±20 % as a common factor on the smooth-wall parameter and on the
Reynolds-scaling coefficient A of the scalar roughness).
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import math
from dataclasses import asdict, dataclass
from pathlib import Path

import numpy as np
import torch

from . import _ctypes_abi as abi
from ._lib import check
from .generators import NE_BOUNDED, make_case
from .model import OceanOnlyModel, stream_handle
from .params import default_params

TUNING_FIELDS = ("tile_rows", "stage_blocks", "surface_block", "surface_blocks", "stage_ms", "surface_ms")


@dataclass
class MemberSpec:
    Nx: int
    Ny: int
    Nz: int
    seed: int
    longitude: tuple
    latitude: tuple
    charnock: float
    roughness_factor: float

    @property
    def cells(self) -> int:
        return self.Nx * self.Ny * self.Nz


def draw_ensemble(n: int = 64, seed: int = 0, smallest=(1, 1, 50), largest=(360, 140, 20),
                  resolution: float = 0.25) -> list[MemberSpec]:
    """n members with sizes log-uniform between `smallest` and `largest` (one uniform draw s ∈ [0, 1]
    per member sets Nx, Ny and Nz together, N = N_small^(1−s)·N_large^s), placed at random inside
    λ ∈ [0, 360), φ ∈ [−70, 70] at `resolution` degrees."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        s = rng.uniform()
        Nx, Ny, Nz = (max(1, int(round(lo ** (1.0 - s) * hi ** s))) for lo, hi in zip(smallest, largest))
        span_x, span_y = Nx * resolution, Ny * resolution
        lon0 = rng.uniform(0.0, 360.0 - span_x)
        lat0 = rng.uniform(-70.0, 70.0 - span_y)
        out.append(MemberSpec(Nx, Ny, Nz, int(rng.integers(1 << 31)), (lon0, lon0 + span_x), (lat0, lat0 + span_y),
                              float(rng.uniform(0.01, 0.02)), float(rng.uniform(0.8, 1.2))))
    return out


def member_case(spec: MemberSpec):
    return make_case(spec.Nx, spec.Ny, spec.Nz, seed=spec.seed, longitude=spec.longitude, latitude=spec.latitude,
                     topology_x=NE_BOUNDED, sigma_cells=min(4.0, max(spec.Nx, spec.Ny) / 8.0))


def member_params(spec: MemberSpec) -> abi.Params:
    p = default_params()
    p.charnock_parameter = spec.charnock
    p.smooth_wall_parameter = p.smooth_wall_parameter * spec.roughness_factor
    p.reynolds_scaling_A = p.reynolds_scaling_A * spec.roughness_factor
    return p


def build_members(specs: list[MemberSpec], precision: str = "fp64", device: int = 0) -> list[OceanOnlyModel]:
    return [OceanOnlyModel(member_case(s), params=member_params(s), precision=precision, device=device)
            for s in specs]


class Ensemble:
    """A batch of models of one precision (``ne_batch``); members stay usable on their own after close()."""

    def __init__(self, models: list[OceanOnlyModel], stream=None, *, precision: str | None = None,
                 device: int = 0):
        from .model import Context
        precisions = {m.precision for m in models}
        if len(precisions) > 1:
            raise ValueError("ensemble members must share one precision")
        self.precision = precision or (models[0].precision if models else "fp64")
        if precisions and precisions != {self.precision}:
            raise ValueError("ensemble members must have the ensemble's precision")
        self.models = list(models)
        self.device = torch.device("cuda", models[0].device.index if models else device)
        self.ctx = models[0].ctx if models else Context.get(self.device.index)
        self.lib = self.ctx.lib
        code = abi.NE_FP64 if self.precision == "fp64" else abi.NE_MIXED_FP32
        arr = (C.c_void_p * max(1, len(models)))(*[m.handle.value for m in models])
        h = C.c_void_p()
        check(self.lib.ne_batch_create(self.ctx.handle, code, len(models), arr, C.byref(h), stream_handle(stream)),
              "ne_batch_create")
        self.handle = h

    def add(self, model: OceanOnlyModel) -> None:
        check(self.lib.ne_batch_add(self.handle, model.handle), "ne_batch_add")
        self.models.append(model)

    def remove(self, model: OceanOnlyModel) -> None:
        check(self.lib.ne_batch_remove(self.handle, model.handle), "ne_batch_remove")
        self.models.remove(model)

    def time_step(self, dt: float, stream=None) -> None:
        check(self.lib.ne_batch_time_step(self.handle, float(dt), stream_handle(stream)), "ne_batch_time_step")

    def use_graph(self, enabled: bool = True) -> None:
        check(self.lib.ne_batch_set_graph(self.handle, int(enabled)), "ne_batch_set_graph")

    @property
    def tuning(self) -> dict:
        t = abi.BatchTuning()
        check(self.lib.ne_batch_get_tuning(self.handle, C.byref(t)), "ne_batch_get_tuning")
        return {k: getattr(t, k) for k in TUNING_FIELDS}

    def set_tuning(self, tuning: dict) -> None:
        t = abi.BatchTuning(*(tuning.get(k, 0) for k in TUNING_FIELDS))
        check(self.lib.ne_batch_set_tuning(self.handle, C.byref(t)), "ne_batch_set_tuning")

    def signature(self) -> str:
        """Key of the autotuning cache: device name, precision and every member's shape."""
        shapes = [m.case.shape for m in self.models]
        name = torch.cuda.get_device_name(self.device)
        return hashlib.sha1(json.dumps([name, self.precision, shapes]).encode()).hexdigest()

    def autotune(self, cache: str | Path | None = None, stream=None) -> dict:
        """Time every candidate launch shape on this ensemble and keep the fastest; with `cache`, reuse
        (and record) the choice for an ensemble of the same shapes on the same device."""
        key = self.signature()
        store = {}
        if cache is not None and Path(cache).exists():
            store = json.loads(Path(cache).read_text())
            if key in store:
                self.set_tuning(store[key])
                return dict(store[key], cached=True)
        t = abi.BatchTuning()
        check(self.lib.ne_batch_autotune(self.handle, C.byref(t), stream_handle(stream)), "ne_batch_autotune")
        chosen = {k: getattr(t, k) for k in TUNING_FIELDS}
        if cache is not None:
            store[key] = chosen
            Path(cache).write_text(json.dumps(store, indent=1))
        return dict(chosen, cached=False)

    def info(self) -> dict:
        n, cells, si, su = C.c_int(), C.c_int64(), C.c_int(), C.c_int()
        check(self.lib.ne_batch_info(self.handle, C.byref(n), C.byref(cells), C.byref(si), C.byref(su)),
              "ne_batch_info")
        return {"members": n.value, "cells": cells.value, "stage_items": si.value, "surface_items": su.value}

    @property
    def cells(self) -> int:
        return sum(math.prod(m.case.shape) for m in self.models)

    def close(self) -> None:
        if getattr(self, "handle", None):
            check(self.lib.ne_batch_destroy(self.handle), "ne_batch_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def specs_json(specs: list[MemberSpec]) -> list[dict]:
    return [asdict(s) for s in specs]
