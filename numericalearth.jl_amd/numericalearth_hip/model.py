"""Host-side mirror of the reference's coupled model over the C ABI.

``OceanOnlyModel`` mirrors NumericalEarth's ``OceanOnlyModel(ocean; atmosphere, radiation)``
(earth_system_model.jl:399-402) and ``time_step!`` (time_step_earth_system_model.jl:14-36):
the state lives in device memory (torch is used only to allocate it and to name the
stream), and every numerical operation is a call into libnumericalearth_hip.so.
``Phases`` exposes the per-phase entry points that replace the reference's kernel
launches one by one.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _ctypes_abi as abi
from ._lib import check, load
from .generators import SurfaceCase
from .params import default_params
from .time_series import ArraySnapshots, SnapshotStream, initial_window, resolve_time_indexing

FLUX_NAMES = ("latent_heat", "sensible_heat", "water_vapor", "x_momentum", "y_momentum",
              "friction_velocity", "temperature_scale", "water_vapor_scale", "interface_temperature")
NET_NAMES = ("tau_x", "tau_y", "J_T", "J_S", "J_w", "J_H")
EXCHANGE_NAMES = ("u", "v", "T", "p", "q", "rain", "snow")
RAD_NAMES = ("upwelling_longwave", "downwelling_longwave", "downwelling_shortwave")


def torch_view(t: torch.Tensor | None) -> abi.View:
    """ne_view of a device tensor whose last dimension is i (element strides)."""
    if t is None:
        return abi.View(None, 0, 0, 0)
    st = t.stride()
    sx = st[-1] if t.dim() >= 1 else 0
    sy = st[-2] if t.dim() >= 2 else 0
    sz = st[-3] if t.dim() >= 3 else 0
    return abi.View(t.data_ptr(), sx, sy, sz)


def stream_handle(stream: torch.cuda.Stream | None = None) -> int | None:
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream or None


class Context:
    """One ne_ctx per device (owns the solver counters and the budget workspace)."""

    _contexts: dict = {}

    def __init__(self, device: int = 0):
        self.lib = load()
        self.device = device
        h = C.c_void_p()
        check(self.lib.ne_ctx_create(device, C.byref(h)), "ne_ctx_create")
        self.handle = h

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        if device not in cls._contexts:
            cls._contexts[device] = Context(device)
        return cls._contexts[device]

    def counters(self, reset: bool = False, stream=None) -> dict:
        c = abi.Counters()
        check(self.lib.ne_ctx_counters(self.handle, C.byref(c), int(reset), stream_handle(stream)), "ne_ctx_counters")
        return {name: getattr(c, name) for name, _ in abi.Counters._fields_}


def _dev(a: np.ndarray, device, dtype=None) -> torch.Tensor:
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


class DeviceGrid:
    """Immersed LatitudeLongitudeGrid on the device: per-row metrics and the first active level of every
    column, computed on the device from the bottom height (GridFittedBottom)."""

    def __init__(self, grid, ctx: Context, device):
        spec = grid.spec
        self.spec = spec
        self.Nx, self.Ny, self.Nz = spec.Nx, spec.Ny, spec.Nz
        f64 = torch.float64
        self.dz = _dev(grid.dz, device, f64)
        self.zc = _dev(grid.z_center, device, f64)
        self.dyx = _dev(grid.dy_xface, device, f64)
        self.dxy = _dev(grid.dx_yface, device, f64)
        self.Az = _dev(grid.Az, device, f64)
        self.lam = _dev(grid.lambda_c, device, f64)
        self.phi = _dev(grid.phi_c, device, f64)
        self.bottom = _dev(grid.bottom_height, device, f64)
        self.zf = _dev(grid.z_face, device, f64)
        self.kmin = torch.empty((self.Ny, self.Nx), dtype=torch.int32, device=device)
        self.struct = abi.Grid(self.Nx, self.Ny, self.Nz, spec.topology_x, spec.topology_y, spec.advection,
                               self.dz.data_ptr(), self.zc.data_ptr(), self.dyx.data_ptr(), self.dxy.data_ptr(),
                               self.Az.data_ptr(), self.lam.data_ptr(), self.phi.data_ptr(), self.kmin.data_ptr(),
                               self.zf.data_ptr())
        check(ctx.lib.ne_grid_immersed_levels(ctx.handle, C.byref(self.struct), torch_view(self.bottom),
                                              self.kmin.data_ptr(), stream_handle()), "ne_grid_immersed_levels")


def fractional_indices(ctx, grid, spec, fi: torch.Tensor, fj: torch.Tensor) -> None:
    """Fractional indices of the ocean grid's cell centres in a source grid
    (prescribed_atmosphere_regridder.jl:51-71): from its regular spacing, or from its node
    coordinates when it has them (an unequally spaced grid such as JRA55's Gaussian latitudes)."""
    g = C.byref(grid.struct)
    s = stream_handle()
    if not spec.irregular:
        check(ctx.lib.ne_fractional_indices_f64(ctx.handle, g, C.byref(_source(spec)), fi.data_ptr(), fj.data_ptr(),
                                                s), "ne_fractional_indices_f64")
        return
    lam, phi = (torch.from_numpy(a).to(fi.device) for a in spec.centres())
    check(ctx.lib.ne_fractional_indices_nodes_f64(ctx.handle, g, lam.data_ptr(), spec.Nx, spec.topology_x,
                                                  phi.data_ptr(), spec.Ny, spec.topology_y, fi.data_ptr(),
                                                  fj.data_ptr(), s), "ne_fractional_indices_nodes_f64")
    torch.cuda.current_stream(fi.device).synchronize()     # the node arrays are freed on return


def _source(spec) -> abi.SourceGrid:
    return abi.SourceGrid(spec.Nx, spec.Ny, spec.topology_x, spec.topology_y, spec.lambda0, spec.dlambda, spec.phi0,
                          spec.dphi)


class OceanOnlyModel:
    """OceanOnlyModel(ocean; atmosphere, radiation) on one MI355X, state resident in HBM.

    precision: "fp64" (the drop-in) or "mixed" (FP32 T, S, velocities and prescribed fields; FP64 flux solve,
    fluxes and budgets).
    """

    def __init__(self, case: SurfaceCase, *, params: abi.Params | None = None, precision: str = "fp64",
                 radiation: bool = True, device: int = 0, time_indexing=None, atmosphere_window: int | None = None,
                 radiation_window: int | None = None, atmosphere_source=None, radiation_source=None,
                 lookahead: int | None = None, copy_stream: bool = False,
                 timestepper: str = "SplitRungeKutta3"):
        """time_indexing: Clamp() (default) or Cyclical(period) of both snapshot series.
        timestepper: "SplitRungeKutta3" (default) or "QuasiAdamsBashforth2" (FP64), as ocean_simulation's
        `timestepper` keyword (ocean_simulation.jl:323).
        atmosphere_window / radiation_window: keep only that many snapshots of the series on the device
        (a ring of slots, time_indices_in_memory) and stream the others from `atmosphere_source` /
        `radiation_source` (an ArraySnapshots or NpySnapshotStore; the case's arrays by default)."""
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        steppers = {"SplitRungeKutta3": abi.NE_TIMESTEPPER_SPLIT_RK3,
                    "QuasiAdamsBashforth2": abi.NE_TIMESTEPPER_QUASI_ADAMS_BASHFORTH2}
        if timestepper not in steppers:
            raise ValueError(f"timestepper must be one of {sorted(steppers)}")
        if timestepper == "QuasiAdamsBashforth2" and precision != "fp64":
            raise ValueError("the Quasi-Adams-Bashforth 2 step is FP64 (the mixed form's stages are the RK ones)")
        self.timestepper = timestepper
        self.case = case
        self.device = torch.device("cuda", device)
        self.ctx = Context.get(device)
        self.lib = self.ctx.lib
        self.precision = precision
        self.real = torch.float64 if precision == "fp64" else torch.float32
        self.params = params if params is not None else default_params()
        self.params.radiation = 1 if radiation else 0
        dev, R, f64 = self.device, self.real, torch.float64
        Nx, Ny, Nz = case.shape
        self.grid = DeviceGrid(case.grid, self.ctx, dev)
        self.T = _dev(case.T, dev, R)
        self.S = _dev(case.S, dev, R)
        self.u = _dev(case.u, dev, R)
        self.v = _dev(case.v, dev, R)
        self.w = _dev(case.w, dev, R)
        self._atimes = np.ascontiguousarray(case.atmos_times, dtype=np.float64)
        self._rtimes = np.ascontiguousarray(case.rad_times, dtype=np.float64)
        self.time_indexing = time_indexing
        self.time_indexing_kind, self.time_period = resolve_time_indexing(time_indexing, self._atimes)
        given = {"atmosphere": atmosphere_source, "radiation": radiation_source}
        sources = {"atmosphere": atmosphere_source or ArraySnapshots(case.atmos, self._atimes),
                   "radiation": radiation_source or ArraySnapshots(case.rad, self._rtimes)}
        windows = {"atmosphere": atmosphere_window, "radiation": radiation_window if radiation else None}
        resident = {}
        for series, arrays, times in (("atmosphere", case.atmos, self._atimes), ("radiation", case.rad, self._rtimes)):
            src = sources[series]
            if given[series] is not None and (src.n_times != len(times) or not np.array_equal(src.times, times)):
                raise ValueError(f"the {series} source's times differ from the case's")
            N = windows[series]
            if N is None:
                continue
            if int(N) < 2:
                raise ValueError(f"{series}_window must hold at least two snapshots")
            resident[series] = initial_window(times, int(N), time_indexing)
        self.atm, self.rad = {}, {}
        for series, arrays, out in (("atmosphere", case.atmos, self.atm), ("radiation", case.rad, self.rad)):
            if series not in resident and given[series] is None:
                out.update({k: _dev(v, dev, R) for k, v in arrays.items()})
                continue
            if series not in resident:           # a source without a window: the whole series, read once
                src = sources[series]
                host = np.empty(src.shape, dtype=np.float64 if R == torch.float64 else np.float32)
                for k in arrays:
                    for n in range(src.n_times):
                        src.read(n, k, host[n])
                    out[k] = torch.from_numpy(host).to(dev)
                continue
            src, N = sources[series], int(windows[series])
            stage = np.empty(src.shape[1:], dtype=np.float64 if R == torch.float64 else np.float32)
            for k in arrays:
                out[k] = torch.zeros((N,) + src.shape[1:], dtype=R, device=dev)
                for p in resident[series]:       # the window as created: position p in slot p mod N
                    src.read(p % src.n_times, k, stage)
                    out[k][p % N].copy_(torch.from_numpy(stage))
        self.src_atmos = _source(case.atmos_grid)
        self.src_rad = _source(case.rad_grid)
        self.afi = torch.empty(Nx, dtype=f64, device=dev)
        self.afj = torch.empty(Ny, dtype=f64, device=dev)
        self.rfi = torch.empty(Nx, dtype=f64, device=dev)
        self.rfj = torch.empty(Ny, dtype=f64, device=dev)
        s = stream_handle()
        fractional_indices(self.ctx, self.grid, case.atmos_grid, self.afi, self.afj)
        fractional_indices(self.ctx, self.grid, case.rad_grid, self.rfi, self.rfj)
        z2 = lambda: torch.zeros((Ny, Nx), dtype=f64, device=dev)  # noqa: E731
        self.fluxes = {k: z2() for k in FLUX_NAMES}
        self.iterations = torch.zeros((Ny, Nx), dtype=torch.int32, device=dev)
        self.net = {k: z2() for k in NET_NAMES}
        self.shortwave_surface_flux = z2()    # J₀ of TwoColorRadiation
        self.radiative = {k: z2() for k in RAD_NAMES}
        self.frazil_heat = z2()
        d = abi.ModelDesc()
        d.precision = abi.NE_FP64 if precision == "fp64" else abi.NE_MIXED_FP32
        d.fuse_freeze_clamp = 0
        d.grid = self.grid.struct
        d.params = self.params
        d.atmosphere = abi.PrescribedAtmosphereDesc(self.src_atmos, *(torch_view(self.atm[k]) for k in
                                                                       ("u", "v", "T", "q", "p", "rain", "snow")),
                                                    abi.TimeInterp(0, 0, 0.0))
        d.radiation = abi.PrescribedRadiationDesc(self.src_rad, torch_view(self.rad["downwelling_shortwave"]),
                                                  torch_view(self.rad["downwelling_longwave"]),
                                                  abi.TimeInterp(0, 0, 0.0))
        d.atmos_fi, d.atmos_fj = self.afi.data_ptr(), self.afj.data_ptr()
        d.radiation_fi, d.radiation_fj = self.rfi.data_ptr(), self.rfj.data_ptr()
        d.ocean = abi.OceanState(torch_view(self.u), torch_view(self.v), torch_view(self.w), torch_view(self.T),
                                 torch_view(self.S))
        d.fluxes = abi.InterfaceFluxes(*(torch_view(self.fluxes[k]) for k in FLUX_NAMES), torch_view(self.iterations))
        d.net = abi.NetOceanFluxes(*(torch_view(self.net[k]) for k in NET_NAMES),
                                   torch_view(self.shortwave_surface_flux))
        d.radiative = abi.RadiativeFluxes(*(torch_view(self.radiative[k]) for k in RAD_NAMES))
        d.frazil_heat = torch_view(self.frazil_heat)
        d.atmosphere_times = self._atimes.ctypes.data_as(C.POINTER(C.c_double))
        d.n_atmosphere_times = len(self._atimes)
        d.radiation_times = self._rtimes.ctypes.data_as(C.POINTER(C.c_double))
        d.n_radiation_times = len(self._rtimes)
        # the RK stage buffers come from torch's allocator, so creating and destroying a model
        # makes no device allocation of its own
        nbytes = C.c_int64()
        check(self.lib.ne_model_workspace_bytes(C.byref(d), C.byref(nbytes)), "ne_model_workspace_bytes")
        self._workspace = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
        d.workspace, d.workspace_bytes = self._workspace.data_ptr(), nbytes.value
        d.time_indexing, d.time_period = self.time_indexing_kind, self.time_period
        d.timestepper = steppers[timestepper]
        d.atmosphere_slots = int(atmosphere_window) if "atmosphere" in resident else 0
        d.radiation_slots = int(radiation_window) if "radiation" in resident else 0
        self.desc = d
        h = C.c_void_p()
        check(self.lib.ne_model_create(self.ctx.handle, C.byref(d), C.byref(h), s), "ne_model_create")
        self.handle = h
        self._budget_out = torch.zeros(2, dtype=f64, device=dev)
        self.streams = {}
        for which, series, fields in ((0, "atmosphere", self.atm), (1, "radiation", self.rad)):
            if series in resident:
                self.streams[series] = SnapshotStream(self, which, sources[series], fields, resident[series],
                                                      lookahead=lookahead, copy_stream=copy_stream)

    # reference API ------------------------------------------------------------
    def time_step(self, dt: float, stream=None) -> None:
        if self.streams:
            self._prepare(self.clock[0] + float(dt), stream)
        check(self.lib.ne_model_time_step(self.handle, float(dt), stream_handle(stream)), "ne_model_time_step")

    def update_state(self, stream=None) -> None:
        if self.streams:
            self._prepare(self.clock[0], stream)
        check(self.lib.ne_model_update_state(self.handle, stream_handle(stream)), "ne_model_update_state")

    def snapshots_at(self, time: float) -> tuple[int, int, int, int]:
        """Stream positions of the atmosphere's and the radiation's two snapshots read at `time`."""
        out = (C.c_int64 * 4)()
        check(self.lib.ne_model_snapshots_at(self.handle, float(time), out), "ne_model_snapshots_at")
        return tuple(out)

    def _prepare(self, time: float, stream) -> None:
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        pos = self.snapshots_at(time)
        for series, st in self.streams.items():
            st.prepare(pos[0:2] if series == "atmosphere" else pos[2:4], s)

    def update_vertical_velocity(self, stream=None) -> None:
        """w from continuity of the current u and v (after the caller sets new horizontal velocities)."""
        check(self.lib.ne_model_update_vertical_velocity(self.handle, stream_handle(stream)),
              "ne_model_update_vertical_velocity")

    @property
    def clock(self):
        t = C.c_double()
        n = C.c_int64()
        check(self.lib.ne_model_clock(self.handle, C.byref(t), C.byref(n)), "ne_model_clock")
        return t.value, n.value

    def timing_begin(self, max_steps: int) -> None:
        """Record HIP events around each kernel of the next max_steps time steps."""
        check(self.lib.ne_model_timing_begin(self.handle, int(max_steps)), "ne_model_timing_begin")

    def timing_end(self) -> tuple[dict, int]:
        """Summed milliseconds per kernel over the recorded steps, and the number of steps."""
        from ._lib import TIMED_KERNELS
        ms = (C.c_double * len(TIMED_KERNELS))()
        n = C.c_int()
        check(self.lib.ne_model_timing_end(self.handle, ms, C.byref(n)), "ne_model_timing_end")
        return dict(zip(TIMED_KERNELS, list(ms))), n.value

    def budget(self, field: torch.Tensor, weight: float = 1.0, stream=None) -> float:
        """Compensated Σ weight·V·c over active cells (returns sum + compensation)."""
        fn = self.lib.ne_budget_f64 if field.dtype == torch.float64 else self.lib.ne_budget_f32
        check(fn(self.ctx.handle, C.byref(self.grid.struct), torch_view(field), float(weight),
                 self._budget_out.data_ptr(), stream_handle(stream)), "ne_budget")
        s, c = self._budget_out.tolist()
        return s + c

    def heat_content(self) -> float:
        return self.budget(self.T, self.params.ocean_reference_density * self.params.ocean_heat_capacity)

    def rounding_residual(self, stream=None) -> tuple[float, float]:
        """Σ V·(x − stored) of T and S over the steps so far (zero in FP64): what the FP32 state of
        the mixed form has lost to rounding, so Σ V·T + this is conserved to FP64 round-off."""
        out = torch.empty(4, dtype=torch.float64, device=self.T.device)
        check(self.lib.ne_model_rounding_residual(self.handle, out.data_ptr(), stream_handle(stream)),
              "ne_model_rounding_residual")
        t, tc, s, sc = out.tolist()
        return t + tc, s + sc

    def conserved_heat(self) -> float:
        """Heat content of the state plus ρ₀c₀ times the rounding residual of T."""
        rho_c = self.params.ocean_reference_density * self.params.ocean_heat_capacity
        return self.heat_content() + rho_c * self.rounding_residual()[0]

    def close(self) -> None:
        for st in getattr(self, "streams", {}).values():
            st.close()
        if getattr(self, "handle", None):
            torch.cuda.synchronize(self.device)
            check(self.lib.ne_model_destroy(self.handle), "ne_model_destroy")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # host copies ----------------------------------------------------------------
    def fluxes_numpy(self) -> dict:
        d = {k: v.cpu().numpy() for k, v in self.fluxes.items()}
        d["iterations"] = self.iterations.cpu().numpy()
        return d

    def net_numpy(self) -> dict:
        return {k: v.cpu().numpy() for k, v in self.net.items()}


def time_step(model: OceanOnlyModel, dt: float) -> None:
    """time_step!(coupled_model, Δt)."""
    model.time_step(dt)


def update_state(model: OceanOnlyModel) -> None:
    """update_state!(coupled_model)."""
    model.update_state()


class Phases:
    """The per-phase entry points of the C ABI on device tensors (FP64)."""

    def __init__(self, model: OceanOnlyModel):
        if model.precision != "fp64":
            raise ValueError("per-phase entry points are FP64")
        self.m = model
        self.lib = model.lib
        self.ctx = model.ctx.handle
        dev = model.device
        Nx, Ny, Nz = model.case.shape
        z2 = lambda: torch.zeros((Ny, Nx), dtype=torch.float64, device=dev)  # noqa: E731
        self.exchange = {k: z2() for k in EXCHANGE_NAMES}
        self.sw, self.lw = z2(), z2()

    def _ex(self):
        ex = abi.ExchangeState(*(torch_view(self.exchange[k]) for k in EXCHANGE_NAMES))
        if self.m.desc.params.radiation:   # the radiation exchanger's state (skin-temperature balance)
            ex.downwelling_shortwave = torch_view(self.sw)
            ex.downwelling_longwave = torch_view(self.lw)
        return ex

    def interp(self, atmos_time: abi.TimeInterp | None = None, rad_time: abi.TimeInterp | None = None):
        m = self.m
        a = abi.PrescribedAtmosphereDesc.from_buffer_copy(m.desc.atmosphere)
        r = abi.PrescribedRadiationDesc.from_buffer_copy(m.desc.radiation)
        if atmos_time is not None:
            a.time = atmos_time
        if rad_time is not None:
            r.time = rad_time
        s = stream_handle()
        check(self.lib.ne_interp_atmos_f64(self.ctx, C.byref(m.grid.struct), C.byref(a), m.afi.data_ptr(),
                                           m.afj.data_ptr(), C.byref(self._ex()), s), "ne_interp_atmos_f64")
        check(self.lib.ne_interp_radiation_f64(self.ctx, C.byref(m.grid.struct), C.byref(r), m.rfi.data_ptr(),
                                               m.rfj.data_ptr(), torch_view(self.sw), torch_view(self.lw), s),
              "ne_interp_radiation_f64")

    def flux_solve(self):
        m = self.m
        check(self.lib.ne_ao_flux_solve_f64(self.ctx, C.byref(m.grid.struct), C.byref(m.params), C.byref(self._ex()),
                                            C.byref(m.desc.ocean), C.byref(m.desc.fluxes), stream_handle()),
              "ne_ao_flux_solve_f64")

    def freeze(self, dt: float):
        m = self.m
        check(self.lib.ne_freeze_clamp_f64(self.ctx, C.byref(m.grid.struct), C.byref(m.params), torch_view(m.T),
                                           torch_view(m.S), torch_view(m.frazil_heat), float(dt), stream_handle()),
              "ne_freeze_clamp_f64")

    def assemble(self):
        m = self.m
        check(self.lib.ne_assemble_net_ocean_fluxes_f64(self.ctx, C.byref(m.grid.struct), C.byref(m.params),
                                                        C.byref(m.desc.fluxes), C.byref(self._ex()),
                                                        torch_view(m.T), C.byref(m.desc.net), stream_handle()),
              "ne_assemble_net_ocean_fluxes_f64")

    def radiation(self):
        m = self.m
        check(self.lib.ne_apply_air_sea_radiation_f64(self.ctx, C.byref(m.grid.struct), C.byref(m.params),
                                                      torch_view(self.sw), torch_view(self.lw),
                                                      torch_view(m.fluxes["interface_temperature"]),
                                                      C.byref(m.desc.net), C.byref(m.desc.radiative),
                                                      stream_handle()), "ne_apply_air_sea_radiation_f64")

    def tendency(self, T: torch.Tensor | None = None, S: torch.Tensor | None = None):
        m = self.m
        T = m.T if T is None else T
        S = m.S if S is None else S
        GT = torch.zeros_like(T)
        GS = torch.zeros_like(S)
        o = abi.OceanState(torch_view(m.u), torch_view(m.v), torch_view(m.w), torch_view(T), torch_view(S))
        check(self.lib.ne_tracer_tendency_weno5_f64(self.ctx, C.byref(m.grid.struct), C.byref(m.params), C.byref(o),
                                                    C.byref(m.desc.net), torch_view(GT), torch_view(GS),
                                                    stream_handle()),
              "ne_tracer_tendency_weno5_f64")
        return GT, GS

    def compute_w(self, u: torch.Tensor, v: torch.Tensor, w: torch.Tensor) -> None:
        m = self.m
        check(self.lib.ne_compute_w_from_continuity_f64(self.ctx, C.byref(m.grid.struct), torch_view(u),
                                                        torch_view(v), torch_view(w), stream_handle()),
              "ne_compute_w_from_continuity_f64")

    def rk_stage(self, T0, S0, Tin, Sin, Tout, Sout, beta_dt: float):
        m = self.m
        check(self.lib.ne_tracer_rk_stage_f64(self.ctx, C.byref(m.grid.struct), C.byref(m.params), torch_view(m.u),
                                              torch_view(m.v),
                                              torch_view(m.w), torch_view(T0), torch_view(S0), torch_view(Tin),
                                              torch_view(Sin), torch_view(Tout), torch_view(Sout),
                                              C.byref(m.desc.net), float(beta_dt), stream_handle()),
              "ne_tracer_rk_stage_f64")


def cell_updates(case: SurfaceCase) -> int:
    Nx, Ny, Nz = case.shape
    return Nx * Ny * Nz


def bytes_per_cell_update(precision: str, Nz: int, n_stages: int = 3,
                          timestepper: str = "SplitRungeKutta3") -> float:
    """Algorithmic HBM bytes per cell-update: B = B₃ + B₂/Nz with B₃ = (7 + 9(n_st − 1))·s₃ for the
    split RK3 step (15·s₃ for the Quasi-Adams-Bashforth 2 step: the tendency reads u, v, w, T, S
    and writes G_T, G_S; the update reads T, S, G, G⁻ and writes T, S) and B₂ = 7.25·s₂ᵢ + 14·s₂ₒ."""
    s3, s2i, s2o = (8, 8, 8) if precision == "fp64" else (4, 4, 8)
    B3 = (15 if timestepper == "QuasiAdamsBashforth2" else 7 + 9 * (n_stages - 1)) * s3
    B2 = 7.25 * s2i + 14 * s2o
    return B3 + B2 / Nz
