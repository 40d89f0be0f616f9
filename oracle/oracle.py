"""Python driver of the CPU oracle (oracle/ne_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker of the HIP path.

``OracleModel`` owns numpy buffers in the same layouts the C ABI uses (i fastest)
and runs the reference algorithm phase by phase or as whole coupled steps.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
sys.path.insert(0, str(REPO / "numericalearth.jl_amd"))

from numericalearth_hip import _ctypes_abi as abi  # noqa: E402  (ABI struct layouts)
from numericalearth_hip.params import default_params  # noqa: E402

LIB_PATH = Path(os.environ.get("NE_ORACLE_LIB", HERE / "_build" / "libne_oracle.so"))


def build(verbose: bool = False) -> Path:
    res = subprocess.run(["make", "-C", str(HERE), "-B"], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    if verbose:
        sys.stdout.write(res.stdout)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        P = C.POINTER
        L.neo_flux_column.argtypes = [P(abi.Params), P(C.c_double), C.c_int, P(C.c_double)]
        L.neo_flux_column.restype = C.c_int
        for name in ("neo_saturation_vapor_pressure", "neo_cp_m", "neo_latent_heat_vapor"):
            getattr(L, name).argtypes = [P(abi.Params), C.c_double]
            getattr(L, name).restype = C.c_double
        for name in ("neo_surface_specific_humidity", "neo_virtual_temperature", "neo_saturation_specific_humidity"):
            getattr(L, name).argtypes = [P(abi.Params), C.c_double, C.c_double]
            getattr(L, name).restype = C.c_double
        L.neo_air_density.argtypes = [P(abi.Params), C.c_double, C.c_double, C.c_double]
        L.neo_air_density.restype = C.c_double
        for name in ("neo_psi_momentum", "neo_psi_scalar", "neo_momentum_roughness"):
            getattr(L, name).argtypes = [P(abi.Params), C.c_double]
            getattr(L, name).restype = C.c_double
        L.neo_latent_heat_sublim.argtypes = [P(abi.Params), C.c_double]
        L.neo_latent_heat_sublim.restype = C.c_double
        L.neo_neutral_drag.argtypes = [P(abi.Params), C.c_double]
        L.neo_neutral_drag.restype = C.c_double
        L.neo_psi_large_yeager.argtypes = [P(abi.Params), C.c_int, C.c_double]
        L.neo_psi_large_yeager.restype = C.c_double
        L.neo_large_yeager_coefficients.argtypes = [P(abi.Params)] + [C.c_double] * 7 + [P(C.c_double)]
        L.neo_pow_int.argtypes = [C.c_double, C.c_int]
        L.neo_pow_int.restype = C.c_double
        L.neo_gustiness_squared.argtypes = [P(abi.Params), C.c_double, C.c_double, C.c_double]
        L.neo_gustiness_squared.restype = C.c_double
        L.neo_scalar_roughness.argtypes = [P(abi.Params), C.c_double, C.c_double, C.c_int]
        L.neo_scalar_roughness.restype = C.c_double
        L.neo_weno5z.argtypes = [C.c_double] * 5
        L.neo_weno5z.restype = C.c_double
        L.neo_weno3z.argtypes = [C.c_double] * 3
        L.neo_weno3z.restype = C.c_double
        L.neo_set_threads.argtypes = [C.c_int]
        L.neo_get_threads.restype = C.c_int
        L.neo_model_update_state.argtypes = [C.c_void_p, C.c_double]
        L.neo_model_ocean_step.argtypes = [C.c_void_p, C.c_double]
        L.neo_budget.argtypes = [P(abi.Grid), abi.View, C.c_double, P(C.c_double)]
        L.neo_tracer_tendency.argtypes = [P(abi.Grid), P(abi.Params), P(abi.OceanState), P(abi.NetOceanFluxes),
                                          abi.View, abi.View]
        L.neo_tracer_tendency_scaled.argtypes = [P(abi.Grid), P(abi.Params), P(abi.OceanState),
                                                 P(abi.NetOceanFluxes), abi.View, abi.View, abi.View, abi.View]
        L.neo_rk_stage.argtypes = [P(abi.Grid), P(abi.Params)] + [abi.View] * 9 + [P(abi.NetOceanFluxes),
                                                                                  C.c_double]
        L.neo_ao_flux_solve.argtypes = [P(abi.Grid), P(abi.Params), P(abi.ExchangeState), P(abi.OceanState),
                                        P(abi.InterfaceFluxes)]
        L.neo_interp_atmos.argtypes = [P(abi.Grid), P(abi.PrescribedAtmosphereDesc), P(C.c_double), P(C.c_double),
                                       P(abi.ExchangeState)]
        L.neo_interp_radiation.argtypes = [P(abi.Grid), P(abi.PrescribedRadiationDesc), P(C.c_double),
                                           P(C.c_double), abi.View, abi.View]
        L.neo_fractional_indices.argtypes = [P(abi.Grid), P(abi.SourceGrid), P(C.c_double), P(C.c_double)]
        L.neo_fractional_indices_nodes.argtypes = [P(abi.Grid), P(C.c_double), C.c_int, C.c_int, P(C.c_double),
                                                   C.c_int, C.c_int, P(C.c_double), P(C.c_double)]
        L.neo_freeze_clamp.argtypes = [P(abi.Grid), P(abi.Params), abi.View, abi.View, abi.View, C.c_double]
        L.neo_assemble_net_ocean_fluxes.argtypes = [P(abi.Grid), P(abi.Params), P(abi.InterfaceFluxes),
                                                    P(abi.ExchangeState), abi.View, P(abi.NetOceanFluxes)]
        L.neo_apply_air_sea_radiation.argtypes = [P(abi.Grid), P(abi.Params), abi.View, abi.View, abi.View,
                                                  P(abi.NetOceanFluxes), P(abi.RadiativeFluxes)]
        L.neo_immersed_levels.argtypes = [P(abi.Grid), abi.View, P(C.c_int32)]
        L.neo_compute_w_from_continuity.argtypes = [P(abi.Grid), abi.View, abi.View, abi.View]
        _lib = L
    return _lib


def np_view(a: np.ndarray | None) -> abi.View:
    """ne_view of a C-ordered numpy array (i fastest); None gives a NULL view."""
    if a is None:
        return abi.View(None, 0, 0, 0)
    it = a.itemsize
    st = [s // it for s in a.strides]
    sx = st[-1] if a.ndim >= 1 else 0
    sy = st[-2] if a.ndim >= 2 else 0
    sz = st[-3] if a.ndim >= 3 else 0
    return abi.View(a.ctypes.data, sx, sy, sz)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def set_threads(n: int) -> None:
    lib().neo_set_threads(int(n))


def threads() -> int:
    return lib().neo_get_threads()


class _ModelStruct(C.Structure):
    # neo_model layout (oracle/ne_oracle.h)
    _fields_ = [
        ("grid", abi.Grid), ("params", abi.Params),
        ("atmosphere", abi.PrescribedAtmosphereDesc), ("radiation", abi.PrescribedRadiationDesc),
        ("atmos_fi", C.c_void_p), ("atmos_fj", C.c_void_p), ("radiation_fi", C.c_void_p), ("radiation_fj", C.c_void_p),
        ("ocean", abi.OceanState), ("exchange", abi.ExchangeState), ("sw", abi.View), ("lw", abi.View),
        ("fluxes", abi.InterfaceFluxes), ("net", abi.NetOceanFluxes), ("radiative", abi.RadiativeFluxes),
        ("frazil_heat", abi.View),
        ("T1", abi.View), ("S1", abi.View), ("T2", abi.View), ("S2", abi.View),
    ]


FLUX_NAMES = ("latent_heat", "sensible_heat", "water_vapor", "x_momentum", "y_momentum",
              "friction_velocity", "temperature_scale", "water_vapor_scale", "interface_temperature")
NET_NAMES = ("tau_x", "tau_y", "J_T", "J_S", "J_w", "J_H")
EXCHANGE_NAMES = ("u", "v", "T", "p", "q", "rain", "snow")
RAD_NAMES = ("upwelling_longwave", "downwelling_longwave", "downwelling_shortwave")


def time_indices(times: np.ndarray, t: float, indexing: int = abi.NE_TIME_CLAMPED, period: float | None = None):
    """cpu_interpolating_time_indices (interpolate_atmospheric_state.jl:57-60), Linear indexing
    clamped at the ends; with cyclical indexing Oceananigans' Cyclical(period) (the indexing the
    reference gives its JRA55 series, DataWrangling/JRA55/JRA55_prescribed_atmosphere.jl:17):
    t is reduced to t¹ + mod(t − t¹, period) (Julia's floored mod, which Python's float % is),
    and between the last snapshot and the first one a period later the two are interpolated."""
    n = len(times)
    if indexing == abi.NE_TIME_CYCLICAL and n > 1:
        t1, tN = float(times[0]), float(times[-1])
        tm = (t - t1) % period + t1
        if tm > tN:
            return n - 1, 0, (tm - tN) / (period - (tN - t1))
        t = tm
    if n <= 1 or t <= times[0]:
        return 0, 0, 0.0
    if t >= times[-1]:
        return n - 1, n - 1, 0.0
    k = int(np.searchsorted(times, t, side="right") - 1)
    return k, k + 1, (t - times[k]) / (times[k + 1] - times[k])


def snapshot_positions(times: np.ndarray, t: float, indexing: int = abi.NE_TIME_CLAMPED,
                       period: float | None = None):
    """Stream positions of the two snapshots read at t: the snapshot indices, or with cyclical
    indexing cycle·Nt + index, the second one in the next cycle across the wrap."""
    n1, n2, _ = time_indices(times, t, indexing, period)
    if indexing != abi.NE_TIME_CYCLICAL or len(times) <= 1:
        return n1, n2
    x = t - float(times[0])
    cycle = round((x - x % period) / period)
    return cycle * len(times) + n1, (cycle + (1 if n2 < n1 else 0)) * len(times) + n2


def fractional_indices(grid: abi.Grid, spec, fi: np.ndarray, fj: np.ndarray) -> None:
    """Fractional indices of the grid's cell centres in a source grid: its regular spacing, or
    its node coordinates when it has them (prescribed_atmosphere_regridder.jl:51-71)."""
    L = lib()
    if getattr(spec, "lambda_nodes", None) is None and getattr(spec, "phi_nodes", None) is None:
        src = abi.SourceGrid(spec.Nx, spec.Ny, spec.topology_x, spec.topology_y, spec.lambda0, spec.dlambda,
                             spec.phi0, spec.dphi)
        L.neo_fractional_indices(C.byref(grid), C.byref(src), _dp(fi), _dp(fj))
        return
    lam, phi = spec.centres()
    L.neo_fractional_indices_nodes(C.byref(grid), _dp(lam), spec.Nx, spec.topology_x, _dp(phi), spec.Ny,
                                   spec.topology_y, _dp(fi), _dp(fj))


class OracleModel:
    """CPU OceanOnlyModel on a SurfaceCase (copies of its arrays)."""

    def __init__(self, case, params: abi.Params | None = None, radiation: bool = True,
                 time_indexing: int = abi.NE_TIME_CLAMPED, time_period: float | None = None,
                 timestepper: str = "SplitRungeKutta3"):
        if timestepper not in ("SplitRungeKutta3", "QuasiAdamsBashforth2"):
            raise ValueError(timestepper)
        self.timestepper = timestepper
        self.case = case
        self.time_indexing, self.time_period = time_indexing, time_period
        g = case.grid
        spec = g.spec
        Nx, Ny, Nz = spec.Nx, spec.Ny, spec.Nz
        self.params = params if params is not None else default_params()
        self.params.radiation = 1 if radiation else 0
        keep = self._keep = []

        def K(a):
            a = np.ascontiguousarray(a)
            keep.append(a)
            return a

        self.dz, self.zc, self.dyx, self.dxy, self.Az = (K(g.dz), K(g.z_center), K(g.dy_xface), K(g.dx_yface), K(g.Az))
        self.lam, self.phi = K(g.lambda_c), K(g.phi_c)
        self.kmin = K(g.kmin.astype(np.int32))
        self.zf = K(g.z_face)
        self.grid = abi.Grid(Nx, Ny, Nz, spec.topology_x, spec.topology_y, spec.advection,
                             self.dz.ctypes.data, self.zc.ctypes.data, self.dyx.ctypes.data, self.dxy.ctypes.data,
                             self.Az.ctypes.data, self.lam.ctypes.data, self.phi.ctypes.data, self.kmin.ctypes.data,
                             self.zf.ctypes.data)
        self.T = K(case.T.copy())
        self.S = K(case.S.copy())
        self.u, self.v, self.w = K(case.u.copy()), K(case.v.copy()), K(case.w.copy())
        self.ocean = abi.OceanState(np_view(self.u), np_view(self.v), np_view(self.w), np_view(self.T), np_view(self.S))
        a = case.atmos_grid
        self.src_atmos = abi.SourceGrid(a.Nx, a.Ny, a.topology_x, a.topology_y, a.lambda0, a.dlambda, a.phi0, a.dphi)
        r = case.rad_grid
        self.src_rad = abi.SourceGrid(r.Nx, r.Ny, r.topology_x, r.topology_y, r.lambda0, r.dlambda, r.phi0, r.dphi)
        self.atm = {k: K(v) for k, v in case.atmos.items()}
        self.rad = {k: K(v) for k, v in case.rad.items()}
        self.atmosphere = abi.PrescribedAtmosphereDesc(self.src_atmos, *(np_view(self.atm[k]) for k in
                                                                          ("u", "v", "T", "q", "p", "rain", "snow")),
                                                       abi.TimeInterp(0, 0, 0.0))
        self.radiation = abi.PrescribedRadiationDesc(self.src_rad, np_view(self.rad["downwelling_shortwave"]),
                                                     np_view(self.rad["downwelling_longwave"]), abi.TimeInterp(0, 0, 0.0))
        self.afi, self.afj = K(np.zeros(Nx)), K(np.zeros(Ny))
        self.rfi, self.rfj = K(np.zeros(Nx)), K(np.zeros(Ny))
        L = lib()
        fractional_indices(self.grid, a, self.afi, self.afj)
        fractional_indices(self.grid, r, self.rfi, self.rfj)
        z2 = lambda: K(np.zeros((Ny, Nx)))  # noqa: E731
        self.exchange_arrays = {k: z2() for k in EXCHANGE_NAMES}
        self.exchange = abi.ExchangeState(*(np_view(self.exchange_arrays[k]) for k in EXCHANGE_NAMES))
        self.sw, self.lw = z2(), z2()
        if self.params.radiation:  # the radiation exchanger's state, read by the skin-temperature balance
            self.exchange.downwelling_shortwave = np_view(self.sw)
            self.exchange.downwelling_longwave = np_view(self.lw)
        self.flux_arrays = {k: z2() for k in FLUX_NAMES}
        self.iterations = K(np.zeros((Ny, Nx), np.int32))
        self.fluxes = abi.InterfaceFluxes(*(np_view(self.flux_arrays[k]) for k in FLUX_NAMES), np_view(self.iterations))
        self.net_arrays = {k: z2() for k in NET_NAMES}
        self.shortwave_surface_flux = z2()
        self.net = abi.NetOceanFluxes(*(np_view(self.net_arrays[k]) for k in NET_NAMES),
                                      np_view(self.shortwave_surface_flux))
        self.rad_arrays = {k: z2() for k in RAD_NAMES}
        self.radiative = abi.RadiativeFluxes(*(np_view(self.rad_arrays[k]) for k in RAD_NAMES))
        self.frazil = z2()
        self.T1, self.S1 = K(case.T.copy()), K(case.S.copy())
        self.T2, self.S2 = K(case.T.copy()), K(case.S.copy())
        self.time = 0.0
        self.iteration = 0
        self._model = _ModelStruct(self.grid, self.params, self.atmosphere, self.radiation,
                                   self.afi.ctypes.data, self.afj.ctypes.data, self.rfi.ctypes.data,
                                   self.rfj.ctypes.data, self.ocean, self.exchange, np_view(self.sw),
                                   np_view(self.lw), self.fluxes, self.net, self.radiative, np_view(self.frazil),
                                   np_view(self.T1), np_view(self.S1), np_view(self.T2), np_view(self.S2))

    # -- clock -------------------------------------------------------------
    def _set_time_indices(self):
        ix = (self.time_indexing, self.time_period)
        n1, n2, tn = time_indices(self.case.atmos_times, self.time, *ix)
        self._model.atmosphere.time = abi.TimeInterp(n1, n2, tn)
        n1, n2, tn = time_indices(self.case.rad_times, self.time, *ix)
        self._model.radiation.time = abi.TimeInterp(n1, n2, tn)

    def update_state(self, dt_frazil: float | None = None):
        if dt_frazil is None:
            dt_frazil = float("inf") if self.iteration == 0 else self.last_dt
        self._set_time_indices()
        lib().neo_model_update_state(C.byref(self._model), dt_frazil)

    def ab2_step(self, dt: float):
        """One Quasi-Adams-Bashforth 2 step of T and S (Oceananigans' QuasiAdamsBashforth2TimeStepper
        with χ = 0.1, the reference's timestepper = :QuasiAdamsBashforth2, ocean_simulation.jl:323):
        c ← c + Δt((3/2 + χ)Gⁿ − (1/2 + χ)G⁻) on active cells with Gⁿ the tendency of the state under
        the fluxes the last update_state computed; forward Euler (χ = −1/2, G⁻ = 0) on the first step
        and on a step whose Δt differs from the last one (ab2_step_field! and time_step!'s euler rule).
        Gⁿ is kept as the next step's G⁻."""
        GT, GS = self.tendency()
        euler = self.iteration == 0 or dt != self.last_dt
        chi = -0.5 if euler else abi.NE_AB2_CHI
        if euler:
            self.GmT, self.GmS = np.zeros_like(GT), np.zeros_like(GS)
        Nz = self.case.shape[2]
        active = np.arange(Nz)[:, None, None] >= self.kmin[None]
        a, b = 1.5 + chi, 0.5 + chi
        Tn = self.T + dt * (a * GT - b * self.GmT)
        Sn = self.S + dt * (a * GS - b * self.GmS)
        self.T[active] = Tn[active]
        self.S[active] = Sn[active]
        self.GmT, self.GmS = GT, GS

    def time_step(self, dt: float):
        if self.timestepper == "QuasiAdamsBashforth2":
            self.ab2_step(dt)
        else:
            lib().neo_model_ocean_step(C.byref(self._model), dt)
        self.time += dt
        self.iteration += 1
        self.last_dt = dt
        self.update_state(dt)

    def reconcile(self):
        """OceanOnlyModel construction: fluxes computed with Δt_frazil = ∞."""
        self.update_state(float("inf"))
        return self

    # -- phases --------------------------------------------------------------
    def interp(self):
        self._set_time_indices()
        L = lib()
        L.neo_interp_atmos(C.byref(self.grid), C.byref(self._model.atmosphere), _dp(self.afi), _dp(self.afj),
                           C.byref(self.exchange))
        L.neo_interp_radiation(C.byref(self.grid), C.byref(self._model.radiation), _dp(self.rfi), _dp(self.rfj),
                               np_view(self.sw), np_view(self.lw))

    def flux_solve(self):
        lib().neo_ao_flux_solve(C.byref(self.grid), C.byref(self.params), C.byref(self.exchange),
                                C.byref(self.ocean), C.byref(self.fluxes))

    def freeze(self, dt):
        lib().neo_freeze_clamp(C.byref(self.grid), C.byref(self.params), np_view(self.T), np_view(self.S),
                               np_view(self.frazil), dt)

    def assemble(self):
        lib().neo_assemble_net_ocean_fluxes(C.byref(self.grid), C.byref(self.params), C.byref(self.fluxes),
                                            C.byref(self.exchange), np_view(self.T), C.byref(self.net))

    def radiation_apply(self):
        lib().neo_apply_air_sea_radiation(C.byref(self.grid), C.byref(self.params), np_view(self.sw),
                                          np_view(self.lw), np_view(self.flux_arrays["interface_temperature"]),
                                          C.byref(self.net), C.byref(self.radiative))

    def tendency(self, T=None, S=None, scales: bool = False):
        """G_T, G_S (and, with scales=True, Σ|terms|/V of each cell for T and S)."""
        Nx, Ny, Nz = self.case.shape
        GT = np.zeros((Nz, Ny, Nx))
        GS = np.zeros((Nz, Ny, Nx))
        sT = np.zeros((Nz, Ny, Nx))
        sS = np.zeros((Nz, Ny, Nx))
        o = self.ocean
        if T is not None:
            T = np.ascontiguousarray(T)
            S = np.ascontiguousarray(S)
            o = abi.OceanState(o.u, o.v, o.w, np_view(T), np_view(S))
        lib().neo_tracer_tendency_scaled(C.byref(self.grid), C.byref(self.params), C.byref(o), C.byref(self.net),
                                         np_view(GT), np_view(GS), np_view(sT), np_view(sS))
        return (GT, GS, sT, sS) if scales else (GT, GS)

    def budget(self, c: np.ndarray, weight: float = 1.0):
        out = (C.c_double * 2)()
        lib().neo_budget(C.byref(self.grid), np_view(np.ascontiguousarray(c)), weight, out)
        return out[0], out[1]

    def fluxes_dict(self):
        d = {k: v.copy() for k, v in self.flux_arrays.items()}
        d["iterations"] = self.iterations.copy()
        return d

    def net_dict(self):
        return {k: v.copy() for k, v in self.net_arrays.items()}
