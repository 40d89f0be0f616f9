"""Halo-padded views and the oracle's ring of the surface phases (shared by the CPU and GPU ring
tests). H = 2: the ring is the first halo layer; the second must stay untouched."""
import ctypes as C

import numpy as np

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip.generators import make_case
from oracle import EXCHANGE_NAMES, FLUX_NAMES, np_view

H = 2


def ring_case(topology: str):
    periodic = topology == "periodic_x"
    lon = (0.0, 360.0) if periodic else (100.0, 160.0)
    case = make_case(48, 20, 6, seed=31, longitude=lon, topology_x=abi.NE_PERIODIC if periodic else abi.NE_BOUNDED)
    return case, periodic


def padded_np(Ny, Nx, dtype=np.float64, fill=np.nan):
    a = np.full((Ny + 2 * H, Nx + 2 * H), fill, dtype=dtype)
    return a, np_view(a[H:, H:])


def halo_filled(c: np.ndarray, periodic_x: bool) -> np.ndarray:
    """A 3-D (Nz, ny, nx) cell field with an H-wide halo filled as fill_halo_regions! fills it:
    wrapped in a periodic x; the edge value repeated in bounded directions (a ring cell there is
    inactive, so any finite value serves)."""
    out = np.pad(c, ((0, 0), (H, H), (H, H)), mode="edge")
    if periodic_x:
        out[:, :, :H] = out[:, :, -2 * H:-H]
        out[:, :, -H:] = out[:, :, H:2 * H]
    return np.ascontiguousarray(out)


def face_halo_periodic_x(u: np.ndarray) -> np.ndarray:
    """x-face field (Nz, Ny, Nx+1) whose face Nx is face 0 (periodic): halo faces −m = Nx−m and
    Nx+m = m."""
    Nx = u.shape[2] - 1
    out = np.pad(u, ((0, 0), (H, H), (H, H)), mode="edge")
    for m in range(1, H + 1):
        out[:, :, H - m] = out[:, :, H + Nx - m]
        out[:, :, H + Nx + m] = out[:, :, H + m]
    return np.ascontiguousarray(out)


def interior3(a: np.ndarray) -> abi.View:
    return np_view(a[:, H:, H:])


def masks(Ny, Nx):
    ring = np.zeros((Ny + 2 * H, Nx + 2 * H), bool)
    ring[H - 1:H + Ny + 1, H - 1:H + Nx + 1] = True
    ring[H:H + Ny, H:H + Nx] = False
    inner = np.zeros_like(ring)
    inner[H:H + Ny, H:H + Nx] = True
    return ring, inner, ~(ring | inner)


def oracle_ring(case, periodic: bool):
    """The oracle's interpolations and flux solve on halo-padded views (grid.halo = H), the ocean
    surface read from halos filled as Oceananigans fills them. Returns the oracle model and
    {name: padded array} of the exchange state, the radiation and the interface fluxes."""
    om = O.OracleModel(case)
    Ny, Nx = case.shape[1], case.shape[0]
    g = abi.Grid.from_buffer_copy(om.grid)
    g.halo = H
    ex = {k: padded_np(Ny, Nx) for k in EXCHANGE_NAMES}
    sw, lw = padded_np(Ny, Nx), padded_np(Ny, Nx)
    exs = abi.ExchangeState(*(ex[k][1] for k in EXCHANGE_NAMES))
    exs.downwelling_shortwave, exs.downwelling_longwave = sw[1], lw[1]
    fl = {k: padded_np(Ny, Nx) for k in FLUX_NAMES}
    it = padded_np(Ny, Nx, np.int32, -7)
    fls = abi.InterfaceFluxes(*(fl[k][1] for k in FLUX_NAMES), it[1])
    T, S = halo_filled(case.T, periodic), halo_filled(case.S, periodic)
    u = face_halo_periodic_x(case.u) if periodic else halo_filled(case.u, False)
    v, w = halo_filled(case.v, periodic), halo_filled(case.w, periodic)
    ocean = abi.OceanState(interior3(u), interior3(v), interior3(w), interior3(T), interior3(S))
    L = O.lib()
    L.neo_interp_atmos(C.byref(g), C.byref(om.atmosphere), O._dp(om.afi), O._dp(om.afj), C.byref(exs))
    L.neo_interp_radiation(C.byref(g), C.byref(om.radiation), O._dp(om.rfi), O._dp(om.rfj), sw[1], lw[1])
    L.neo_ao_flux_solve(C.byref(g), C.byref(om.params), C.byref(exs), C.byref(ocean), C.byref(fls))
    out = {f"exchange {k}": ex[k][0] for k in EXCHANGE_NAMES}
    out["shortwave"], out["longwave"] = sw[0], lw[0]
    out.update({k: fl[k][0] for k in FLUX_NAMES})
    out["iterations"] = it[0]
    return om, out
