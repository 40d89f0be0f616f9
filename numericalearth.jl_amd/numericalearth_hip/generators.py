"""Seeded synthetic inputs of the coupled surface step (numpy, FP64).

The workloads follow the benchmark configurations of this path:

* LatitudeLongitudeGrid, λ ∈ [0, 360) Periodic, φ ∈ [−70, 70] Bounded,
  z ∈ [−3000, 0] with Nz uniform levels;
* bathymetry: seeded white noise, Gaussian-smoothed (σ = 4 cells at 1440 columns,
  scaled with resolution), mapped affinely to depth D ∈ [−400, 3000] m; columns
  with D < 50 m are land (bottom height 0) and wet columns are deepened to ≥ 200 m,
  as the reference's minimum-depth step does (Bathymetry/regrid_bathymetry.jl:267-275);
* T = 2 + 26 cos²φ exp(z/500) + N(0, 0.1) °C, S = 35 + 0.5 N(0, 1) psu,
  u, v = 0.1 N(0, 1) m/s on faces (zero on faces touching an inactive cell) minus
  their depth mean over the face's open levels, so that w, diagnosed from discrete
  continuity with w = 0 at the bottom, also vanishes at the top of every column;
* a prescribed atmosphere on a 2× coarser grid: u₁₀ = 8 sin 2φ, v₁₀ = 2 N(0, 1),
  Tᵃ = 2 + 26 cos²φ + 273.15 − 1 + N(0, 1) K, q = 0.8 qₛₐₜ(Tᵃ, p), p = 101325 Pa,
  rain = snow = 0; prescribed radiation SW = 300 cos φ, LW = 350 W m⁻²
  (SST is the analytic profile at the surface-cell centre).

Arrays are C-ordered with i fastest: 3-D (Nz, Ny, Nx), u (Nz, Ny, Nx+1),
v (Nz, Ny+1, Nx), w (Nz+1, Ny, Nx); 2-D (Ny, Nx); time series (Nt, Nyᵃ, Nxᵃ).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

EARTH_RADIUS = 6_371_000.0  # Oceananigans default sphere radius

NE_PERIODIC, NE_BOUNDED, NE_FLAT = 0, 1, 2


@dataclass
class GridSpec:
    Nx: int
    Ny: int
    Nz: int
    longitude: tuple = (0.0, 360.0)
    latitude: tuple = (-70.0, 70.0)
    z: tuple = (-3000.0, 0.0)
    topology_x: int = NE_PERIODIC
    topology_y: int = NE_BOUNDED
    radius: float = EARTH_RADIUS

    @property
    def advection(self) -> int:
        # ocean_simulation.jl:367-379: no advection on single-column grids
        return 0 if (self.Nx == 1 and self.Ny == 1) else 1


@dataclass
class GridArrays:
    spec: GridSpec
    dz: np.ndarray
    z_center: np.ndarray
    z_face: np.ndarray
    dy_xface: np.ndarray
    dx_yface: np.ndarray
    Az: np.ndarray
    lambda_c: np.ndarray
    phi_c: np.ndarray
    phi_f: np.ndarray
    bottom_height: np.ndarray
    kmin: np.ndarray


@dataclass
class SourceSpec:
    Nx: int
    Ny: int
    topology_x: int
    topology_y: int
    lambda0: float
    dlambda: float
    phi0: float
    dphi: float
    # Cell-centre coordinates of a source grid that is not equally spaced (ascending, degrees),
    # e.g. the Gaussian latitudes of JRA55's grid; None: the regular spacing above.
    lambda_nodes: np.ndarray | None = None
    phi_nodes: np.ndarray | None = None

    @property
    def irregular(self) -> bool:
        return self.lambda_nodes is not None or self.phi_nodes is not None

    def centres(self) -> tuple[np.ndarray, np.ndarray]:
        """Cell-centre longitudes and latitudes (the given nodes, or those of the regular spacing)."""
        lam = (self.lambda_nodes if self.lambda_nodes is not None
               else self.lambda0 + (np.arange(self.Nx) + 0.5) * self.dlambda)
        phi = self.phi_nodes if self.phi_nodes is not None else self.phi0 + (np.arange(self.Ny) + 0.5) * self.dphi
        return np.ascontiguousarray(lam, dtype=np.float64), np.ascontiguousarray(phi, dtype=np.float64)


def source_grid_from_interfaces(lambda_faces, phi_faces, topology_x: int = NE_PERIODIC,
                                topology_y: int = NE_BOUNDED) -> SourceSpec:
    """A source grid given by its cell interfaces (ascending, degrees), as the reference builds
    a LatitudeLongitudeGrid from a dataset's coordinate bounds: the cell centres are the
    midpoints of consecutive interfaces."""
    lf = np.asarray(lambda_faces, dtype=np.float64)
    pf = np.asarray(phi_faces, dtype=np.float64)
    nx, ny = len(lf) - 1, len(pf) - 1
    if nx < 1 or ny < 1 or np.any(np.diff(lf) <= 0) or np.any(np.diff(pf) <= 0):
        raise ValueError("interfaces must be ascending, at least two per direction")
    return SourceSpec(nx, ny, topology_x, topology_y, float(lf[0]), float((lf[-1] - lf[0]) / nx), float(pf[0]),
                      float((pf[-1] - pf[0]) / ny), lambda_nodes=0.5 * (lf[:-1] + lf[1:]),
                      phi_nodes=0.5 * (pf[:-1] + pf[1:]))


def gaussian_latitudes(n: int) -> np.ndarray:
    """The n Gaussian latitudes (degrees, ascending): arcsin of the roots of the Legendre
    polynomial P_n, the latitudes of a TL spectral model's grid such as JRA55's (n = 320)."""
    x, _ = np.polynomial.legendre.leggauss(n)
    return np.degrees(np.arcsin(x))


def gaussian_latitude_bounds(n: int) -> np.ndarray:
    """n + 1 cell bounds (degrees) around the n Gaussian latitudes: each cell holds the area
    its Gauss weight stands for, sin φ_{j+1/2} = −1 + Σ_{k≤j} w_k."""
    _, w = np.polynomial.legendre.leggauss(n)
    s = np.concatenate([[-1.0], -1.0 + np.cumsum(w)])
    s[-1] = 1.0
    return np.degrees(np.arcsin(np.clip(s, -1.0, 1.0)))


def jra55_source_grid(lat_bounds=None, lon_bounds=None) -> SourceSpec:
    """The JRA55-do atmosphere grid as the reference builds it
    (DataWrangling/JRA55/JRA55_metadata.jl:38-48): longitude interfaces are the files' lower
    `lon_bnds` closed by the first + 360, latitude interfaces their lower `lat_bnds` closed by
    90°, and the cell centres are the interfaces' midpoints. Pass the files' lower bounds
    (640 longitudes, 320 latitudes) when they are at hand; by default the longitudes are
    0.5625° cells centred on 0°, 0.5625°, … and the latitude cells are the Gauss-weight
    cells around the 320 Gaussian latitudes."""
    lon = -0.28125 + 0.5625 * np.arange(640) if lon_bounds is None else np.asarray(lon_bounds, dtype=np.float64)
    lat = gaussian_latitude_bounds(320)[:-1] if lat_bounds is None else np.asarray(lat_bounds, dtype=np.float64)
    return source_grid_from_interfaces(np.append(lon, lon[0] + 360.0), np.append(lat, 90.0))


@dataclass
class SurfaceCase:
    grid: GridArrays
    T: np.ndarray
    S: np.ndarray
    u: np.ndarray
    v: np.ndarray
    w: np.ndarray
    atmos_grid: SourceSpec
    atmos: dict            # u v T q p rain snow → (Nt, Nyᵃ, Nxᵃ)
    atmos_times: np.ndarray
    rad_grid: SourceSpec
    rad: dict              # downwelling_shortwave / downwelling_longwave
    rad_times: np.ndarray
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def shape(self):
        g = self.grid.spec
        return (g.Nx, g.Ny, g.Nz)


def latlon_metrics(spec: GridSpec):
    """Exact spherical metrics of a regular LatitudeLongitudeGrid."""
    Nx, Ny, Nz = spec.Nx, spec.Ny, spec.Nz
    lon0, lon1 = spec.longitude
    lat0, lat1 = spec.latitude
    dlon = (lon1 - lon0) / Nx
    dlat = (lat1 - lat0) / Ny
    R = spec.radius
    lambda_c = lon0 + (np.arange(Nx) + 0.5) * dlon
    phi_f = lat0 + np.arange(Ny + 1) * dlat
    phi_c = lat0 + (np.arange(Ny) + 0.5) * dlat
    z0, z1 = spec.z
    z_face = np.linspace(z0, z1, Nz + 1)
    dz = np.diff(z_face)
    z_center = 0.5 * (z_face[:-1] + z_face[1:])
    dlon_r = math.radians(dlon)
    dlat_r = math.radians(dlat)
    dy_xface = np.full(Ny, R * dlat_r)
    dx_yface = R * np.cos(np.radians(phi_f)) * dlon_r
    Az = R * R * dlon_r * (np.sin(np.radians(phi_f[1:])) - np.sin(np.radians(phi_f[:-1])))
    return dict(lambda_c=lambda_c, phi_c=phi_c, phi_f=phi_f, z_face=z_face, dz=dz, z_center=z_center,
                dy_xface=dy_xface, dx_yface=dx_yface, Az=Az)


def immersed_levels(z_center: np.ndarray, bottom_height: np.ndarray) -> np.ndarray:
    """First active level per column: cell (i,j,k) is immersed iff z_center[k] <= bottom."""
    return np.sum(z_center[None, None, :] <= bottom_height[:, :, None], axis=2).astype(np.int32)


def _smooth(noise: np.ndarray, sigma: float, periodic_x: bool) -> np.ndarray:
    if sigma <= 0:
        return noise
    from scipy.ndimage import gaussian_filter
    modes = ("reflect", "wrap" if periodic_x else "reflect")
    return gaussian_filter(noise, sigma=sigma, mode=modes)


def bathymetry(spec: GridSpec, rng: np.random.Generator, sigma_cells: float | None = None,
               depth_range=(-400.0, 3000.0), land_depth=50.0, minimum_depth=200.0) -> np.ndarray:
    Ny, Nx = spec.Ny, spec.Nx
    noise = rng.standard_normal((Ny, Nx))
    if sigma_cells is None:
        sigma_cells = 4.0 * Nx / 1440.0
    sm = _smooth(noise, sigma_cells, spec.topology_x == NE_PERIODIC)
    lo, hi = sm.min(), sm.max()
    span = hi - lo if hi > lo else 1.0
    D = depth_range[0] + (sm - lo) / span * (depth_range[1] - depth_range[0])
    land = D < land_depth
    D = np.maximum(D, minimum_depth)
    zmin = spec.z[0]
    D = np.minimum(D, -zmin)
    bottom = np.where(land, 0.0, -D)
    return bottom


def saturation_specific_humidity(T, p, *, R=8.3144598, Md=0.02897, Mv=0.018015, cpv=1859.0, cpl=4181.0,
                                 LH=2500800.0, T0=273.16, Ttr=273.16, ptr=611.657):
    """Liquid-phase qᵥ⁺ (interface_states.jl:79-90) with the Thermodynamics.jl pₛₐₜ form."""
    Rv = R / Mv
    dcp = cpv - cpl
    psat = ptr * np.power(T / Ttr, dcp / Rv) * np.exp((LH - dcp * T0) / Rv * (1.0 / Ttr - 1.0 / T))
    edv = 1.0 / (Md / Mv)
    pv = np.minimum(psat, 0.999 * p)
    return edv * pv / (p - (1.0 - edv) * pv)


def _face_velocities(spec: GridSpec, kmin: np.ndarray, rng: np.random.Generator, amp=0.1):
    Nx, Ny, Nz = spec.Nx, spec.Ny, spec.Nz
    k = np.arange(Nz)[:, None, None]
    active = k >= kmin[None, :, :]                        # (Nz, Ny, Nx)
    u = amp * rng.standard_normal((Nz, Ny, Nx + 1))
    v = amp * rng.standard_normal((Nz, Ny + 1, Nx))
    # x faces: face i between cells i-1 and i
    if spec.topology_x == NE_PERIODIC:
        west = np.roll(active, 1, axis=2)
        ok = west & active
        u[:, :, :Nx] = np.where(ok, u[:, :, :Nx], 0.0)
        u[:, :, Nx] = u[:, :, 0]
    else:
        okx = np.zeros((Nz, Ny, Nx + 1), bool)
        okx[:, :, 1:Nx] = active[:, :, :-1] & active[:, :, 1:]
        u = np.where(okx, u, 0.0)
    if spec.topology_y == NE_PERIODIC:
        south = np.roll(active, 1, axis=1)
        ok = south & active
        v[:, :Ny, :] = np.where(ok, v[:, :Ny, :], 0.0)
        v[:, Ny, :] = v[:, 0, :]
    else:
        oky = np.zeros((Nz, Ny + 1, Nx), bool)
        oky[:, 1:Ny, :] = active[:, :-1, :] & active[:, 1:, :]
        v = np.where(oky, v, 0.0)
    if spec.Nx == 1 and spec.Ny == 1:
        u[:] = 0.0
        v[:] = 0.0
    return u, v, active


def remove_barotropic_transport(u: np.ndarray, v: np.ndarray, dz: np.ndarray):
    """Subtract from every face its thickness-weighted mean over the face's open levels, so the
    depth-integrated transport through every face vanishes. Continuity then gives w = 0 at the
    top of every column: with a zero advective flux through the top face (the top boundary
    carries only the prescribed flux), a column whose integrated divergence is not zero would
    act as a mass source in its top cell and its tracer would grow or decay exponentially
    (e-folding time Δz/w_top, hours at 1/4°)."""
    out = []
    for f in (u, v):
        open_ = f != 0.0
        w = dz[:, None, None] * open_
        depth = w.sum(axis=0)
        mean = np.where(depth > 0, (dz[:, None, None] * f).sum(axis=0) / np.where(depth > 0, depth, 1.0), 0.0)
        out.append(np.where(open_, f - mean[None], 0.0))
    return out[0], out[1]


def continuity_w(g: GridArrays, u: np.ndarray, v: np.ndarray, active: np.ndarray) -> np.ndarray:
    """w on (C,C,F) faces from discrete continuity, w = 0 at the bottom of each column."""
    spec = g.spec
    Nx, Ny, Nz = spec.Nx, spec.Ny, spec.Nz
    w = np.zeros((Nz + 1, Ny, Nx))
    for k in range(Nz):
        Ax = g.dy_xface[:, None] * g.dz[k]
        Ay = g.dx_yface[:, None] * g.dz[k]
        Fu = Ax * u[k]                                     # (Ny, Nx+1)
        Fv = Ay * v[k]                                     # (Ny+1, Nx)
        div = (Fu[:, 1:] - Fu[:, :-1]) + (Fv[1:, :] - Fv[:-1, :])
        wk = w[k] - div / g.Az[:, None]
        w[k + 1] = np.where(active[k], wk, 0.0)
    return w


def make_case(Nx: int, Ny: int, Nz: int, seed: int = 0, *, longitude=(0.0, 360.0), latitude=(-70.0, 70.0),
              z=(-3000.0, 0.0), topology_x=NE_PERIODIC, atmos_coarsening=2, n_times=1, dt_times=3600.0 * 6,
              sigma_cells=None) -> SurfaceCase:
    """Seeded synthetic coupled-surface workload (see module docstring)."""
    rng = np.random.default_rng(seed)
    spec = GridSpec(Nx, Ny, Nz, longitude=longitude, latitude=latitude, z=z, topology_x=topology_x)
    m = latlon_metrics(spec)
    bottom = bathymetry(spec, rng, sigma_cells)
    if Nx == 1 and Ny == 1:
        bottom = np.full((1, 1), z[0] - 1.0)               # a single column is never land
    kmin = immersed_levels(m["z_center"], bottom)
    g = GridArrays(spec=spec, bottom_height=bottom, kmin=kmin, **m)

    zc = m["z_center"][:, None, None]
    phi = np.radians(m["phi_c"])[None, :, None]
    T = 2.0 + 26.0 * np.cos(phi) ** 2 * np.exp(zc / 500.0) + 0.1 * rng.standard_normal((Nz, Ny, Nx))
    S = 35.0 + 0.5 * rng.standard_normal((Nz, Ny, Nx))
    u, v, active = _face_velocities(spec, kmin, rng)
    u, v = remove_barotropic_transport(u, v, m["dz"])
    w = continuity_w(g, u, v, active)

    # Prescribed atmosphere on a coarser grid covering the same domain.
    Nxa = max(1, Nx // atmos_coarsening)
    Nya = max(1, Ny // atmos_coarsening)
    dlon_a = (longitude[1] - longitude[0]) / Nxa
    dlat_a = (latitude[1] - latitude[0]) / Nya
    topo_xa = topology_x if Nxa > 1 else NE_FLAT
    topo_ya = NE_BOUNDED if Nya > 1 else NE_FLAT
    src = SourceSpec(Nxa, Nya, topo_xa, topo_ya, longitude[0], dlon_a, latitude[0], dlat_a)
    phia = np.radians(latitude[0] + (np.arange(Nya) + 0.5) * dlat_a)[:, None]
    shape = (n_times, Nya, Nxa)
    ua = np.broadcast_to(8.0 * np.sin(2.0 * phia), shape).copy()
    va = 2.0 * rng.standard_normal(shape)
    # SST of the analytic profile at the centre of the surface cell
    sst = 2.0 + 26.0 * np.cos(phia) ** 2 * np.exp(m["z_center"][-1] / 500.0)
    Ta = np.broadcast_to(sst + 273.15 - 1.0, shape) + rng.standard_normal(shape)
    pa = np.full(shape, 101325.0)
    qa = 0.8 * saturation_specific_humidity(Ta, pa)
    zero = np.zeros(shape)
    atmos = dict(u=ua, v=va, T=Ta, q=qa, p=pa, rain=zero.copy(), snow=zero.copy())
    sw = np.broadcast_to(300.0 * np.cos(phia), shape).copy()
    lw = np.full(shape, 350.0)
    times = np.arange(n_times) * dt_times
    return SurfaceCase(grid=g, T=T, S=S, u=u, v=v, w=w, atmos_grid=src, atmos=atmos, atmos_times=times,
                       rad_grid=src, rad=dict(downwelling_shortwave=sw, downwelling_longwave=lw),
                       rad_times=times.copy(), seed=seed,
                       meta=dict(land_fraction=float(np.mean(kmin >= Nz))))


CONFIGS = {
    # name: (Nx, Ny, Nz)
    "cpu_reference_90x40x10": (90, 40, 10),
    "quarter_degree_1440x560x10": (1440, 560, 10),
}
