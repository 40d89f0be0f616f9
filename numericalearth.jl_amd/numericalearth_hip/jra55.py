"""JRA55-do forcing files as snapshot sources.

The reference reads each JRA55 variable from its own file — ``RYF.<short>.1990_1991.nc`` for
the repeat-year forcing (DataWrangling/JRA55/JRA55_metadata.jl:92-96) — one 2-D slice
``ds[name][:, :, n]`` per snapshot (JRA55_field_time_series.jl:13-33), on the grid built from
the files' ``lon_bnds`` / ``lat_bnds`` (JRA55_metadata.jl:38-48). ``JRA55Snapshots`` is that
reader as a snapshot source for ``OceanOnlyModel(atmosphere_source=…, radiation_source=…)``:
one variable per file, laid out (time, lat, lon) as NetCDF stores it, which is the (Nt, Ny,
Nx) order, i fastest, of the model's prescribed fields.

Files are opened with the ``netCDF4`` module when it is installed (NetCDF-4/HDF5, the format
JRA55-do is distributed in) and otherwise with ``scipy.io.netcdf_file``, which reads the
classic NetCDF-3 formats only (``nccopy -k classic`` converts a NetCDF-4 file).
"""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np

from .generators import NE_BOUNDED, SurfaceCase, source_grid_from_interfaces

# JRA55_dataset_variable_names (JRA55_metadata.jl:135-147), by the model's field names
ATMOSPHERE_VARIABLES = {"u": "uas", "v": "vas", "T": "tas", "q": "huss", "p": "psl", "rain": "prra",
                        "snow": "prsn"}
RADIATION_VARIABLES = {"downwelling_shortwave": "rsds", "downwelling_longwave": "rlds"}

_UNIT_SECONDS = {"second": 1.0, "seconds": 1.0, "s": 1.0, "minute": 60.0, "minutes": 60.0, "hour": 3600.0,
                 "hours": 3600.0, "h": 3600.0, "day": 86400.0, "days": 86400.0, "d": 86400.0}


def repeat_year_filename(short: str) -> str:
    """File of one repeat-year JRA55-do variable (JRA55_metadata.jl:92-96)."""
    return f"RYF.{short}.1990_1991.nc"


def open_dataset(path):
    """An open NetCDF file: netCDF4.Dataset if that module is installed, else scipy's
    classic-format reader (memory-mapped)."""
    try:
        import netCDF4
    except ImportError:
        netCDF4 = None
    if netCDF4 is not None:
        ds = netCDF4.Dataset(str(path))
        ds.set_auto_mask(False)
        return ds
    with open(path, "rb") as f:
        magic = f.read(8)
    if magic.startswith(b"\x89HDF"):
        raise OSError(f"{path} is NetCDF-4 (HDF5): reading it needs the netCDF4 module, or convert it with "
                      "`nccopy -k classic`")
    from scipy.io import netcdf_file
    return netcdf_file(str(path), "r", mmap=True)


def _attr(var, name, default=None):
    v = getattr(var, name, default)
    return v.decode() if isinstance(v, bytes) else v


def _seconds_per_unit(units: str) -> float:
    m = re.match(r"\s*(\w+)\s+since\b", units or "")
    if not m or m.group(1).lower() not in _UNIT_SECONDS:
        raise ValueError(f"time units {units!r} are not '<unit> since <date>'")
    return _UNIT_SECONDS[m.group(1).lower()]


def restrict_interfaces(lo: float, hi: float, interfaces: np.ndarray) -> tuple[int, int]:
    """Interfaces i0 … i1 (0-based, inclusive) of the cells that keep [lo, hi] of an ascending
    interface vector: the cell whose centre is at or below lo, the one whose centre is at or above
    hi, and those between, so the box lies inside the hull of the kept centres; a box edge in a
    cell's first half, or on a face, takes one more cell, an edge past the centre none; clamped at
    the vector's ends, at least one cell (metadata_field.jl:47-56)."""
    f = np.asarray(interfaces, dtype=np.float64)
    n = len(f)
    k = min(max(int(np.searchsorted(f, lo, side="right")) - 1, 0), n - 2)
    i0 = k if 0.5 * (f[k] + f[k + 1]) <= lo else max(k - 1, 0)
    m = min(max(int(np.searchsorted(f, hi, side="left")), 1), n - 1)
    i1 = m if 0.5 * (f[m - 1] + f[m]) >= hi else min(m + 1, n - 1)
    if i1 - i0 < 1:
        i1 = min(i0 + 1, n - 1)
        i0 = i1 - 1
    return i0, i1


class JRA55Snapshots:
    """Snapshots of JRA55 variables, one NetCDF file per variable, read one 2-D slice at a time.

    fields: model field name → JRA55 short name (``ATMOSPHERE_VARIABLES`` or
    ``RADIATION_VARIABLES``); filename: short name → file name in `directory`. Snapshot n is
    at n·interval seconds, as the reference places the repeat-year record on its 3-hourly
    dates whatever each file's own stamps (JRA55_metadata.jl:55-61,
    JRA55_field_time_series.jl:13-26); ``interval=None`` takes the times, in seconds from
    the first, from the first file's ``time`` variable. Every file holds the same number of
    snapshots; the grid comes from the first file's bounds."""

    def __init__(self, directory, fields: dict, filename=repeat_year_filename, interval: float | None = 10800.0,
                 region=None):
        if region is not None:
            (west, east), (south, north) = region
            if not (east > west and north > south):
                raise ValueError(f"region {region} is empty")
        self.directory = Path(directory)
        self.short = dict(fields)
        self._ds = {k: open_dataset(self.directory / filename(s)) for k, s in self.short.items()}
        self._var = {k: ds.variables[self.short[k]] for k, ds in self._ds.items()}
        first = next(iter(self._ds.values()))
        nt = self._var[next(iter(self._var))].shape[0]
        if interval is not None:                 # the repeat-year convention: snapshot n at n·3 h
            self.times = interval * np.arange(nt, dtype=np.float64)
        else:                                    # the first file's own time axis
            t = first.variables["time"]
            raw = np.asarray(t[:], dtype=np.float64)
            self.times = np.ascontiguousarray((raw - raw[0]) * _seconds_per_unit(_attr(t, "units")))
        for k, v in self._var.items():
            if v.shape[0] != nt:
                raise ValueError(f"{self.short[k]}: {v.shape[0]} snapshots, {nt} in the first file")
            if v.dimensions != ("time", "lat", "lon"):
                raise ValueError(f"{self.short[k]}: dimensions {self._var[k].dimensions}, expected (time, lat, lon)")
        lat = np.asarray(first.variables["lat"][:], dtype=np.float64)
        self._flip = len(lat) > 1 and lat[0] > lat[-1]         # stored north to south: served south to north
        self._scale = {k: (float(_attr(v, "scale_factor", 1.0)), float(_attr(v, "add_offset", 0.0)))
                       for k, v in self._var.items()}
        self._lf, self._pf = self._interfaces(first)
        nx, ny = len(self._lf) - 1, len(self._pf) - 1
        self._cols, self._rows = np.arange(nx), slice(0, ny)
        self._lon_shift, self._periodic = 0.0, True
        if region is not None:
            self._restrict(region)
        self.shape = (len(self.times), self._rows.stop - self._rows.start, len(self._cols))

    def _interfaces(self, ds):
        """Full-record interfaces as the reference builds them: lower lon/lat bounds closed by
        +360 and 90° (JRA55_metadata.jl:38-48); files without bounds give interfaces halfway
        between their coordinates."""
        v = ds.variables
        lon = np.asarray(v["lon"][:], dtype=np.float64)
        lat = np.asarray(v["lat"][:], dtype=np.float64)
        if "lon_bnds" in v and "lat_bnds" in v:
            lf = np.asarray(v["lon_bnds"][:], dtype=np.float64).min(axis=1)
            pf = np.asarray(v["lat_bnds"][:], dtype=np.float64).min(axis=1)
            if self._flip:
                pf = pf[::-1]
            return np.append(lf, lf[0] + 360.0), np.append(pf, 90.0)
        lat = np.sort(lat)
        lf = np.concatenate([[lon[0] - 0.5 * (lon[1] - lon[0])], 0.5 * (lon[1:] + lon[:-1])])
        pf = np.concatenate([[max(-90.0, lat[0] - 0.5 * (lat[1] - lat[0]))], 0.5 * (lat[1:] + lat[:-1]),
                             [min(90.0, lat[-1] + 0.5 * (lat[-1] - lat[-2]))]])
        return np.append(lf, lf[0] + 360.0), pf

    def _restrict(self, region):
        """Keep the cells of a bounding box ((west, east), (south, north)) as the reference restricts
        a series given by interface vectors (restrict(bbox, interfaces::AbstractVector, N),
        metadata_field.jl:47-56, through restrict_longitude, :66-67): the cells whose centres
        bracket the box (restrict_interfaces). Rows and columns are read only there, and the grid
        is the sub-grid, bounded in longitude unless it closes the circle, its longitudes in the
        box's own convention (a box across 0° given as −20…30 stays −20…30 whatever the files'
        0…360). A box that crosses the record's seam (the files' first longitude) is continued
        round the circle: the reference's vector restriction clamps at the record's end there and
        would leave the part of the box beyond the seam outside its grid."""
        (west, east), (south, north) = region
        lf, pf = self._lf, self._pf
        nx = len(lf) - 1
        j0, j1 = restrict_interfaces(south, north, pf)
        self._rows = slice(j0, j1)
        if east - west >= 360.0:
            return
        turn = np.floor((west - lf[0]) / 360.0)
        w = west - 360.0 * turn                                 # the box's west edge within the record's turn
        e = w + (east - west)
        # the interfaces continued one turn either side: the same cells as the record's own vector
        # wherever the bracketing stays inside the record, and the neighbours across the seam where not
        faces = np.concatenate([lf[:-1] - 360.0, lf[:-1], lf + 360.0])
        i0, i1 = restrict_interfaces(w, e, faces)
        if i1 - i0 >= nx:                                       # the box closes the circle
            return
        self._cols = np.arange(i0, i1) % nx
        self._lon_faces = faces[i0:i1 + 1] + 360.0 * turn
        self._periodic = False

    @property
    def n_times(self) -> int:
        return len(self.times)

    def read(self, n: int, name: str, out: np.ndarray) -> None:
        """Copy snapshot n of field `name` into `out` (Ny, Nx), converted to out's dtype."""
        a = np.asarray(self._var[name][n])
        if self._flip:
            a = a[::-1]
        a = a[self._rows]
        if len(self._cols) != a.shape[1]:
            a = a[:, self._cols]
        scale, offset = self._scale[name]
        if scale != 1.0 or offset != 0.0:
            a = a * scale + offset
        np.copyto(out, a, casting="unsafe")

    def grid(self):
        """The source grid: the record's interfaces (centres at their midpoints), or the region's
        sub-grid."""
        pf = self._pf[self._rows.start:self._rows.stop + 1]
        if self._periodic:
            return source_grid_from_interfaces(self._lf, pf)
        return source_grid_from_interfaces(self._lon_faces, pf, topology_x=NE_BOUNDED)

    def close(self) -> None:
        self._var.clear()                        # scipy's reader closes only once no variable refers to its map
        for ds in self._ds.values():
            ds.close()
        self._ds.clear()


def with_forcing(case: SurfaceCase, atmosphere: JRA55Snapshots, radiation: JRA55Snapshots | None = None) -> SurfaceCase:
    """A copy of `case` forced by the given sources: their grids and times, and for each field a
    read-only zero placeholder of the series' shape. The model reads the snapshots themselves
    from the sources passed as atmosphere_source / radiation_source (whole, or through
    windows)."""
    import dataclasses

    def placeholders(src):
        return {k: np.broadcast_to(np.zeros((), np.float32), src.shape) for k in src.short}

    out = dataclasses.replace(case, atmos_grid=atmosphere.grid(), atmos=placeholders(atmosphere),
                              atmos_times=atmosphere.times)
    if radiation is not None:
        out = dataclasses.replace(out, rad_grid=radiation.grid(), rad=placeholders(radiation),
                                  rad_times=radiation.times)
    return out
