"""Small NetCDF-3 files in the JRA55-do repeat-year layout, written with scipy for the tests:
one variable per ``RYF.<short>.1990_1991.nc`` file, (time, lat, lon) float32, with lon/lat,
their bounds and a ``time`` axis in days since 1900-01-01 (3-hourly)."""
import numpy as np
from scipy.io import netcdf_file

from numericalearth_hip.generators import gaussian_latitude_bounds, gaussian_latitudes
from numericalearth_hip.jra55 import repeat_year_filename


def write_jra55_files(directory, fields: dict, n_lon: int, n_lat: int, north_first=False, scale=None):
    """fields: short name → array (Nt, n_lat, n_lon) south to north. Returns the lower bounds used."""
    dlon = 360.0 / n_lon
    lon = dlon * np.arange(n_lon)
    lon_b = np.stack([lon - 0.5 * dlon, lon + 0.5 * dlon], axis=1)
    lat = gaussian_latitudes(n_lat)
    b = gaussian_latitude_bounds(n_lat)
    lat_b = np.stack([b[:-1], b[1:]], axis=1)
    order = slice(None, None, -1) if north_first else slice(None)
    for short, data in fields.items():
        nt = data.shape[0]
        with netcdf_file(str(directory / repeat_year_filename(short)), "w", version=2) as f:
            f.createDimension("time", None)
            f.createDimension("lat", n_lat)
            f.createDimension("lon", n_lon)
            f.createDimension("bnds", 2)
            f.createVariable("lon", "f8", ("lon",))[:] = lon
            f.createVariable("lat", "f8", ("lat",))[:] = lat[order]
            f.createVariable("lon_bnds", "f8", ("lon", "bnds"))[:] = lon_b
            f.createVariable("lat_bnds", "f8", ("lat", "bnds"))[:] = lat_b[order][:, ::-1] if north_first else lat_b
            t = f.createVariable("time", "f8", ("time",))
            t.units = "days since 1900-01-01 00:00:00"
            t[:] = 32872.0 + np.arange(nt) / 8.0
            v = f.createVariable(short, "f4", ("time", "lat", "lon"))
            if scale is not None:
                v.scale_factor, v.add_offset = scale
                v[:] = ((data - scale[1]) / scale[0]).astype(np.float32)[:, order]
            else:
                v[:] = data.astype(np.float32)[:, order]
    return lon_b[:, 0], b[:-1]
