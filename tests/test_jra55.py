"""JRA55-do files as snapshot sources, on the CPU (NetCDF-3 files in the repeat-year layout,
written with scipy): snapshot slices in model order, repeat-year times, the grid from the
files' bounds, north-to-south storage, packed variables, bounding-box regions read as sub-blocks,
and a clear refusal of NetCDF-4 when no NetCDF-4 reader is installed."""
import numpy as np
import pytest

from jra55_files import write_jra55_files
from numericalearth_hip.generators import (NE_BOUNDED, NE_PERIODIC, gaussian_latitude_bounds, jra55_source_grid,
                                           make_case)
from numericalearth_hip.jra55 import (ATMOSPHERE_VARIABLES, RADIATION_VARIABLES, JRA55Snapshots, open_dataset,
                                      with_forcing)

NLON, NLAT, NT = 16, 8, 5


def _fields(rng, names):
    return {s: rng.standard_normal((NT, NLAT, NLON)).astype(np.float32) + 10.0 for s in names}


def test_slices_times_and_grid(tmp_path):
    rng = np.random.default_rng(1)
    data = _fields(rng, ATMOSPHERE_VARIABLES.values())
    write_jra55_files(tmp_path, data, NLON, NLAT)
    src = JRA55Snapshots(tmp_path, ATMOSPHERE_VARIABLES)
    assert src.shape == (NT, NLAT, NLON) and src.n_times == NT
    assert np.array_equal(src.times, 10800.0 * np.arange(NT))
    out = np.empty((NLAT, NLON))
    for name, short in ATMOSPHERE_VARIABLES.items():
        for n in range(NT):
            src.read(n, name, out)
            assert np.array_equal(out, data[short][n].astype(np.float64)), (name, n)
    # the files' own time axis (days since 1900, 3-hourly) gives the same seconds
    own = JRA55Snapshots(tmp_path, {"T": "tas"}, interval=None)
    assert np.allclose(own.times, 10800.0 * np.arange(NT), rtol=0, atol=1e-6)
    # the grid the reference builds: lower bounds closed by +360 and 90°, midpoint centres
    g = src.grid()
    lam, phi = g.centres()
    b = gaussian_latitude_bounds(NLAT)
    assert (g.Nx, g.Ny) == (NLON, NLAT) and g.irregular
    assert np.allclose(lam, 22.5 * np.arange(NLON), atol=1e-12)
    assert np.allclose(phi, 0.5 * (b[:-1] + b[1:]), atol=1e-12)
    # the same construction at JRA55's size is jra55_source_grid
    full = jra55_source_grid()
    assert np.allclose(full.centres()[0], 0.5625 * np.arange(640), atol=1e-12)
    src.close()
    own.close()


def test_north_to_south_storage_is_served_south_to_north(tmp_path):
    rng = np.random.default_rng(2)
    data = _fields(rng, ["tas"])
    (tmp_path / "ns").mkdir()
    (tmp_path / "sn").mkdir()
    write_jra55_files(tmp_path / "ns", data, NLON, NLAT, north_first=True)
    write_jra55_files(tmp_path / "sn", data, NLON, NLAT)
    ns = JRA55Snapshots(tmp_path / "ns", {"T": "tas"})
    sn = JRA55Snapshots(tmp_path / "sn", {"T": "tas"})
    a, b = np.empty((NLAT, NLON)), np.empty((NLAT, NLON))
    for n in range(NT):
        ns.read(n, "T", a)
        sn.read(n, "T", b)
        assert np.array_equal(a, b)
    ga, gb = ns.grid(), sn.grid()
    assert np.array_equal(ga.centres()[1], gb.centres()[1])
    ns.close()
    sn.close()


def test_packed_variable_is_unpacked(tmp_path):
    rng = np.random.default_rng(3)
    data = _fields(rng, ["psl"])
    data["psl"] = 1.0e5 + 100.0 * data["psl"]
    write_jra55_files(tmp_path, data, NLON, NLAT, scale=(0.5, 1.0e5))
    src = JRA55Snapshots(tmp_path, {"p": "psl"})
    out = np.empty((NLAT, NLON))
    src.read(2, "p", out)
    assert np.allclose(out, data["psl"][2], rtol=0, atol=0.26)
    src.close()


def test_netcdf4_without_reader_is_refused(tmp_path):
    pytest.importorskip("scipy")
    try:
        import netCDF4  # noqa: F401
        pytest.skip("a NetCDF-4 reader is installed")
    except ImportError:
        pass
    f = tmp_path / "RYF.tas.1990_1991.nc"
    f.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(OSError, match="NetCDF-4"):
        open_dataset(f)


def test_with_forcing_builds_the_case(tmp_path):
    rng = np.random.default_rng(4)
    write_jra55_files(tmp_path, _fields(rng, list(ATMOSPHERE_VARIABLES.values()) + list(RADIATION_VARIABLES.values())),
                      NLON, NLAT)
    atm = JRA55Snapshots(tmp_path, ATMOSPHERE_VARIABLES)
    rad = JRA55Snapshots(tmp_path, RADIATION_VARIABLES)
    case = with_forcing(make_case(24, 12, 3, seed=2), atm, rad)
    assert set(case.atmos) == set(ATMOSPHERE_VARIABLES) and set(case.rad) == set(RADIATION_VARIABLES)
    assert case.atmos["T"].shape == (NT, NLAT, NLON) and np.array_equal(case.atmos_times, atm.times)
    assert case.atmos_grid.irregular and case.rad_grid.irregular
    only_atm = with_forcing(make_case(24, 12, 3, seed=2), atm)
    assert not only_atm.rad_grid.irregular                     # radiation left as the case had it
    atm.close()
    rad.close()


def test_region_reads_a_sub_block_across_the_date_line(tmp_path):
    rng = np.random.default_rng(5)
    data = _fields(rng, ["tas"])
    write_jra55_files(tmp_path, data, NLON, NLAT)
    full = JRA55Snapshots(tmp_path, {"T": "tas"})
    box = ((-30.0, 40.0), (-20.0, 30.0))                         # across 0°, files in 0…360
    sub = JRA55Snapshots(tmp_path, {"T": "tas"}, region=box)
    g, gf = sub.grid(), full.grid()
    lam, phi = g.centres()
    lam_f, phi_f = gf.centres()
    assert g.topology_x == NE_BOUNDED and sub.shape == (NT, g.Ny, g.Nx) and g.Nx < NLON and g.Ny < NLAT
    # the box inside the hull of the kept centres, its longitudes in the box's convention
    assert lam[0] <= -30.0 < 40.0 <= lam[-1] and phi[0] <= -20.0 < 30.0 <= phi[-1]
    assert np.all(np.diff(lam) > 0)
    cols = np.round((lam % 360.0) / 22.5).astype(int) % NLON
    rows = np.searchsorted(phi_f, phi)
    assert np.allclose(phi_f[rows], phi, atol=1e-12) and np.allclose(lam_f[cols] % 360.0, lam % 360.0, atol=1e-12)
    a, b = np.empty((NLAT, NLON)), np.empty((g.Ny, g.Nx))
    for n in range(NT):
        full.read(n, "T", a)
        sub.read(n, "T", b)
        assert np.array_equal(b, a[np.ix_(rows, cols)])
    # a box that closes the circle keeps the whole record, periodic
    ring = JRA55Snapshots(tmp_path, {"T": "tas"}, region=((0.0, 360.0), (-20.0, 30.0)))
    assert ring.grid().topology_x == NE_PERIODIC and ring.grid().Nx == NLON
    with pytest.raises(ValueError):
        JRA55Snapshots(tmp_path, {"T": "tas"}, region=((10.0, 10.0), (0.0, 5.0)))
    for s in (full, sub, ring):
        s.close()


def test_region_cells_bracket_the_box_with_centres():
    """restrict_interfaces keeps the cells whose centres bracket [lo, hi] and no more: the first kept
    centre is at or below lo and the second above it, the last at or above hi and the one before below
    it; an edge on a face or in a cell's first half takes that neighbour, an edge past the centre does
    not; the ends of the vector clamp (metadata_field.jl:28-33, :47-56)."""
    from numericalearth_hip.jra55 import restrict_interfaces
    rng = np.random.default_rng(3)
    faces = np.cumsum(np.concatenate([[-80.0], rng.uniform(0.5, 2.0, 120)]))   # stretched, like JRA55's latitudes
    centres = 0.5 * (faces[1:] + faces[:-1])
    for _ in range(500):
        lo, hi = np.sort(rng.uniform(faces[2], faces[-3], 2))
        i0, i1 = restrict_interfaces(lo, hi, faces)
        c = centres[i0:i1]
        assert c[0] <= lo and c[-1] >= hi                       # the kept centres bracket the box
        assert len(c) == 1 or (c[1] > lo and c[-2] < hi)        # and no cell more than that needs
    k = 40
    mid = centres[k]
    assert restrict_interfaces(faces[k], mid + 0.1, faces)[0] == k - 1          # edge on a face: one more
    assert restrict_interfaces(0.5 * (faces[k] + mid), mid + 0.1, faces)[0] == k - 1   # first half: one more
    assert restrict_interfaces(mid + 1e-9, mid + 0.1, faces)[0] == k              # past the centre: none
    assert restrict_interfaces(mid, mid + 0.1, faces)[0] == k                     # on the centre: none
    assert restrict_interfaces(faces[0] - 5.0, faces[3], faces)[0] == 0           # clamped at the ends
    assert restrict_interfaces(faces[-4], faces[-1] + 5.0, faces)[1] == len(faces) - 1
    assert restrict_interfaces(-200.0, -150.0, faces) == (0, 1)                   # outside: one cell


def test_region_across_the_seam_continues_round_the_circle(tmp_path):
    """A box that crosses the files' first longitude keeps the cells on both sides of it (the
    reference's vector restriction would clamp at the record's end), bracketed the same way."""
    rng = np.random.default_rng(6)
    data = _fields(rng, ["tas"])
    write_jra55_files(tmp_path, data, NLON, NLAT)
    full = JRA55Snapshots(tmp_path, {"T": "tas"})
    sub = JRA55Snapshots(tmp_path, {"T": "tas"}, region=((350.0, 380.0), (-10.0, 10.0)))
    lam, _ = sub.grid().centres()
    assert lam[0] <= 350.0 and lam[-1] >= 380.0 and np.all(np.diff(lam) > 0)
    assert np.allclose(np.diff(lam), 360.0 / NLON)
    full.close()
    sub.close()
