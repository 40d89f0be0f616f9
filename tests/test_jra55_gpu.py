"""Models forced from JRA55-layout files on the GPU: the whole series read from the files, and
a ring of snapshots streamed from them, must step bit for bit like a model holding the same
(float32-stored) snapshots as arrays on the grid the files' bounds define; that grid's
fractional indices equal the oracle's."""
import dataclasses

import numpy as np
import pytest
import torch

import oracle as O
from jra55_files import write_jra55_files
from helpers import rel_err
from numericalearth_hip.generators import NE_BOUNDED, make_case
from numericalearth_hip.jra55 import ATMOSPHERE_VARIABLES, RADIATION_VARIABLES, JRA55Snapshots, with_forcing
from numericalearth_hip.model import OceanOnlyModel
from numericalearth_hip.time_series import Cyclical
from test_time_series_gpu import _assert_same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_models_forced_from_files_step_alike(require_gpu, tmp_path, precision):
    case = make_case(64, 32, 6, seed=17, n_times=6, dt_times=10800.0)
    print(f"[jra55] files 32x16x6, ocean (64, 32, 6), precision={precision}", flush=True)
    files = {s: case.atmos[k] for k, s in ATMOSPHERE_VARIABLES.items()}
    files.update({s: case.rad[k] for k, s in RADIATION_VARIABLES.items()})
    write_jra55_files(tmp_path, files, 32, 16)
    atm = JRA55Snapshots(tmp_path, ATMOSPHERE_VARIABLES)
    rad = JRA55Snapshots(tmp_path, RADIATION_VARIABLES)
    forced = with_forcing(case, atm, rad)
    f32 = lambda d: {k: v.astype(np.float32).astype(np.float64) for k, v in d.items()}  # noqa: E731
    arrays = dataclasses.replace(forced, atmos=f32(case.atmos), rad=f32(case.rad))
    kw = dict(precision=precision, time_indexing=Cyclical())
    ref = OceanOnlyModel(arrays, **kw)
    whole = OceanOnlyModel(forced, atmosphere_source=atm, radiation_source=rad, **kw)
    ring = OceanOnlyModel(forced, atmosphere_source=atm, radiation_source=rad, atmosphere_window=3,
                          radiation_window=2, **kw)
    # the grid from the files' bounds: fractional indices as the oracle computes them
    om = O.OracleModel(arrays)
    torch.cuda.synchronize()
    for a, b in ((ref.afi, om.afi), (ref.afj, om.afj), (ref.rfi, om.rfi), (ref.rfj, om.rfj)):
        assert np.array_equal(a.cpu().numpy(), b)
    _assert_same(whole, ref, "after construction")
    _assert_same(ring, ref, "after construction")
    for n in range(24):                          # a day of hourly steps: through the 18-hour record's wrap
        for m in (ref, whole, ring):
            m.time_step(3600.0)
        _assert_same(whole, ref, f"after step {n + 1}")
        _assert_same(ring, ref, f"after step {n + 1}")
    assert ring.streams["atmosphere"].uploads > 0
    for m in (ref, whole, ring):
        m.close()
    atm.close()
    rad.close()


def test_regional_model_reads_only_its_box(require_gpu, tmp_path):
    """A regional model across 0° forced from a bounding-box region of the files steps like the
    same model forced from the whole record. The sub-grid's fractional indices differ from the
    whole grid's by an integer offset, so the interpolation weights agree only to rounding: each
    field must agree to 1e-12 of its largest value, and elementwise to 1e-9 (a near-zero stress
    amplifies the rounding's relative size)."""
    glob = make_case(64, 32, 2, seed=19, n_times=4, dt_times=10800.0)
    files = {s: glob.atmos[k] for k, s in ATMOSPHERE_VARIABLES.items()}
    files.update({s: glob.rad[k] for k, s in RADIATION_VARIABLES.items()})
    write_jra55_files(tmp_path, files, 32, 16)
    ocean = make_case(40, 24, 4, seed=31, longitude=(-30.0, 40.0), latitude=(-20.0, 30.0), topology_x=NE_BOUNDED)
    box = ((-30.0, 40.0), (-20.0, 30.0))
    print(f"[jra55] regional ocean (40, 24, 4) in box {box}", flush=True)
    srcs = {r: (JRA55Snapshots(tmp_path, ATMOSPHERE_VARIABLES, region=r),
                JRA55Snapshots(tmp_path, RADIATION_VARIABLES, region=r)) for r in (None, box)}
    assert srcs[box][0].shape[1:] < srcs[None][0].shape[1:]
    models = {r: OceanOnlyModel(with_forcing(ocean, a, b), atmosphere_source=a, radiation_source=b,
                                time_indexing=Cyclical()) for r, (a, b) in srcs.items()}
    full, sub = models[None], models[box]
    for step in range(4):
        torch.cuda.synchronize()
        same = (full.iterations == sub.iterations).float().mean().item()
        assert same >= 0.99, same
        ok = (full.iterations == sub.iterations).cpu().numpy()
        pairs = [(sub.fluxes[k].cpu().numpy()[ok], full.fluxes[k].cpu().numpy()[ok], k) for k in full.fluxes]
        pairs += [(sub.T.cpu().numpy(), full.T.cpu().numpy(), "T"), (sub.S.cpu().numpy(), full.S.cpu().numpy(), "S")]
        for a, b, k in pairs:
            elem, field = rel_err(a, b)
            assert field <= 1e-12 and elem <= 1e-9, (k, step, elem, field)
        full.time_step(3600.0)
        sub.time_step(3600.0)
        print(f"[jra55] step {step + 1}", flush=True)
    for m in models.values():
        m.close()
    for a, b in srcs.values():
        a.close()
        b.close()
