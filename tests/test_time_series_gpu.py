"""Snapshot windows streamed from the host, and cyclical time indexing, on the GPU.

A model that keeps only a ring of N snapshots on the device (snapshots uploaded one at a
time by SnapshotStream, in stream order or on a copy stream, ahead of the steps that read them) must step bit
for bit like the same model holding the whole series, through the wrap of a cyclical series
and past the end of a clamped one. A cyclical model's fluxes inside the wrap interval match
the oracle's with the same indexing.
"""
import numpy as np
import pytest
import torch

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip._lib import NEError
from numericalearth_hip.ensemble import Ensemble
from numericalearth_hip.generators import make_case
from numericalearth_hip.model import FLUX_NAMES, NET_NAMES, OceanOnlyModel, stream_handle
from numericalearth_hip.params import to_params
from numericalearth_hip.time_series import Clamp, Cyclical, NpySnapshotStore
from test_gpu_parity import _compare_fluxes_on_gpu_state

pytestmark = pytest.mark.gpu

SHAPE = (48, 24, 6)
HOUR = 3600.0


def _case(n_times, seed=11):
    case = make_case(*SHAPE, seed=seed, n_times=n_times, dt_times=HOUR)
    print(f"[case] shape={SHAPE} seed={seed} n_times={n_times}", flush=True)
    return case


def _fields(m):
    out = {"T": m.T, "S": m.S}
    out.update({f"flux.{k}": v for k, v in m.fluxes.items()})
    out.update({f"net.{k}": v for k, v in m.net.items()})
    return out


def _assert_same(a, b, when):
    fa, fb = _fields(a), _fields(b)
    for k in fa:
        assert torch.equal(fa[k], fb[k]), f"{k} differs {when}"


@pytest.mark.parametrize("precision,indexing,aw,rw,n_times,copy_stream", [
    ("fp64", Clamp(), 2, 3, 5, False),
    ("fp64", Cyclical(), 2, 2, 5, False),       # odd Nt, ring of two: the wrap reads positions Nt-1 and Nt
    ("mixed", Cyclical(5.5 * HOUR), 3, 4, 5, False),
    ("fp64", Cyclical(), 4, None, 6, False),    # radiation held whole
    ("fp64", Cyclical(), 3, 3, 5, True),        # copies on their own stream, ordered by events
])
def test_streamed_window_steps_like_resident_series(require_gpu, precision, indexing, aw, rw, n_times, copy_stream):
    print(f"[window] precision={precision} indexing={indexing} atmosphere={aw} radiation={rw} "
          f"copy_stream={copy_stream}", flush=True)
    case = _case(n_times)
    full = OceanOnlyModel(case, precision=precision, time_indexing=indexing)
    win = OceanOnlyModel(case, precision=precision, time_indexing=indexing, atmosphere_window=aw,
                         radiation_window=rw, copy_stream=copy_stream)
    assert win.atm["T"].shape[0] == aw and full.atm["T"].shape[0] == n_times
    _assert_same(win, full, "after construction")
    for n in range(30):                          # 12.5 h: more than two cycles of the cyclical series
        full.time_step(1500.0)
        win.time_step(1500.0)
        _assert_same(win, full, f"after step {n + 1}")
    assert win.clock == full.clock
    if isinstance(indexing, Clamp):              # each snapshot after the first window once, then held
        assert win.streams["atmosphere"].uploads == n_times - aw
    else:
        assert win.streams["atmosphere"].uploads >= 8
    full.close()
    win.close()


class _SlowSource:
    """A snapshot source whose reads take a while (a file system under load)."""

    def __init__(self, inner, delay=0.002):
        self.inner, self.delay = inner, delay
        self.n_times, self.times, self.shape = inner.n_times, inner.times, inner.shape

    def read(self, n, name, out):
        import time
        time.sleep(self.delay)
        return self.inner.read(n, name, out)


def test_stream_survives_skipped_positions(require_gpu):
    """Steps three snapshot intervals long skip positions the read-ahead had already requested:
    their staging buffers are given back (before, once a handful of positions had been skipped no
    buffer was left free and the next request failed)."""
    from numericalearth_hip.time_series import ArraySnapshots
    case = _case(9, seed=13)
    src = _SlowSource(ArraySnapshots(case.atmos, case.atmos_times))
    full = OceanOnlyModel(case, time_indexing=Cyclical())
    win = OceanOnlyModel(case, time_indexing=Cyclical(), atmosphere_window=3, atmosphere_source=src)
    st = win.streams["atmosphere"]
    print(f"[window] dt = 3 snapshot intervals, staging buffers {st.n_staging}", flush=True)
    for n in range(24):
        full.time_step(3 * HOUR)
        win.time_step(3 * HOUR)
        _assert_same(win, full, f"after step {n + 1}")
        assert len(st.free) + len(st.reads) == st.n_staging
        assert len(st.reads) <= st.read_ahead, f"stale reads held after step {n + 1}: {sorted(st.reads)}"
    full.close()
    win.close()


def test_window_streams_from_an_npy_store(require_gpu, tmp_path):
    case = _case(7, seed=12)
    atm = NpySnapshotStore(NpySnapshotStore.write(tmp_path / "atm", case.atmos, case.atmos_times), tuple(case.atmos))
    rad = NpySnapshotStore(NpySnapshotStore.write(tmp_path / "rad", case.rad, case.rad_times), tuple(case.rad))
    full = OceanOnlyModel(case, time_indexing=Cyclical())
    win = OceanOnlyModel(case, time_indexing=Cyclical(), atmosphere_window=3, radiation_window=2,
                         atmosphere_source=atm, radiation_source=rad, lookahead=1)
    for n in range(20):
        full.time_step(2400.0)
        win.time_step(2400.0)
    torch.cuda.synchronize()
    _assert_same(win, full, "after 20 steps")
    full.close()
    win.close()


def test_step_without_its_snapshots_is_refused(require_gpu):
    case = _case(5)
    m = OceanOnlyModel(case, atmosphere_window=2, radiation_window=2)
    T0 = m.T.clone()
    # straight through the C ABI, without the stream's upload: the step would end between
    # snapshots 1 and 2, and slot 0 still holds snapshot 0
    st = m.lib.ne_model_time_step(m.handle, 1.5 * HOUR, stream_handle())
    assert st == abi.NE_ERR_STATE
    assert m.clock == (0.0, 0)
    assert torch.equal(m.T, T0)
    assert m.lib.ne_model_snapshot_resident(m.handle, 2, 0) == abi.NE_ERR_SHAPE
    assert m.lib.ne_model_snapshot_resident(m.handle, 0, 9) == abi.NE_ERR_SHAPE     # past a clamped series
    m.time_step(1.5 * HOUR)                      # the model's own step uploads them first
    assert m.clock == (1.5 * HOUR, 1)
    resident = OceanOnlyModel(case)
    assert resident.lib.ne_model_snapshot_resident(resident.handle, 0, 0) == abi.NE_ERR_STATE
    with pytest.raises(NEError) as e:
        Ensemble([m])
    assert e.value.status == abi.NE_ERR_UNSUPPORTED
    m.close()
    resident.close()


def test_cyclical_model_matches_oracle_inside_the_wrap(require_gpu):
    case = _case(3, seed=13)
    p = to_params()
    gm = OceanOnlyModel(case, params=abi.Params.from_buffer_copy(p), time_indexing=Cyclical(),
                        atmosphere_window=2, radiation_window=2)
    period = 3 * HOUR
    for _ in range(5):                           # t = 2.5 h: between the last snapshot and the first
        gm.time_step(1800.0)
    assert gm.snapshots_at(gm.clock[0]) == (2, 3, 2, 3)
    om = O.OracleModel(case, params=abi.Params.from_buffer_copy(p), time_indexing=abi.NE_TIME_CYCLICAL,
                       time_period=period)
    om.time = gm.clock[0]
    _compare_fluxes_on_gpu_state(gm, om, envelope=True)
    # one period later the same snapshots are read, a cycle further on
    assert gm.snapshots_at(gm.clock[0] + period) == (5, 6, 5, 6)
    gm.close()
