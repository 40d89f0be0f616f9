"""Snapshot time indexing (Clamp / Cyclical) and snapshot stores, on the CPU.

The oracle's time indices are pinned against numpy's own interpolation (np.interp clamps at
the ends as Clamp() does, and with `period=` interpolates on the circle as Cyclical(period)
does), against exact dyadic periodicity, and against the wrap's end points; the library's
ne_snapshot_time_indices (host code, no device) must then reproduce the oracle bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from numericalearth_hip import _ctypes_abi as abi
from numericalearth_hip._lib import load
from numericalearth_hip.time_series import (ArraySnapshots, Clamp, Cyclical, NpySnapshotStore, initial_window,
                                            interpolating_time_indices, resolve_time_indexing)

CYC = abi.NE_TIME_CYCLICAL
CLAMP = abi.NE_TIME_CLAMPED


def _value(times, f, t, indexing=CLAMP, period=None):
    n1, n2, w = O.time_indices(times, t, indexing, period)
    return (1.0 - w) * f[n1] + w * f[n2]


def _series(rng, n, uniform=False, t0=0.0):
    if uniform:
        return t0 + 3600.0 * np.arange(n)
    return t0 + np.cumsum(rng.uniform(600.0, 7200.0, n)) - 600.0


@pytest.mark.parametrize("uniform", [True, False])
def test_clamped_oracle_is_numpy_interp(uniform):
    rng = np.random.default_rng(1)
    times = _series(rng, 9, uniform)
    f = rng.standard_normal(9)
    for t in np.concatenate([rng.uniform(times[0] - 5e3, times[-1] + 5e3, 400), times]):
        assert abs(_value(times, f, t) - np.interp(t, times, f)) <= 1e-14 * max(1.0, np.max(np.abs(f)))


@pytest.mark.parametrize("n,t0,extra", [(5, 0.0, 3600.0), (8, 7200.0, 1800.0), (3, -5e4, 9000.0)])
def test_cyclical_oracle_is_numpy_periodic_interp(n, t0, extra):
    rng = np.random.default_rng(n)
    times = _series(rng, n, uniform=False, t0=t0)
    period = times[-1] - times[0] + extra      # the last-to-first interval is `extra` long
    f = rng.standard_normal(n)
    ts = np.concatenate([rng.uniform(times[0] - 3 * period, times[0] + 4 * period, 800), times, times + period])
    for t in ts:
        want = np.interp(t, times, f, period=period)
        assert abs(_value(times, f, t, CYC, period) - want) <= 1e-12, t


def test_cyclical_wrap_end_points_and_positions():
    times = np.array([0.0, 0.25, 0.75, 1.0])    # dyadic: every sum below is exact
    period = 1.5
    f = np.array([1.0, 2.0, 4.0, 8.0])
    # inside the wrap interval (1.0, 1.5): from the last snapshot to the first one of the next cycle
    assert O.time_indices(times, 1.25, CYC, period) == (3, 0, 0.5)
    assert O.snapshot_positions(times, 1.25, CYC, period) == (3, 4)
    assert _value(times, f, 1.25, CYC, period) == 4.5
    assert _value(times, f, 1.0 + 2 ** -20, CYC, period) == 8.0 - 7.0 * 2 ** -19    # just past the last snapshot
    # at a period past the first snapshot the weight is back on snapshot 0 (continuity of the wrap)
    assert O.time_indices(times, 1.5, CYC, period) == (0, 0, 0.0)
    assert O.snapshot_positions(times, 1.5, CYC, period) == (4, 4)
    assert _value(times, f, 1.5 - 2 ** -40, CYC, period) == pytest.approx(f[0] + 7 * 2 ** -39, abs=1e-15)
    # before the first snapshot: the previous cycle's wrap
    assert O.time_indices(times, -0.25, CYC, period) == (3, 0, 0.5)
    assert O.snapshot_positions(times, -0.25, CYC, period) == (-1, 0)
    # k periods later: the same indices and weight, positions k·Nt further on
    for t in (0.125, 0.5, 0.875, 1.0, 1.25, 1.375):
        base_i, base_p = O.time_indices(times, t, CYC, period), O.snapshot_positions(times, t, CYC, period)
        for k in (-3, -1, 1, 2, 7):
            assert O.time_indices(times, t + k * period, CYC, period) == base_i
            assert O.snapshot_positions(times, t + k * period, CYC, period) == (base_p[0] + 4 * k, base_p[1] + 4 * k)


def test_cyclical_with_period_of_the_span_never_reads_across_the_wrap():
    times = np.array([0.0, 1.0, 2.0])
    for t in np.linspace(-6.0, 6.0, 97):
        n1, n2, w = O.time_indices(times, t, CYC, 2.0)
        assert n2 >= n1 and 0.0 <= w <= 1.0


def _c_indices(times, t, kind, period):
    lib = load()
    ti = abi.TimeInterp()
    pos = (C.c_int64 * 2)()
    times = np.ascontiguousarray(times, dtype=np.float64)
    st = lib.ne_snapshot_time_indices(times.ctypes.data_as(C.POINTER(C.c_double)), len(times), kind, period, t,
                                      C.byref(ti), pos)
    return st, (ti.n1, ti.n2, ti.tilde_n), (pos[0], pos[1])


@pytest.mark.parametrize("kind", [CLAMP, CYC])
def test_library_time_indices_match_oracle_bit_for_bit(kind):
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 7, 2920):
        times = _series(rng, n, uniform=(n == 2920), t0=rng.uniform(-1e5, 1e5))
        period = float(times[-1] - times[0] + rng.uniform(100.0, 9000.0)) if kind == CYC else 0.0
        span = max(period, times[-1] - times[0], 1.0)
        ts = np.concatenate([rng.uniform(times[0] - 3 * span, times[0] + 5 * span, 300), times[:50],
                             times[:50] + period, [times[0] + period, times[-1]]])
        for t in ts:
            st, got, pos = _c_indices(times, float(t), kind, period)
            assert st == abi.NE_OK
            want = O.time_indices(times, float(t), kind, period if kind == CYC else None)
            assert got == want, (n, t, got, want)
            assert pos == O.snapshot_positions(times, float(t), kind, period if kind == CYC else None)


def test_library_time_indices_reject_bad_series():
    times = np.array([0.0, 10.0, 20.0])
    assert _c_indices(times, 1.0, CYC, 15.0)[0] == abi.NE_ERR_SHAPE       # shorter than the series
    assert _c_indices(times, 1.0, CYC, -30.0)[0] == abi.NE_ERR_SHAPE
    assert _c_indices(times, 1.0, 7, 30.0)[0] == abi.NE_ERR_UNSUPPORTED
    assert _c_indices(np.zeros(0), 1.0, CLAMP, 0.0)[0] == abi.NE_ERR_SHAPE


def test_cyclical_period_inference_matches_repeat_year_forcing():
    # 2920 three-hourly snapshots of the repeat-year JRA55 forcing: Cyclical() cycles over 365 days
    times = 10800.0 * np.arange(2920)
    assert resolve_time_indexing(Cyclical(), times) == (CYC, 365 * 86400.0)
    assert resolve_time_indexing(Cyclical(100.0), times[:2]) == (CYC, 100.0)
    assert resolve_time_indexing(None, times) == (CLAMP, 0.0)
    assert resolve_time_indexing(Clamp(), times) == (CLAMP, 0.0)
    with pytest.raises(ValueError):
        resolve_time_indexing(Cyclical(), times[:1])


def test_windows_and_the_wrapped_snapshot():
    nt = 10
    times = 3600.0 * np.arange(nt)
    assert initial_window(times, 4) == [0, 1, 2, 3]
    assert initial_window(times, 4, time=8.5 * 3600.0) == [6, 7, 8, 9]          # moved inside the series
    assert initial_window(times, 20) == list(range(10))
    assert initial_window(times, 4, Cyclical(), time=9.5 * 3600.0) == [9, 10, 11, 12]
    assert initial_window(times, 3, Cyclical(), time=-0.5 * 3600.0) == [-1, 0, 1]
    # a window starting two before the end of the record holds (Nt-2, Nt-1, Nt, 1) in 1-based
    # indices, and its last slot is the first snapshot again
    f = {"x": np.arange(nt, dtype=np.float64)[:, None, None] * np.ones((nt, 2, 3))}
    src = ArraySnapshots(f, times)
    window = [p % nt for p in range(nt - 3, nt + 1)]
    assert [w + 1 for w in window] == [nt - 2, nt - 1, nt, 1]
    a, b = np.empty((2, 3)), np.empty((2, 3))
    src.read(window[-1], "x", a)
    src.read(0, "x", b)
    assert np.array_equal(a, b)


def test_npy_store_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    fields = {k: rng.standard_normal((6, 4, 5)) for k in ("u", "v", "T")}
    times = 600.0 * np.arange(6)
    NpySnapshotStore.write(tmp_path / "atm", fields, times)
    store = NpySnapshotStore(tmp_path / "atm", ("u", "v", "T"))
    assert store.n_times == 6 and store.shape == (6, 4, 5)
    out32 = np.empty((4, 5), np.float32)
    for n in range(6):
        for k in fields:
            out = np.empty((4, 5))
            store.read(n, k, out)
            assert np.array_equal(out, fields[k][n])
            store.read(n, k, out32)
            assert np.array_equal(out32, fields[k][n].astype(np.float32))
    with pytest.raises(ValueError):
        ArraySnapshots({"u": np.zeros((3, 2, 2)), "v": np.zeros((4, 2, 2))}, np.arange(3.0))
    with pytest.raises(ValueError):
        ArraySnapshots({"u": np.zeros((3, 2, 2))}, np.arange(4.0))


def test_interpolating_time_indices_wrapper():
    times = np.array([0.0, 0.25, 0.75, 1.0])
    ti, pos = interpolating_time_indices(times, 1.25, Cyclical(1.5))
    assert (ti.n1, ti.n2, ti.tilde_n) == (3, 0, 0.5) and pos == (3, 4)
    ti, pos = interpolating_time_indices(times, 5.0)
    assert (ti.n1, ti.n2, ti.tilde_n) == (3, 3, 0.0) and pos == (3, 3)
