"""Snapshot time indexing of prescribed forcing, snapshot stores, and host→device streaming.

``Clamp`` and ``Cyclical`` name Oceananigans.OutputReaders' time indexing of a
FieldTimeSeries; the reference builds its JRA55 atmosphere and radiation series with
``Cyclical()`` (DataWrangling/JRA55/JRA55_prescribed_atmosphere.jl:17,
JRA55_prescribed_radiation.jl:14) and evaluates the indices on the host
(``cpu_interpolating_time_indices``, Atmospheres/interpolate_atmospheric_state.jl:57-60).
``interpolating_time_indices`` is that function, evaluated by the library
(``ne_snapshot_time_indices``) exactly as the models and the batched device clock evaluate it.

The reference keeps ``time_indices_in_memory`` snapshots of each series on the device (an
InMemory(N) backend) and refills the whole window through ``set!`` when the clock leaves it
(JRA55_field_time_series.jl:64-80). Here a window is a ring of N slots: the snapshot at stream
position p (the snapshot index, or cycle·Nt + index with cyclical indexing) fills slot p mod N,
and ``SnapshotStream`` uploads one snapshot at a time on a copy stream, ahead of the step that
first reads it, while the steps before it run.
"""
from __future__ import annotations

import ctypes as C
import os
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _ctypes_abi as abi
from ._lib import check, load


@dataclass(frozen=True)
class Clamp:
    """Linear between the two snapshots around t; the first / last snapshot outside the series."""


@dataclass(frozen=True)
class Cyclical:
    """Oceananigans Cyclical(period): t is taken modulo the period from the first snapshot time,
    and between the last snapshot and the first one a period later the two are interpolated.
    ``period=None`` is Cyclical() with the period inferred from the times as Oceananigans infers
    it: last − first plus the last interval."""

    period: float | None = None


def resolve_time_indexing(indexing, times) -> tuple[int, float]:
    """(NE_TIME_*, period) of a Clamp / Cyclical object (None is Clamp())."""
    if indexing is None or isinstance(indexing, Clamp):
        return abi.NE_TIME_CLAMPED, 0.0
    if isinstance(indexing, Cyclical):
        period = indexing.period
        if period is None:
            times = np.asarray(times, dtype=np.float64)
            if len(times) < 2:
                raise ValueError("Cyclical() infers its period from at least two snapshot times")
            period = float(times[-1] - times[0] + (times[-1] - times[-2]))
        return abi.NE_TIME_CYCLICAL, float(period)
    raise TypeError(f"time indexing must be Clamp() or Cyclical(period), not {indexing!r}")


def interpolating_time_indices(times, t: float, indexing=None) -> tuple[abi.TimeInterp, tuple[int, int]]:
    """The two snapshots read at time t, the weight of the second, and their stream positions."""
    times = np.ascontiguousarray(times, dtype=np.float64)
    kind, period = resolve_time_indexing(indexing, times)
    ti = abi.TimeInterp()
    pos = (C.c_int64 * 2)()
    check(load().ne_snapshot_time_indices(times.ctypes.data_as(C.POINTER(C.c_double)), len(times), kind, period,
                                          float(t), C.byref(ti), pos), "ne_snapshot_time_indices")
    return ti, (pos[0], pos[1])


def initial_window(times, slots: int, indexing=None, time: float = 0.0) -> list[int]:
    """Stream positions a freshly created window of `slots` holds: N consecutive positions from the
    first one read at `time` (moved to fit inside the series with clamped indexing)."""
    nt = len(times)
    _, (start, _) = interpolating_time_indices(times, time, indexing)
    count = slots
    if resolve_time_indexing(indexing, times)[0] != abi.NE_TIME_CYCLICAL:
        start = max(0, min(start, nt - slots))
        count = min(slots, nt - start)
    return list(range(start, start + count))


# -- snapshot sources -----------------------------------------------------------------------------
class ArraySnapshots:
    """Snapshots of named fields held in host arrays shaped (Nt, Ny, Nx), i fastest."""

    def __init__(self, fields: dict, times):
        self.fields = fields
        self.times = np.ascontiguousarray(times, dtype=np.float64)
        shapes = {np.shape(a) for a in fields.values()}
        if len(shapes) != 1:
            raise ValueError(f"snapshot fields differ in shape: {shapes}")
        self.shape = shapes.pop()
        if self.shape[0] != len(self.times):
            raise ValueError(f"{self.shape[0]} snapshots but {len(self.times)} times")

    @property
    def n_times(self) -> int:
        return len(self.times)

    def read(self, n: int, name: str, out: np.ndarray) -> None:
        """Copy snapshot n of `name` into `out` (converted to out's dtype)."""
        np.copyto(out, self.fields[name][n], casting="same_kind")


class NpySnapshotStore(ArraySnapshots):
    """A directory of ``<field>.npy`` arrays (Nt, Ny, Nx) and ``times.npy``, memory-mapped read-only, so a
    series far larger than host memory is read one snapshot at a time."""

    def __init__(self, directory, names):
        d = Path(directory)
        fields = {k: np.load(d / f"{k}.npy", mmap_mode="r") for k in names}
        super().__init__(fields, np.load(d / "times.npy"))

    @staticmethod
    def write(directory, fields: dict, times) -> "Path":
        d = Path(directory)
        d.mkdir(parents=True, exist_ok=True)
        for k, a in fields.items():
            np.save(d / f"{k}.npy", np.ascontiguousarray(a))
        np.save(d / "times.npy", np.ascontiguousarray(times, dtype=np.float64))
        return d


# -- streaming ------------------------------------------------------------------------------------
class SnapshotStream:
    """Uploads the snapshots of one windowed series of a model into its ring of slots.

    ``prepare(positions, stream)`` is called before each step (or update_state) with the positions the
    step reads: those not resident are copied into their slots, in stream order on `stream` itself (the
    default), or with ``copy_stream=True`` on a stream of their own that first waits for the work already
    on `stream` (the slot may still be read there), `stream` then waiting for the copy. In-stream copies
    cost the step loop their own duration (a 7-field snapshot of 11 MB: about 0.25 ms); the cross-stream
    event waits have been measured to cost more than that on ROCm (single steps several milliseconds
    long), so the copy stream is an option, not the default. Snapshots are read
    from the source by worker threads (one task per field) into pinned staging buffers `read_ahead`
    positions ahead of the step that needs them; the next `lookahead` positions (at most slots − 2, so
    the two slots the next step reads are never overwritten) are copied as soon as they have been read,
    so the host's reads overlap the steps before them. The read_ahead + 2 pinned
    staging buffers are allocated once; a buffer whose copy is still queued is passed over, so the host
    waits for the device only when every buffer's copy is. Positions are marked resident in the model when
    their copy is enqueued; every step of the model must be prepared (OceanOnlyModel.time_step does it)."""

    def __init__(self, model, which: int, source: ArraySnapshots, device_fields: dict, resident: list[int],
                 lookahead: int | None = None, read_ahead: int | None = None, workers: int | None = None,
                 copy_stream: bool = False):
        self.model = model
        self.which = which
        self.source = source
        self.fields = device_fields
        self.names = tuple(device_fields)
        first = next(iter(device_fields.values()))
        self.slots = first.shape[0]
        self.device = first.device
        self.dtype = first.dtype
        self.cyclical = model.time_indexing_kind == abi.NE_TIME_CYCLICAL
        self.lookahead = max(0, self.slots - 2) if lookahead is None else min(int(lookahead), self.slots - 2)
        self.read_ahead = max(self.lookahead + 2, 2) if read_ahead is None else max(int(read_ahead), self.lookahead)
        self.copy_stream = torch.cuda.Stream(device=self.device) if copy_stream else None
        self.pool = ThreadPoolExecutor(max_workers=workers or min(len(self.names), default_workers()))
        self.slot_position = {int(p % self.slots): int(p) for p in resident}
        self.reads: dict[int, tuple[torch.Tensor, list[Future]]] = {}   # position -> staging buffer, field reads
        self.copies: dict[int, torch.cuda.Event] = {}                    # position -> event after its copy
        # pinned staging buffers, allocated here once (a pinned allocation of a snapshot takes
        # milliseconds of host time, which inside the step loop would let the device run dry)
        shape = (len(self.names),) + tuple(first.shape[1:])
        self.n_staging = self.read_ahead + 2
        self.free: list[tuple[torch.Tensor, torch.cuda.Event | None]] = [
            (torch.empty(shape, dtype=self.dtype, pin_memory=True), None) for _ in range(self.n_staging)]
        self.uploads = 0                        # snapshots copied to the device
        self.stalls = 0                         # copies that waited for their read

    def _valid(self, p: int) -> bool:
        return self.cyclical or 0 <= p < self.source.n_times

    def _staging(self) -> torch.Tensor:
        for k, (buf, ev) in enumerate(self.free):
            if ev is None or ev.query():        # its previous copy has left the buffer
                del self.free[k]
                return buf
        buf, ev = self.free.pop(0)              # every buffer's copy still queued: wait for the oldest
        ev.synchronize()
        return buf

    def _request(self, p: int) -> None:
        if p in self.reads or p in self.copies or self.slot_position.get(p % self.slots) == p:
            return
        buf = self._staging()
        host = buf.numpy()
        n = p % self.source.n_times
        futs = [self.pool.submit(self.source.read, n, name, host[k]) for k, name in enumerate(self.names)]
        self.reads[p] = (buf, futs)

    def _read_done(self, p: int) -> bool:
        return p in self.reads and all(f.done() for f in self.reads[p][1])

    def _copy(self, p: int, stream: torch.cuda.Stream) -> None:
        buf, futs = self.reads.pop(p)
        if not all(f.done() for f in futs):
            self.stalls += 1
        for f in futs:
            f.result()
        slot = p % self.slots
        cs = self.copy_stream or stream
        if cs is not stream:
            ready = torch.cuda.Event()
            ready.record(stream)                # what is enqueued so far may still read the slot
            cs.wait_event(ready)
        with torch.cuda.stream(cs):
            for k, name in enumerate(self.names):
                self.fields[name][slot].copy_(buf[k], non_blocking=True)
        done = torch.cuda.Event()
        done.record(cs)
        self.copies = {q: e for q, e in self.copies.items() if q % self.slots != slot}
        self.copies[p] = done
        self.free.append((buf, done))
        self.slot_position[slot] = p
        check(load().ne_model_snapshot_resident(self.model.handle, self.which, p), "ne_model_snapshot_resident")
        self.uploads += 1

    def _drop_stale(self, first: int, last: int) -> None:
        """Give back the staging buffers of read-ahead positions the clock will not read: those
        before `first` (a step longer than the snapshot interval, or a clock set back) or beyond
        the read-ahead window after `last`. Their reads are waited for, since the workers are
        still writing into the buffer."""
        for p in [q for q in self.reads if q < first or q > last + self.read_ahead]:
            buf, futs = self.reads.pop(p)
            for f in futs:
                f.cancel()
            for f in futs:
                if not f.cancelled():
                    f.exception()               # wait; a failed read of a dropped position is moot
            self.free.append((buf, None))

    def prepare(self, positions, stream: torch.cuda.Stream) -> None:
        """Make the positions read next resident and ordered before `stream`'s next work."""
        need = sorted({int(p) for p in positions})
        self._drop_stale(need[0], need[-1])
        for p in need:
            if self.slot_position.get(p % self.slots) != p:
                self._request(p)
                self._copy(p, stream)
        for p in need:
            ev = self.copies.pop(p, None)
            if ev is not None:
                stream.wait_event(ev)
        last = need[-1]
        for q in range(last + 1, last + 1 + self.read_ahead):
            if self._valid(q):
                self._request(q)
        for q in range(last + 1, last + 1 + self.lookahead):
            if self._read_done(q):
                self._copy(q, stream)

    def close(self) -> None:
        self.pool.shutdown(wait=True)
        self.reads.clear()


def default_workers() -> int:
    return max(1, min(4, (os.cpu_count() or 1) // 4))
