"""Request serving with continuous batching on one MI355X.

A request is an `steps`-step solve (``time_step!`` repeated) of a model of its own: a grid,
a prescribed atmosphere and an initial ocean state, in FP64 or in the mixed form. The server

* admits waiting requests, first come first served, while the device bytes of every resident
  request stay within an HBM budget and their cells within a work budget (`AdmissionQueue`):
  the work budget bounds the size of a batch step, so that under overload requests wait
  for a slot instead of all sharing ever slower steps;
* uploads an admitted request's state on an upload stream and creates its model there
  (``ne_model_create``, which also runs the construction-time flux solve);
* keeps one batch per (precision, Δt) lane (``ne_batch``) on a stream of its own; a request
  joins its lane at the next batch step once its upload has completed, every batch step
  advances all members by one ``time_step!``, and a request leaves its lane after its last
  step, when its results are copied to host memory on the lane's stream;
* keeps at most two steps of each lane in flight, so that a request admitted now joins
  within two steps, and the FP64 and mixed lanes overlap on the device.

A request's results do not depend on what else is being served: every member of a batch
step computes exactly what stepping its model alone computes (the kernels share the
per-tile and per-column code), whichever members share the step.

``serve_plain`` is the per-request path the server is compared with: requests one after
another, each with its own model stepped by ``ne_model_time_step``.
"""
from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from .generators import SurfaceCase
from .model import OceanOnlyModel

RESULT_FIELDS = ("T", "S", "sensible_heat", "latent_heat", "x_momentum", "y_momentum")
SURFACE_RESULT_FIELDS = ("SST", "sensible_heat", "latent_heat")


def _result_sources(model: OceanOnlyModel) -> dict:
    src = {"T": model.T, "S": model.S, "SST": model.T[-1]}
    src.update(model.fluxes)
    return src


def request_bytes(shape, precision: str, n_times: int = 1, atmos_coarsening: int = 2) -> int:
    """Device bytes of one resident request: T, S and the four RK stage buffers, the face
    velocities, the interface/net/radiative fluxes and the other 2-D fields, the prescribed
    atmosphere and radiation, and the grid metrics."""
    Nx, Ny, Nz = shape
    s3 = 8 if precision == "fp64" else 4
    n3 = Nx * Ny * Nz
    cells3 = 6 * n3 + Nz * Ny * (Nx + 1) + Nz * (Ny + 1) * Nx + (Nz + 1) * Ny * Nx
    fields2 = 9 + 6 + 3 + 1                       # interface, net, radiative fluxes, frazil heat (FP64)
    two_d = fields2 * 8 * Nx * Ny + 4 * Nx * Ny * 2 + 8 * Nx * Ny     # + iterations, kmin, bottom height
    Nxa, Nya = max(1, Nx // atmos_coarsening), max(1, Ny // atmos_coarsening)
    atmos = 9 * s3 * Nxa * Nya * n_times
    metrics = 8 * (4 * Nx + 6 * Ny + 4 * Nz + 8)
    return cells3 * s3 + two_d + atmos + metrics


@dataclass
class Request:
    id: int
    case: SurfaceCase
    precision: str = "fp64"
    steps: int = 100
    dt: float = 1200.0
    arrival: float = 0.0                 # seconds on the serving clock
    params: object = None
    # filled in by the server
    nbytes: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    result: dict = field(default_factory=dict)

    @property
    def cells(self) -> int:
        Nx, Ny, Nz = self.case.shape
        return Nx * Ny * Nz

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class AdmissionQueue:
    """First-come-first-served admission under a byte budget and a cell (work) budget (host
    logic only). A request larger than the cell budget is admitted when nothing else is
    resident."""

    def __init__(self, budget_bytes: float, cell_budget: float = math.inf):
        self.budget = float(budget_bytes)
        self.cell_budget = float(cell_budget)
        self.resident = 0.0
        self.resident_cells = 0
        self.waiting: deque = deque()

    def offer(self, req: Request) -> None:
        if req.nbytes > self.budget:
            raise ValueError(f"request {req.id} needs {req.nbytes} bytes, more than the whole budget")
        self.waiting.append(req)

    def _fits(self, r: Request) -> bool:
        if self.resident + r.nbytes > self.budget:
            return False
        return self.resident_cells == 0 or self.resident_cells + r.cells <= self.cell_budget

    def admit(self) -> list:
        out = []
        while self.waiting and self._fits(self.waiting[0]):
            r = self.waiting.popleft()
            self.resident += r.nbytes
            self.resident_cells += r.cells
            out.append(r)
        return out

    def release(self, req: Request) -> None:
        self.resident -= req.nbytes
        self.resident_cells -= req.cells


class _Lane:
    def __init__(self, precision: str, dt: float, device: torch.device):
        from .ensemble import Ensemble
        self.precision, self.dt = precision, dt
        self.stream = torch.cuda.Stream(device=device)
        self.ens = Ensemble([], stream=self.stream, precision=precision, device=device.index)
        self.members: dict = {}          # request id → [request, model, steps done]
        self.joining: list = []          # (request, model, upload event)
        self.inflight: deque = deque()   # events of enqueued steps

    def busy(self) -> bool:
        return bool(self.members or self.joining)


class Server:
    def __init__(self, budget_bytes: float = 200e9, cell_budget: float = 32e6, device: int = 0,
                 max_inflight: int = 2, result_fields=RESULT_FIELDS):
        self.device = torch.device("cuda", device)
        self.queue = AdmissionQueue(budget_bytes, cell_budget)
        self.upload = torch.cuda.Stream(device=self.device)
        self.lanes: dict = {}
        self.finishing: list = []        # (request, model, lane, event)
        self.done: list = []
        self.max_inflight = max_inflight
        self.result_fields = result_fields
        self.t0 = None
        self.steps_issued = 0

    # clock ------------------------------------------------------------------------
    def now(self) -> float:
        return time.perf_counter() - self.t0

    def start(self) -> None:
        self.t0 = time.perf_counter()

    # request flow ---------------------------------------------------------------------
    def submit(self, req: Request) -> None:
        if not req.nbytes:
            req.nbytes = request_bytes(req.case.shape, req.precision, len(req.case.atmos_times))
        self.queue.offer(req)

    def _lane(self, req: Request) -> _Lane:
        key = (req.precision, req.dt)
        if key not in self.lanes:
            self.lanes[key] = _Lane(req.precision, req.dt, self.device)
        return self.lanes[key]

    def _admit(self) -> None:
        for req in self.queue.admit():
            req.admitted = self.now()
            with torch.cuda.stream(self.upload):
                model = OceanOnlyModel(req.case, params=req.params, precision=req.precision,
                                       device=self.device.index)
                ev = torch.cuda.Event()
                ev.record(self.upload)
            self._lane(req).joining.append((req, model, ev))

    def _step_lane(self, lane: _Lane) -> bool:
        """Issue one batch step of the lane if it has room; True if a step was issued."""
        while lane.inflight and lane.inflight[0].query():
            lane.inflight.popleft()
        if len(lane.inflight) >= self.max_inflight:
            return False
        still = []
        for req, model, ev in lane.joining:
            if ev.query():
                lane.stream.wait_event(ev)
                lane.ens.add(model)
                lane.members[req.id] = [req, model, 0]
            else:
                still.append((req, model, ev))
        lane.joining = still
        if not lane.members:
            return False
        lane.ens.time_step(lane.dt, stream=lane.stream)
        self.steps_issued += 1
        finished = []
        for rid, entry in lane.members.items():
            entry[2] += 1
            if entry[2] >= entry[0].steps:
                finished.append(rid)
        for rid in finished:
            req, model, _ = lane.members.pop(rid)
            lane.ens.remove(model)
            with torch.cuda.stream(lane.stream):
                req.result = self._copy_results(model)
                ev = torch.cuda.Event()
                ev.record(lane.stream)
            self.finishing.append((req, model, lane, ev))
        ev = torch.cuda.Event()
        ev.record(lane.stream)
        lane.inflight.append(ev)
        return True

    def _copy_results(self, model: OceanOnlyModel) -> dict:
        src = _result_sources(model)
        out = {}
        for k in self.result_fields:
            t = src[k]
            h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
            h.copy_(t, non_blocking=True)
            out[k] = h
        return out

    def _retire(self) -> None:
        still = []
        for req, model, lane, ev in self.finishing:
            if ev.query():
                req.finished = self.now()
                req.result = {k: v.numpy() for k, v in req.result.items()}
                model.close()
                self.queue.release(req)
                self.done.append(req)
            else:
                still.append((req, model, lane, ev))
        self.finishing = still

    def busy(self) -> bool:
        return bool(self.queue.waiting or self.finishing or any(l.busy() for l in self.lanes.values()))

    def poll(self) -> bool:
        """One scheduler iteration; True if it issued device work."""
        self._retire()
        self._admit()
        issued = False
        for lane in self.lanes.values():
            issued |= self._step_lane(lane)
        return issued

    def run(self, requests: list, idle_sleep: float = 50e-6) -> list:
        """Serve `requests` (their `arrival` times on this server's clock, which starts now) and
        return them completed, in completion order."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self.start()
        while pending or self.busy():
            now = self.now()
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            if not self.poll() and not self.finishing:
                if pending and not self.busy():
                    wait = pending[0].arrival - self.now()
                    if wait > 0:
                        time.sleep(min(wait, 1e-3))
                else:
                    time.sleep(idle_sleep)
        torch.cuda.synchronize(self.device)
        return self.done

    def close(self) -> None:
        for lane in self.lanes.values():
            lane.ens.close()
        self.lanes.clear()


def serve_plain(requests: list, device: int = 0, result_fields=RESULT_FIELDS) -> list:
    """The per-request path: each request, in arrival order, gets its own model, stepped
    `steps` times with ne_model_time_step, then its results are copied back."""
    dev = torch.device("cuda", device)
    t0 = time.perf_counter()
    done = []
    for req in sorted(requests, key=lambda r: r.arrival):
        while time.perf_counter() - t0 < req.arrival:
            time.sleep(50e-6)
        req.admitted = time.perf_counter() - t0
        m = OceanOnlyModel(req.case, params=req.params, precision=req.precision, device=device)
        for _ in range(req.steps):
            m.time_step(req.dt)
        src = _result_sources(m)
        req.result = {k: src[k].cpu().numpy() for k in result_fields}
        torch.cuda.synchronize(dev)
        req.finished = time.perf_counter() - t0
        m.close()
        done.append(req)
    return done


def poisson_arrivals(rate: float, duration: float, rng: np.random.Generator) -> np.ndarray:
    """Arrival times of a Poisson process of `rate` per second over [0, duration)."""
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return np.asarray(out)
        out.append(t)


def latency_summary(done: list) -> dict:
    lat = np.asarray([r.latency for r in done])
    cells = sum(r.cells * r.steps for r in done)
    span = max(r.finished for r in done) - min(r.arrival for r in done)
    return {"requests": len(done), "p50_ms": 1e3 * float(np.percentile(lat, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat, 99)), "max_ms": 1e3 * float(lat.max()),
            "cell_updates_per_s": cells / span, "span_s": span}
