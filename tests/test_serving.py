"""Serving: admission by HBM budget (host logic, CPU) and, on the GPU, continuous batching whose
results are the same for a request served alone, in a mix of requests of both precisions, and on
the plain per-request path."""
import copy

import numpy as np
import pytest

from numericalearth_hip.ensemble import draw_ensemble, member_case, member_params
from numericalearth_hip.serving import (AdmissionQueue, Request, latency_summary, poisson_arrivals,
                                        request_bytes)


class _Shape:
    def __init__(self, cells):
        self.shape = (cells, 1, 1)


def _req(i, nbytes, cells=1):
    r = Request(id=i, case=_Shape(cells))
    r.nbytes = nbytes
    return r


def test_admission_is_fifo_within_the_budget():
    q = AdmissionQueue(100)
    for i, b in enumerate([40, 50, 30, 5]):
        q.offer(_req(i, b))
    first = q.admit()
    assert [r.id for r in first] == [0, 1]          # 30 does not fit behind 40 + 50, and 5 waits its turn
    assert q.resident == 90
    q.release(first[0])
    assert [r.id for r in q.admit()] == [2, 3]
    assert q.resident == 85
    with pytest.raises(ValueError):
        q.offer(_req(9, 101))


def test_admission_work_budget_bounds_resident_cells():
    q = AdmissionQueue(1e12, cell_budget=100)
    for i, c in enumerate([60, 50, 30, 500]):
        q.offer(_req(i, 1, cells=c))
    assert [r.id for r in q.admit()] == [0]          # 60 + 50 > 100
    first = q.waiting[0]
    q.release(_req(0, 1, cells=60))
    assert [r.id for r in q.admit()] == [1, 2]       # 50 + 30 ≤ 100; 500 waits
    assert first.id == 1 and q.resident_cells == 80
    q.release(_req(1, 1, cells=50))
    q.release(_req(2, 1, cells=30))
    assert [r.id for r in q.admit()] == [3]          # larger than the budget: admitted alone


def test_request_bytes_grow_with_size_and_precision():
    small = request_bytes((1, 1, 50), "fp64")
    big = request_bytes((720, 280, 10), "fp64")
    assert 0 < small < big
    assert request_bytes((720, 280, 10), "mixed") < big
    # 3-D state of 720×280×10: T, S, 4 stage buffers and 3 face velocities ≈ 9 fields
    assert 9 * 720 * 280 * 10 * 8 < big < 12 * 720 * 280 * 10 * 8


def test_poisson_arrivals_rate():
    rng = np.random.default_rng(0)
    t = poisson_arrivals(200.0, 20.0, rng)
    assert np.all(np.diff(t) > 0) and t[-1] < 20.0
    assert abs(len(t) / 20.0 - 200.0) < 10.0


def _requests(n, steps, seed):
    specs = draw_ensemble(n, seed, largest=(120, 50, 10))
    out = []
    for k, s in enumerate(specs):
        out.append(Request(id=k, case=member_case(s), params=member_params(s),
                           precision="fp64" if k % 2 == 0 else "mixed", steps=steps, arrival=0.0))
    return out


def _fresh(r, arrival=0.0):
    c = copy.copy(r)
    c.result = {}
    c.arrival = arrival
    c.finished = c.admitted = float("nan")
    return c


@pytest.mark.gpu
def test_same_result_alone_in_the_mix_and_on_the_plain_path(require_gpu):
    from numericalearth_hip.serving import Server, serve_plain
    reqs = _requests(8, steps=12, seed=3)
    print("[serve] shapes", [(r.case.shape, r.precision) for r in reqs], flush=True)
    for target in (reqs[2], reqs[5]):          # one FP64, one mixed
        srv = Server(budget_bytes=8e9)
        alone = srv.run([_fresh(target)])[0]
        srv.close()
        mix_in = [_fresh(r, arrival=0.002 * k) for k, r in enumerate(reqs) if r is not target]
        mix_in.insert(3, _fresh(target, arrival=0.005))
        srv = Server(budget_bytes=8e9)
        done = srv.run(mix_in)
        srv.close()
        assert len(done) == len(reqs)
        mixed = next(r for r in done if r.id == target.id)
        plain = serve_plain([_fresh(target)])[0]
        for k in alone.result:
            assert np.array_equal(alone.result[k], mixed.result[k], equal_nan=True), (target.precision, k)
            assert np.array_equal(alone.result[k], plain.result[k], equal_nan=True), (target.precision, k)
        print(f"[serve] request {target.id} ({target.precision}) identical alone / in the mix / plain; "
              f"mix {latency_summary(done)}", flush=True)


@pytest.mark.gpu
def test_budget_bounds_residency(require_gpu):
    from numericalearth_hip.serving import Server
    reqs = _requests(6, steps=5, seed=4)
    one = max(request_bytes(r.case.shape, r.precision) for r in reqs)
    srv = Server(budget_bytes=one)              # room for one request at a time
    peak = []
    orig = srv._admit

    def admit():
        orig()
        peak.append(srv.queue.resident)
    srv._admit = admit
    done = srv.run([_fresh(r) for r in reqs])
    srv.close()
    assert len(done) == len(reqs)
    assert max(peak) <= one
    assert srv.queue.resident == 0
