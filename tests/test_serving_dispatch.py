"""The serving loop's dispatch on the CPU: a Server whose device work is
replaced by a timed stand-in (no device contexts) shows the order in which
prepared requests enter the slots, the bound on prepared requests and the
report of a request whose host work fails."""
import dataclasses
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import pytest

from kore_amd import backend, serving

KEY = (1, 1)


class StandInServer(serving.Server):
    """Server with the dispatch state only; a request's slot time is a sleep."""

    def __init__(self, slots, eigen_first, service=0.05, lookahead=8, fail=()):
        self.slots = slots
        self.mem_budget = 1 << 50
        self.precision = backend.PREC_C64
        self.block = 1
        self.free_slots = queue.Queue()
        for i in range(slots):
            self.free_slots.put(i)
        self.mem_in_use = 0
        self.mem_peak = 0
        self.max_inflight = 0
        self._inflight = 0
        self._alone = False
        self._cv = threading.Condition()
        self.asm_pool = ThreadPoolExecutor(max_workers=2)
        self.lookahead = lookahead
        self.log = None
        self.eigen_first = eigen_first
        self.service = service
        self.fail = set(fail)
        self.starts = []
        self._count = threading.Lock()
        self.prepared = 0
        self.max_waiting = 0

    def _prepare(self, req):
        if req.rid in self.fail:
            raise ValueError("assembly failed for request %d" % req.rid)
        with self._count:
            self.prepared += 1
            self.max_waiting = max(self.max_waiting, self.prepared - len(self.starts))
        return req.params(), None, None, None, (0.0, 0.0)

    def _execute(self, slot, req, par, A0, B0, t0, plan=None):
        with self._count:
            self.starts.append(req.rid)
        time.sleep(self.service)
        return serving.Result(req.rid, req.kind, time.perf_counter() - (t0 + req.arrival), self.service, 0.0, True)

    def close(self):
        self.asm_pool.shutdown(wait=True)


@pytest.fixture
def stand_in(monkeypatch):
    # one tuned class and a nominal footprint: no library calls
    monkeypatch.setattr(serving, "tuner_key", lambda par, prec: KEY)
    monkeypatch.setattr(serving, "footprint", lambda par, prec, nev=6, block=1: 1)
    monkeypatch.setattr(serving, "_TUNED", {KEY})


def _trace(kinds, arrivals):
    base = serving.serving_trace(nreq=len(kinds), seed=3, e_range=(1e-4, 1e-3), n_cap=24)
    return [dataclasses.replace(r, rid=i, kind=k, arrival=a) for i, (r, k, a) in enumerate(zip(base, kinds, arrivals))]


@pytest.mark.parametrize("eigen_first", [False, True])
def test_order_when_a_slot_frees(stand_in, eigen_first):
    """One slot busy with request 0; requests 1-5 arrive meanwhile.  With
    eigen_first the two eigen requests (4, 5) enter the slot next; otherwise
    the forced ones that completed their host work earlier do."""
    kinds = ["forced", "forced", "forced", "forced", "eigen", "eigen"]
    arrivals = [0.0, 0.005, 0.006, 0.007, 0.008, 0.009]
    srv = StandInServer(slots=1, eigen_first=eigen_first, service=0.08)
    out = srv._run(_trace(kinds, arrivals))
    srv.close()
    assert all(r.ok for r in out)
    assert srv.starts[0] == 0
    nxt = set(srv.starts[1:3])
    assert nxt == ({4, 5} if eigen_first else {1, 2})
    assert sorted(srv.starts) == list(range(6))


def test_prepared_requests_stay_within_the_window(stand_in):
    """All requests arrive at once on two slots: at most `lookahead` are
    prepared and not yet dispatched at any time (counted here from the start of
    their slot work, a dispatched request's worker may not have started yet:
    at most one per slot more), and every request is served."""
    n = 20
    srv = StandInServer(slots=2, eigen_first=True, service=0.01, lookahead=3)
    out = srv._run(_trace(["forced", "eigen"] * (n // 2), [0.0] * n))
    srv.close()
    assert len(out) == n and all(r.ok for r in out)
    assert srv.max_waiting <= 3 + 2
    assert srv.max_inflight <= 2


def test_failed_host_work_is_reported(stand_in):
    srv = StandInServer(slots=2, eigen_first=True, fail={2})
    out = srv._run(_trace(["forced", "eigen", "forced", "eigen"], [0.0, 0.001, 0.002, 0.003]))
    srv.close()
    assert not out[2].ok and "assembly failed" in out[2].error
    assert all(out[i].ok for i in (0, 1, 3))
    assert sorted(srv.starts) == [0, 1, 3]


def test_slot_count_is_bounded():
    """More slots than the process's hardware queues are refused before any
    device context is made."""
    with pytest.raises(ValueError):
        serving.Server(slots=serving.MAX_SLOTS + 1)
    with pytest.raises(ValueError):
        serving.Server(slots=0)
