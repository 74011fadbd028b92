"""Serving scheduler on the CPU with a fake device executor: admission by
statevector bytes against an injected budget, rejection of requests that
can never fit, lane routing, overlap of the two lanes, continuous admission
of late arrivals, and the timeout."""
import time

import numpy as np
import pytest


class FakeJob:
    """Completes `dur` seconds after launch; samples = rid repeated."""
    log = []

    def __init__(self, reqs, dur):
        self.reqs = reqs if isinstance(reqs, list) else [reqs]
        self.dur = dur
        self.t = None

    def launch(self):
        self.t = time.perf_counter()
        FakeJob.log.append(("launch", [r.rid for r in self.reqs], self.t))

    def done(self):
        return time.perf_counter() - self.t >= self.dur

    def finish(self, now):
        for r in self.reqs:
            r.samples = np.full(r.shots, r.rid)
            r.t_done = now


def make_server(budget, dur_lds=0.002, dur_hbm=0.01, **kw):
    from blackwell_backend.serving import Server
    FakeJob.log = []
    return Server(budget_bytes=budget, lds_factory=lambda rs: FakeJob(rs, dur_lds),
                  hbm_factory=lambda r: FakeJob(r, dur_hbm), **kw)


def reqs(ns, gap=0.0, shots=16):
    from blackwell_backend.serving import Request
    return [Request(i, i * gap, n, shots, 1000 + i, program=("su4", n, n, i)) for i, n in enumerate(ns)]


def test_lane_routing_and_all_served():
    s = make_server(1 << 40)
    rs = reqs([10, 14, 15, 24, 12, 20])
    st = s.serve(rs)
    assert st["served"] == 6 and st["rejected"] == 0
    assert [r.lane for r in rs] == ["lds", "lds", "hbm", "hbm", "lds", "hbm"]
    assert all(np.all(r.samples == r.rid) for r in rs)
    assert st["p99_latency_s"] >= st["p50_latency_s"] > 0


def test_budget_rejects_what_never_fits_and_bounds_residency():
    # budget fits one 20-qubit state (8 MiB + shots + 1 MiB slack) but not a 24-qubit one
    budget = (8 << 20) + (2 << 20)
    s = make_server(budget, dur_hbm=0.005)
    rs = reqs([20, 24, 20, 20, 11])
    st = s.serve(rs)
    assert st["rejected"] == 1 and rs[1].t_done is None
    assert st["served"] == 4
    assert s.peak_in_use <= budget
    # the three 20-qubit jobs could not overlap under this budget
    hbm = sorted((r.t_admit, r.t_done) for r in rs if r.lane == "hbm")
    for (a0, d0), (a1, d1) in zip(hbm, hbm[1:]):
        assert a1 >= d0 - 1e-3


def test_late_arrivals_join_while_others_run():
    s = make_server(1 << 40, dur_lds=0.02)
    rs = reqs([10] * 6, gap=0.005)
    s.serve(rs)
    launches = [e for e in FakeJob.log if e[0] == "launch"]
    assert len(launches) >= 2                       # not one batch after the last arrival
    assert launches[0][1][0] == 0
    assert all(r.t_admit >= r.arrival - 1e-4 for r in rs)


def test_timeout():
    s = make_server(1 << 40, dur_lds=10.0)
    with pytest.raises(TimeoutError):
        s.serve(reqs([10]), timeout_s=0.05)


def test_splitmix_matches_oracle():
    from blackwell_backend.serving import splitmix_uniforms
    from oracle import statevector as O
    assert np.array_equal(splitmix_uniforms(12345, 1000), O.splitmix_uniforms(12345, 1000))


def test_su4_program_matches_circuit_program():
    from blackwell_backend import serving as S, workloads as W
    ops, cf = S.su4_program(9, 9, 4)
    ops2, cf2 = S.circuit_program(W.su4_volume(9, 9, 4))
    assert np.array_equal(ops, ops2) and np.allclose(cf, cf2)


def test_hbm_lane_batches_payload_requests_within_budget():
    """HBM-lane requests that carry gate payloads go out in batches (at most
    max_hbm_batch, within the budget); a request without a payload runs
    alone; every request is answered once."""
    from blackwell_backend.serving import Request, Server

    class FakeHbm(FakeJob):
        def __init__(self, req):
            super().__init__(req, 0.003)
            self.req, self.prepared = req, None

    batches = []

    def batch_factory(jobs):
        batches.append([j.req.rid for j in jobs])
        return FakeJob([j.req for j in jobs], 0.003)

    def one(req):
        return FakeHbm(req)

    budget = 3 * ((1 << 19) * 8 + 8 * 16 + (1 << 20))      # three 19-qubit requests at a time
    s = Server(budget_bytes=budget, lds_factory=lambda rs: FakeJob(rs, 0.001), hbm_factory=one,
               hbm_batch_factory=batch_factory, max_hbm_batch=8, planner_threads=0)
    rs = [Request(i, 0.0, 19, 16, 7 + i, gates=([(0, 1)], [None])) for i in range(10)]
    rs.append(Request(10, 0.0, 19, 16, 99, program=("su4", 19, 19, 1)))   # no payload
    st = s.serve(rs)
    assert st["served"] == 11
    assert all(1 <= len(b) <= 3 for b in batches)          # the budget caps each batch
    assert sorted(sum(batches, [])) == list(range(10))
    assert s.peak_in_use <= budget
    assert rs[10].samples is not None and 10 not in sum(batches, [])


def test_hbm_lane_batches_capped_by_state_bytes():
    """A batch holds at most max_hbm_batch_bytes of states; a circuit larger
    than the cap still runs, alone."""
    from blackwell_backend.serving import Request, Server

    class FakeHbm(FakeJob):
        def __init__(self, req):
            super().__init__(req, 0.002)
            self.req, self.prepared = req, None

    batches = []

    def batch_factory(jobs):
        batches.append([j.req.n_qubits for j in jobs])
        return FakeJob([j.req for j in jobs], 0.002)

    cap = 4 * (1 << 18) * 8                                  # four 18-qubit states
    s = Server(budget_bytes=1 << 40, lds_factory=lambda rs: FakeJob(rs, 0.001), hbm_factory=FakeHbm,
               hbm_batch_factory=batch_factory, max_hbm_batch=64, planner_threads=0,
               max_hbm_batch_bytes=cap)
    ns = [18] * 9 + [20] + [17] * 4
    rs = [Request(i, 0.0, n, 16, 5 + i, gates=([(0, 1)], [None])) for i, n in enumerate(ns)]
    st = s.serve(rs)
    assert st["served"] == len(rs)
    assert sorted(sum(batches, [])) == sorted(ns)
    for b in batches:
        assert len(b) == 1 or sum(8 << n for n in b) <= cap
    assert [20] in batches                                   # larger than the cap: alone


def test_graph_cache_policy():
    """_GraphCache: a circuit's first run is not captured, its second is; at
    most max_graphs graphs stay cached (least recently used dropped, and the
    circuit's own graph slot cleared with it); circuits that are too large or
    measure mid-circuit are never captured; run counts go away with their
    circuits."""
    import gc
    from blackwell_backend.serving import _GraphCache

    class FakeGraph:
        def __init__(self, cc):
            self.cc = cc

    class FakeCC:
        def __init__(self, n=10, stochastic=False):
            self.n_qubits = n
            self.has_stochastic = stochastic
            self._graphs = {}

        def resolve_mode(self, mode):
            return mode

        def graphed(self, mode, device):
            return self._graphs.setdefault((mode, device), FakeGraph(self))

    cache = _GraphCache(max_graphs=2, max_qubits=20)
    a, b, c = FakeCC(), FakeCC(), FakeCC()
    assert cache.get(a, "cpu") is None                  # first run: eager
    ga = cache.get(a, "cpu")
    assert ga is not None and ga.cc is a and cache.get(a, "cpu") is ga
    assert cache.get(b, "cpu") is None and cache.get(b, "cpu").cc is b
    cache.get(a, "cpu")                                  # a most recently used
    assert cache.get(c, "cpu") is None and cache.get(c, "cpu").cc is c
    assert set(id(x) for x in (a, c)) == set(cache.graphs) and b._graphs == {}
    assert cache.captures == 3
    big, noisy = FakeCC(n=21), FakeCC(stochastic=True)
    for cc in (big, noisy):
        assert cache.get(cc, "cpu") is None and cache.get(cc, "cpu") is None
    d = FakeCC()
    assert cache.get(d, "cpu") is None
    n_seen = len(cache.seen)
    del d
    gc.collect()
    assert len(cache.seen) == n_seen - 1


def test_batch_programs_stacked_equals_per_request_programs():
    """The LDS lane's stacked assembly of gate-payload requests gives byte for
    byte the ops and coefficients of concatenating each request's
    gates_program() with its coefficient offset, and programs passed as
    (ops, coefs) tuples or su4 tags go through the per-request path."""
    from blackwell_backend import serving as S
    rs = [S.su4_request(i, 0.0, n, 64, 1000 + i) for i, n in enumerate([10, 12, 14, 11, 13, 10])]
    n_ops, n_cf, ops, cf = S._batch_programs(rs)
    want_ops, want_cf, off = [], [], 0
    for r in rs:
        o, c = S.gates_program(*r.gates)
        o = o.copy()
        o["coef"] += off
        off += len(c)
        want_ops.append(o)
        want_cf.append(np.asarray(c, dtype=np.complex64))
    assert ops.tobytes() == np.concatenate(want_ops).tobytes()
    assert cf.tobytes() == np.concatenate(want_cf).tobytes()
    assert list(n_ops) == [len(r.gates[0]) for r in rs] and list(n_cf) == [16 * len(r.gates[0]) for r in rs]
    as_prog = [S.Request(r.rid, 0.0, r.n_qubits, r.shots, r.seed, program=S.gates_program(*r.gates))
               for r in rs]
    as_tag = [S.Request(r.rid, 0.0, r.n_qubits, r.shots, r.seed, program=("su4", r.n_qubits, r.n_qubits, r.seed))
              for r in rs]
    for alt in (as_prog, as_tag):
        a_ops, a_cf, o2, c2 = S._batch_programs(alt)
        assert o2.tobytes() == ops.tobytes() and c2.tobytes() == cf.tobytes()
