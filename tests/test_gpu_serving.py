"""Continuous-batching server on the GPU: every request is answered, its
samples are bitwise the same when it runs alone and inside a mixed batch,
and the sampled distribution matches the FP64 statevector (chi-square)."""
import numpy as np
import pytest
import torch

from oracle import statevector as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim():
    from blackwell_backend import BlackwellSimulator
    return BlackwellSimulator("cuda:0")


def _reqs(ns, shots=2048, base_seed=17):
    from blackwell_backend import serving as S
    return [S.su4_request(i, 0.0, n, shots, base_seed * 1000 + i) for i, n in enumerate(ns)]


def test_alone_equals_mixed_and_matches_fp64(sim):
    from blackwell_backend import serving as S, workloads as W
    from blackwell_backend.circuit import instructions
    ns = [10, 12, 14, 11, 16, 13, 10, 18, 12, 15]
    mixed = _reqs(ns)
    st = S.Server(sim).serve(mixed, realtime=False)
    assert st["served"] == len(ns)
    for i, r in enumerate(mixed):
        alone = _reqs([n for n in ns])[i]
        S.Server(sim).serve([alone], realtime=False)
        assert np.array_equal(alone.samples, r.samples), (i, r.n_qubits)
        assert r.lane == ("lds" if r.n_qubits <= 14 else "hbm")
    # distributions against FP64 for a few requests
    for r in mixed[:3] + [mixed[4]]:
        qc = W.su4_volume(r.n_qubits, r.n_qubits, r.program[3])
        psi = O.run_instructions(r.n_qubits, [(nm, q, p) for nm, q, _, p in instructions(qc)])
        p, _, _ = O.chi_square_p(r.samples, O.probs(psi), r.shots)
        assert p > 1e-3, (r.n_qubits, p)
        # the samples are the inverse CDF of the seeded uniforms on the FP64 state,
        # except where the FP32 state moves a CDF boundary past a uniform
        u = O.splitmix_uniforms(r.seed, r.shots)
        assert np.mean(O.sample(psi, u) == r.samples) > 0.99


def test_plain_path_matches_distribution(sim):
    from blackwell_backend import serving as S
    reqs = _reqs([10, 13, 15])
    S.serve_plain(reqs, sim, realtime=False)
    served = _reqs([10, 13, 15])
    S.Server(sim).serve(served, realtime=False)
    for a, b in zip(reqs, served):
        # same uniforms, different (but both FP32-compensated) arithmetic
        assert np.mean(a.samples == b.samples) > 0.99


def test_budget_and_lanes_on_device(sim):
    from blackwell_backend import serving as S
    reqs = _reqs([20, 10, 21, 11])
    srv = S.Server(sim, budget_bytes=(8 << 21) + (4 << 20))
    st = srv.serve(reqs, realtime=False)
    assert st["rejected"] == 0 and st["served"] == 4
    assert srv.peak_in_use <= srv.budget


def test_gate_payload_equals_circuit_path(sim):
    """An HBM-lane request planned from its gate payload (native planner on
    a host thread, one qsim_apply_blocks call) and the same circuit compiled
    from its circuit object give bitwise the same samples."""
    from blackwell_backend import serving as S
    for n in (15, 17):
        a = S.su4_request(0, 0.0, n, 2048, 4242 + n)
        b = S.Request(1, 0.0, n, 2048, 4242 + n, program=("su4", n, n, 4242 + n))
        assert b.gates is None
        S.Server(sim).serve([a], realtime=False)
        S.Server(sim, planner_threads=0).serve([b], realtime=False)
        assert a.lane == b.lane == "hbm"
        assert np.array_equal(a.samples, b.samples), n


def test_served_states_match_fp64_alone_and_mixed(sim):
    """keep_state requests in both lanes (LDS n <= 14, HBM batched and
    unbatched) return their final statevectors: fidelity >= 1 - 1e-7 and
    max |d amp| <= 1e-5 against the FP64 statevector of the same gates, and
    the state of a request served alone equals, bit for bit, its state
    served inside the mix."""
    from blackwell_backend import serving as S, workloads as W
    ns = [10, 14, 12, 16, 19, 11, 20, 15, 17]

    def mk():
        reqs = [S.su4_request(i, 0.0, n, 256, 31 * 1000 + i, keep_state=True)
                for i, n in enumerate(ns[:-1])]
        # no gate payload: the HBM lane runs it alone, compiled from its circuit
        i, n = len(ns) - 1, ns[-1]
        reqs.append(S.Request(i, 0.0, n, 256, 31 * 1000 + i, program=("su4", n, n, 31 * 1000 + i),
                              keep_state=True))
        return reqs
    mixed = mk()
    S.Server(sim).serve(mixed, realtime=False)
    for r in mixed:
        assert r.state is not None and r.state.shape == (1 << r.n_qubits,)
        pairs, us = W.su4_arrays(r.n_qubits, r.n_qubits, r.seed)
        psi = O.zero_state(r.n_qubits)
        for (a, b), u in zip(pairs, us):
            psi = O.apply_matrix(psi, np.asarray(u, dtype=np.complex128), [int(a), int(b)])
        f = O.fidelity(psi, r.state)
        d = float(np.max(np.abs(psi - r.state)))
        assert f >= 1 - 1e-7 and d <= 1e-5, (r.n_qubits, r.lane, f, d)
    for i in (0, 3, 6, 8):
        alone = mk()[i]
        S.Server(sim).serve([alone], realtime=False)
        assert np.array_equal(alone.state.view(np.uint32), mixed[i].state.view(np.uint32)), ns[i]


def test_repeated_circuits_replay_as_graphs(sim):
    """HBM-lane requests that repeat a circuit (no gate payload) replay a
    captured HIP graph from the second occurrence on; their samples and
    states equal, bit for bit, those of the same requests served with graph
    replay turned off, and every result matches the FP64 statevector."""
    from blackwell_backend import serving as S, workloads as W

    def mk():
        out = []
        for i in range(6):
            n, cseed = (16, 808) if i % 2 == 0 else (18, 909)
            out.append(S.Request(i, 0.0, n, 512, 5000 + i, program=("su4", n, n, cseed), keep_state=True))
        return out
    with_graphs, without = mk(), mk()
    srv = S.Server(sim)
    srv.serve(with_graphs, realtime=False)
    S.Server(sim, graph_repeats=0).serve(without, realtime=False)
    assert srv.graphs.captures == 2 and srv.graphs.replays == 4, (srv.graphs.captures, srv.graphs.replays)
    for a, b in zip(with_graphs, without):
        assert a.lane == b.lane == "hbm"
        assert np.array_equal(a.samples, b.samples), a.rid
        assert np.array_equal(a.state.view(np.uint32), b.state.view(np.uint32)), a.rid
    for r in with_graphs[:2]:
        qc = W.su4_volume(r.n_qubits, r.n_qubits, r.program[3])
        psi = O.run_instructions(r.n_qubits, [(nm, q, p) for nm, q, _, p in instructions_of(qc)])
        assert O.fidelity(psi, r.state) >= 1 - 1e-7


def instructions_of(qc):
    from blackwell_backend.circuit import instructions
    return instructions(qc)
