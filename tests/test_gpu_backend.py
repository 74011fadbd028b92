"""BlackwellSimulator on the GPU against the FP64 oracle.

Bars: fidelity >= 1 - 1e-7 (complex128 up-cast before normalising) and the
phase-aligned max |delta amp| reported and bounded; permutation circuits
bit-exact; shot counts pass chi-square (p > 1e-3) against the FP64
distribution; Chebyshev evolution within 1e-5 of scipy's expm_multiply.
"""
import math

import numpy as np
import pytest
import torch

from oracle import hamiltonian as OH
from oracle import statevector as O

pytestmark = pytest.mark.gpu

MODES = ("auto", "plain", "tiles", "blocks")


@pytest.fixture(scope="module")
def sim():
    from blackwell_backend import BlackwellSimulator
    return BlackwellSimulator("cuda:0")


def oracle_sv(qc):
    from blackwell_backend.circuit import instructions
    return O.run_instructions(qc.num_qubits, [(nm, q, p) for nm, q, _, p in instructions(qc)])


def check(ref, sv, f_min=1 - 1e-7, d_max=2e-6):
    f = O.fidelity(ref, sv)
    d = O.max_abs_diff_phase_aligned(ref, sv)
    assert f >= f_min and d <= d_max, (f, d)
    return f, d


@pytest.mark.parametrize("mode", MODES)
def test_config1_rotation_ladder_12q(sim, mode):
    from blackwell_backend import workloads as W
    qc = W.c1_rotation_ladder(12, 12, seed=1234)
    ref = oracle_sv(qc)
    sv, _ = sim.run_statevector(qc, mode=mode)
    f, d = check(ref, sv)
    print(f"C1 {mode}: fidelity {f:.12f} max|d| {d:.3e}")


@pytest.mark.parametrize("mode", ["plain", "blocks", "tiles"])
def test_su4_volume_20q(sim, mode):
    from blackwell_backend import workloads as W
    qc = W.su4_volume(20, 20, seed=2024)
    ref = oracle_sv(qc)
    sv, _ = sim.run_statevector(qc, mode=mode)
    f, d = check(ref, sv, d_max=5e-6)
    print(f"SU4-20 {mode}: fidelity {f:.12f} max|d| {d:.3e}")


@pytest.mark.parametrize("mode", ["plain", "blocks", "tiles"])
def test_rotation_mcx_16q(sim, mode):
    from blackwell_backend import workloads as W
    qc = W.c3_rotation_mcx(16, 3, seed=77)
    ref = oracle_sv(qc)
    sv, _ = sim.run_statevector(qc, mode=mode)
    f, d = check(ref, sv)
    print(f"rot-MCX-16 {mode}: fidelity {f:.12f} max|d| {d:.3e}")


def test_random_circuits_all_modes(sim):
    from test_compiler import random_circuit
    for seed, n in [(1, 6), (2, 11), (3, 14)]:
        qc = random_circuit(n, 200, seed)
        ref = oracle_sv(qc)
        for mode in MODES:
            sv, _ = sim.run_statevector(qc, mode=mode)
            check(ref, sv)


def test_permutation_circuit_bit_exact(sim):
    from blackwell_backend.circuit import QuantumCircuit
    rng = np.random.default_rng(0)
    n = 14
    qc = QuantumCircuit(n)
    qc.x(3).x(0).x(13)
    for _ in range(60):
        qs = [int(q) for q in rng.choice(n, 4, replace=False)]
        r = rng.integers(4)
        if r == 0:
            qc.cx(qs[0], qs[1])
        elif r == 1:
            qc.ccx(qs[0], qs[1], qs[2])
        elif r == 2:
            qc.swap(qs[0], qs[1])
        else:
            qc.cswap(qs[0], qs[1], qs[2])
    ref = oracle_sv(qc).astype(np.complex64)
    for mode in ("plain", "tiles", "auto"):
        sv, _ = sim.run_statevector(qc, mode=mode)
        assert np.array_equal(sv.view(np.uint32), ref.view(np.uint32)), mode


def test_stress_repeated_gates_through_the_abi(bk, dev):
    n = 8
    psi = np.zeros(1 << n, dtype=np.complex64)
    psi[0] = 1
    st = torch.from_numpy(psi.copy()).to(dev)
    h = bk.hadamard_gate(dev)
    for _ in range(1000):
        bk.apply_gate(st, h, 3)
    got = st.cpu().numpy()
    assert abs(got[0] - 1) < 1e-5 and np.sum(np.abs(got[1:])) < 1e-5
    st = torch.from_numpy(psi.copy()).to(dev)
    for _ in range(999):
        bk.apply_mcx(st, [], 5)
    ref = np.zeros_like(psi)
    ref[1 << 5] = 1
    assert np.array_equal(st.cpu().numpy(), ref)
    st = torch.from_numpy(psi.copy()).to(dev)
    bk.apply_mcx(st, [], 1)
    for _ in range(1000):
        bk.apply_mcx(st, [1], 4)
    ref = np.zeros_like(psi)
    ref[2] = 1
    assert np.array_equal(st.cpu().numpy(), ref)
    st = torch.from_numpy((np.ones(1 << n) / 16).astype(np.complex64)).to(dev)
    d = bk.rz_gate_diag(0.001, dev)
    for _ in range(10000):
        bk.apply_diagonal(st, d, 2)
    ref = O.apply_diag(np.ones(1 << n) / 16, [np.exp(-5j), np.exp(5j)], 2)
    assert O.fidelity(ref, st.cpu().numpy()) > 0.99999


def test_run_counts_and_chi_square(sim):
    from blackwell_backend.circuit import QuantumCircuit
    qc = QuantumCircuit(2, 2, name="bell")
    qc.h(0).cx(0, 1).measure_all()
    res = sim.run(qc, shots=50000, seed=7)
    counts = res.get_counts()
    assert set(counts) <= {"0x0", "0x3"}
    assert sum(counts.values()) == 50000
    assert res.results[0].header["n_qubits"] == 2
    again = sim.run(qc, shots=50000, seed=7).get_counts()
    assert again == counts
    from blackwell_backend import workloads as W
    qc = W.su4_volume(10, 10, seed=5)
    ref = oracle_sv(qc)
    res = sim.run(qc, shots=200000, seed=3)
    idx = np.concatenate([[int(k, 16)] * v for k, v in res.get_counts().items()])
    p, _, _ = O.chi_square_p(idx, O.probs(ref), 200000)
    assert p > 1e-3


def test_midcircuit_measure_and_reset(sim):
    from blackwell_backend.circuit import QuantumCircuit
    qc = QuantumCircuit(3, 1)
    qc.h(0).cx(0, 1).measure(0, 0).reset(2).h(2)
    for seed in range(4):
        torch.manual_seed(seed)
        sv, _ = sim.run_statevector(qc)
        p = np.abs(sv) ** 2
        # qubits 0 and 1 collapsed together; qubit 2 in |+>
        assert np.isclose(p.sum(), 1, atol=1e-6)
        on = {i for i in range(8) if p[i] > 1e-6}
        assert on in ({0, 4}, {3, 7})


def _csr_cases():
    from blackwell_backend import workloads as W
    rng = np.random.default_rng(11)
    labels = ["".join(rng.choice(list("IXYZ"), 10)) for _ in range(40)]
    labels += labels[:6]                                   # repeated strings sum
    coeffs = rng.uniform(-1, 1, 46) + 1j * rng.uniform(-1, 1, 46)
    h18 = W.heisenberg_chain(18, seed=2)
    return {
        "random_10q": (labels, coeffs, 10),
        "heisenberg_12q": (W.heisenberg_chain(12, seed=3).labels, W.heisenberg_chain(12, seed=3).coeffs, 12),
        "heisenberg_18q": (h18.labels, h18.coeffs, 18),    # 1024 row blocks: multi-chunk scan
        "cancel_2q": (["XX", "YY", "ZZ", "ZZ"], [1.0, 1.0, 0.5, -0.5], 2),
        "all_zero_3q": (["ZIX"], [0.0], 3),
        "single_1q": (["Y"], [0.3 + 0.2j], 1),
    }


@pytest.mark.parametrize("case", sorted(_csr_cases()))
def test_pauli_csr_device_matches_oracle(case):
    """Device CSR = the oracle's grouped CSR: row_ptr and col_idx equal,
    values equal after rounding the FP64 sums to complex64."""
    from blackwell_backend.hamiltonian import pauli_to_csr_device
    labels, coeffs, n = _csr_cases()[case]
    print(f"[csr] {case} n={n} terms={len(labels)}", flush=True)
    rp, ci, v = pauli_to_csr_device(labels, coeffs, n, "cuda:0")
    rp0, ci0, v0 = OH.pauli_csr_grouped(labels, coeffs, n)
    assert np.array_equal(rp.cpu().numpy(), rp0)
    assert np.array_equal(ci.cpu().numpy(), ci0)
    assert np.array_equal(v.cpu().numpy(), v0.astype(np.complex64))


@pytest.mark.parametrize("case", ["random_10q", "heisenberg_12q", "single_1q"])
def test_hamiltonian_expectation_matches_dense(case):
    """<psi|H|psi> from the grouped FP64 Pauli expectations equals the dense
    FP64 sandwich with the oracle's Kronecker-product matrix."""
    from blackwell_backend.hamiltonian import SparsePauliOp, expectation
    labels, coeffs, n = _csr_cases()[case]
    if case != "single_1q":
        coeffs = np.real(coeffs)                  # Hermitian: real total
    rng = np.random.default_rng(5)
    psi = (rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)).astype(np.complex64)
    st = torch.from_numpy(psi).to("cuda:0")
    got = expectation(st, SparsePauliOp(labels, coeffs))
    p64 = psi.astype(np.complex128)
    want = np.vdot(p64, OH.pauli_dense(labels, coeffs) @ p64)
    if case == "single_1q":
        assert isinstance(got, complex) and abs(got - want) < 1e-12 * max(1, abs(want))
    else:
        assert isinstance(got, float) and abs(got - want.real) < 1e-11 * max(1, abs(want))


def test_hamiltonian_evolution_matches_expm(sim):
    from blackwell_backend import workloads as W
    from blackwell_backend.hamiltonian import HamiltonianCSR
    from blackwell_backend.chebyshev import chebyshev_evolve_batch
    spo = W.heisenberg_chain(10, seed=3)
    h = OH.pauli_sparse(spo.labels, spo.coeffs)
    psi0 = np.zeros(1 << 10, dtype=np.complex128)
    psi0[0] = 1
    for t in (0.5, 2.0):
        sv, info = sim.run_hamiltonian(spo, t, precision=1e-8)
        ref = OH.expm_reference(h, psi0, t)
        assert np.max(np.abs(sv - ref)) < 1e-5, (t, info)
        assert info["engine"] == "chebyshev-beta" and info["chebyshev_degree"] >= 2
    hc = HamiltonianCSR.from_sparse_pauli_op(spo, device="cuda:0")
    rng = np.random.default_rng(1)
    X = rng.standard_normal((4, 1 << 10)) + 1j * rng.standard_normal((4, 1 << 10))
    X /= np.linalg.norm(X, axis=1, keepdims=True)
    out, info = chebyshev_evolve_batch(torch.from_numpy(X.astype(np.complex64)).cuda(), hc, 1.5, 1e-8)
    out = out.cpu().numpy()
    for r in range(4):
        ref = OH.expm_reference(h, X[r].astype(np.complex64).astype(np.complex128), 1.5)
        assert np.max(np.abs(out[r] - ref)) < 1e-5
    # H = Z: exact phases (the reference's analytic case)
    from blackwell_backend.hamiltonian import SparsePauliOp
    sv, _ = sim.run_hamiltonian(SparsePauliOp.from_list([("Z", 1.0)]), 1.3,
                                initial_state=np.array([0.6, 0.8j], dtype=np.complex64))
    assert np.allclose(sv, [0.6 * np.exp(-1.3j), 0.8j * np.exp(1.3j)], atol=1e-6)
    res, info = sim.run_hamiltonian(spo, 0.7, shots=4096)
    assert sum(res.get_counts().values()) == 4096


def test_run_and_time_stats(sim):
    from blackwell_backend import workloads as W
    res, stats = sim.run_and_time(W.c1_rotation_ladder(10, 2), shots=64, warmup=1, repeats=3)
    assert len(stats["times_s"]) == 3 and stats["median_s"] > 0
    assert sum(res.get_counts().values()) == 64


def test_graph_replays_match_eager():
    """run_statevector / run replay HIP graphs of small circuits: every replay,
    with eager work and synchronisations between replays, gives the eager
    result bit for bit; plans a graph cannot hold run eagerly."""
    from blackwell_backend import BlackwellSimulator, workloads as W
    from blackwell_backend.circuit import QuantumCircuit
    eager = BlackwellSimulator("cuda:0", use_graphs=False)
    graphed = BlackwellSimulator("cuda:0", use_graphs=True)
    other = W.su4_volume(12, 6, 5)
    mixed = QuantumCircuit(11)
    mixed.h(0).cx(0, 1).ccx(0, 1, 2).unitary(np.eye(8), [3, 5, 7]).rz(0.4, 9)
    cases = [(W.su4_volume(14, 14, 3), "auto"), (W.qft(12), "auto"), (W.qft(12), "plain"),
             (W.c3_rotation_mcx(16, 2, 4), "blocks"), (W.c3_rotation_mcx(12, 2, 9), "tiles"),
             (mixed, "plain")]
    for qc, mode in cases:
        print(f"[graph] {qc.name} mode={mode}", flush=True)
        ref, _ = eager.run_statevector(qc, mode=mode)
        for rep in range(3):
            got, _ = graphed.run_statevector(qc, mode=mode)
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (qc.name, mode, rep)
            eager.run_statevector(other)                 # eager work between replays
            torch.cuda.synchronize()
    cc = graphed.compile(mixed)
    assert cc.graphed("plain", graphed.device) is None    # 3-qubit dense op: host-uploaded table
    assert graphed.compile(W.qft(12)).graphed("plain", graphed.device) is not None


def test_global_phase_multiplies_the_state():
    """global_phase = phi gives e^{i phi} times the phase-free state in every
    mode, eager and graphed; changing it on the same circuit object is seen."""
    from blackwell_backend import BlackwellSimulator, workloads as W
    from blackwell_backend.circuit import QuantumCircuit
    phi = 0.7
    ph = np.exp(1j * phi)
    qc = QuantumCircuit(1)
    qc.h(0)
    qc.global_phase = phi
    for graphs in (False, True):
        sim = BlackwellSimulator("cuda:0", use_graphs=graphs)
        sv, _ = sim.run_statevector(qc)
        assert np.allclose(sv, ph * np.array([1, 1]) / np.sqrt(2), atol=1e-7), (graphs, sv)
        base = W.su4_volume(12, 6, 8)
        ref, _ = sim.run_statevector(base, mode="plain")
        for mode in ("plain", "blocks", "tiles", "whole"):
            c = base.copy()
            c.global_phase = phi
            got, _ = sim.run_statevector(c, mode=mode)
            assert np.max(np.abs(got - ph * ref)) < 5e-6, (graphs, mode)
        qc2 = QuantumCircuit(2)
        qc2.x(1)
        a, _ = sim.run_statevector(qc2)
        qc2.global_phase = np.pi
        b, _ = sim.run_statevector(qc2)
        assert a[2] == 1 and abs(b[2] + 1) < 1e-7


def test_tune_mode_picks_a_measured_mode(sim):
    from blackwell_backend import BlackwellSimulator, workloads as W
    tuned = BlackwellSimulator("cuda:0", mode="tune")
    qc = W.deep_rz(12, depth=20)
    sv, _ = tuned.run_statevector(qc)
    ref = oracle_sv(qc)
    check(ref, sv)
    mode = tuned.compile(qc).tuned_mode(tuned.device)
    assert mode in ("plain", "whole", "blocks", "blocks/4")
    res = tuned.run(qc, shots=1000, seed=3)
    assert sum(res.get_counts().values()) == 1000
    # above 13 qubits the tuner also times tile shapes; every shape is exact
    qc = W.qft(15)
    ref = oracle_sv(qc)
    mode = tuned.compile(qc).tuned_mode(tuned.device)
    assert mode == "plain" or mode.startswith("blocks") or mode.startswith("tiles/")
    for m, low in ((13, 3), (12, 4), (11, 5)):
        sv, _ = tuned.run_statevector(qc, mode=f"tiles/{m}/{low}")
        check(ref, sv)
    with pytest.raises(ValueError):
        tuned.run_statevector(qc, mode="tiles/14/4")


def test_two_kernel_builds_side_by_side(tmp_path):
    """The reference's kernel A/B harness (tests/kernel_ab.py:100-230) runs
    two simulators on two kernel builds: here a copy of this build at another
    path; each simulator's plans launch through its own library, and the two
    agree bit for bit."""
    import os
    import shutil
    import sys
    from blackwell_backend import BlackwellSimulator, workloads as W
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "qiskit-blackwell_amd")
    other = tmp_path / "build_b"
    shutil.copytree(os.path.join(pkg, "blackwell_kernels"), other / "blackwell_kernels",
                    ignore=shutil.ignore_patterns("__pycache__"))
    try:
        sim_a = BlackwellSimulator("cuda:0", kernel_path=pkg)
        sim_b = BlackwellSimulator("cuda:0", kernel_path=str(other))
        assert sim_b.bk._lib.LIB_PATH.startswith(str(other)) and sim_a.bk is not sim_b.bk
        for qc in (W.qft(10), W.su4_volume(14, 8, 3), W.c3_rotation_mcx(12, 2, 5)):
            assert sim_b.compile(qc).bk is sim_b.bk
            a, _ = sim_a.run_statevector(qc)
            b, _ = sim_b.run_statevector(qc)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), qc.name
    finally:
        if str(other) in sys.path:
            sys.path.remove(str(other))
        from blackwell_backend.simulator import _load_kernels
        _load_kernels(pkg)


def test_block_widths_and_width_tuner(sim):
    """Circuits run as fused blocks of 4, 5 and 6 qubits match the FP64
    oracle; the width tuner times all three and returns the fastest, and
    every width it timed is a valid mode."""
    from test_compiler import random_circuit
    from blackwell_backend import workloads as W
    for seed, n in [(4, 13), (5, 16)]:
        qc = random_circuit(n, 160, seed)
        ref = oracle_sv(qc)
        for mode in ("blocks/4", "blocks", "blocks/6"):
            sv, _ = sim.run_statevector(qc, mode=mode)
            check(ref, sv)
    qc = W.su4_volume(16, 16, 7)
    ref = oracle_sv(qc)
    cc = sim.compile(qc)
    best, times = cc.tuned_block_width(sim.device)
    assert set(times) == {4, 5, 6} and best in ("blocks/4", "blocks", "blocks/6")
    from blackwell_backend.executor import block_width
    assert times[block_width(best)] == min(times.values())
    sv, _ = sim.run_statevector(qc, mode=best)
    check(ref, sv)
    with pytest.raises(ValueError):
        sim.run_statevector(W.qft(12), mode="blocks/6")       # k = 6 needs n >= 13


def test_chebyshev_split_term_16q_matches_fp64():
    """At 16 qubits and 32 RHS, split=True runs each Chebyshev term in two
    passes (HamiltonianCSR.split); the result matches the FP64 Chebyshev oracle
    (fidelity >= 1 - 1e-7, max |d amp| <= 1e-5 max |a|) and the one-pass term
    within 1e-6."""
    from blackwell_backend import workloads as W
    from blackwell_backend.chebyshev import chebyshev_evolve_batch
    from blackwell_backend.hamiltonian import HamiltonianCSR
    from oracle import hamiltonian as OH
    n, R = 16, 32
    op = W.heisenberg_chain(n, 11)
    h = HamiltonianCSR.from_sparse_pauli_op(op, device="cuda:0")
    rng = np.random.default_rng(12)
    X = (rng.standard_normal((R, 1 << n)) + 1j * rng.standard_normal((R, 1 << n))).astype(np.complex64)
    X /= np.linalg.norm(X, axis=1, keepdims=True).astype(np.float32)
    Xd = torch.from_numpy(X).cuda()
    Y, info = chebyshev_evolve_batch(Xd, h, 2.0, 1e-8, split=True)
    assert info["split_term"]
    Y1, info1 = chebyshev_evolve_batch(Xd, h, 2.0, 1e-8)
    assert not info1["split_term"]
    with pytest.raises(ValueError):
        chebyshev_evolve_batch(Xd[:4], h, 2.0, 1e-8, split=True)
    got, one = Y.cpu().numpy(), Y1.cpu().numpy()
    assert np.max(np.abs(got - one)) <= 1e-6
    ref, _ = OH.chebyshev_evolve_c(op.labels, op.coeffs, n, X[:4].astype(np.complex128), 2.0,
                                   h.spectral_bound, precision=1e-13)
    for k in range(4):
        f = O.fidelity(ref[k], got[k])
        d = float(np.max(np.abs(ref[k] - got[k])))
        assert f >= 1 - 1e-7 and d <= 1e-5 * np.max(np.abs(ref[k])), (k, f, d)
