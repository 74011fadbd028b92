"""The reference's stress suite (tests/stress_test.py:134-836 of the
reference), run through BlackwellSimulator on the GPU and checked against the
FP64 oracle instead of Aer.

The twelve groups keep the reference's circuits and sizes: identity round
trips, all-qubit sweep, CX/CZ on every pair, norm preservation, gate
identities, phase accumulation, entanglement structures, measurement
statistics, multi-controlled edge cases, adversarial patterns, large-qubit
scaling and repeated-gate stress.  The bars are the reference's or stricter:
fidelity >= 1 - 1e-7 where it asks 0.99999 (0.999 / 0.99 for its long
FP32 accumulation chains), permutation-only circuits bit-exact, chi-square
p > 1e-3, and no spurious Bell outcomes.
"""
import math

import numpy as np
import pytest

from oracle import statevector as O

pytestmark = pytest.mark.gpu

MODES = ("auto", "plain")
F_MIN = 1 - 1e-7


@pytest.fixture(scope="module")
def sim():
    from blackwell_backend import BlackwellSimulator
    return BlackwellSimulator("cuda:0")


def QC(n):
    from blackwell_backend.circuit import QuantumCircuit
    return QuantumCircuit(n)


def ref_sv(qc):
    from blackwell_backend.circuit import instructions
    return O.run_instructions(qc.num_qubits, [(nm, q, p) for nm, q, _, p in instructions(qc)])


def check(sim, qc, mode, f_min=F_MIN, ref=None):
    sv, _ = sim.run_statevector(qc, mode=mode)
    ref = ref_sv(qc) if ref is None else ref
    f = O.fidelity(ref, sv)
    assert f >= f_min, (qc.name, mode, f)
    return sv, ref


@pytest.mark.parametrize("mode", MODES)
def test_identity_roundtrips(sim, mode):
    n = 6
    pairs = {"h": ("h", "h"), "x": ("x", "x"), "y": ("y", "y"), "z": ("z", "z"), "s": ("s", "sdg"),
             "t": ("t", "tdg"), "sx": ("sx", "sxdg")}
    for q in (0, n // 2, n - 1):
        for a, b in pairs.values():
            qc = QC(n)
            getattr(qc, a)(q)
            getattr(qc, b)(q)
            sv, _ = check(sim, qc, mode)
            assert abs(abs(sv[0]) - 1) < 1e-6
        for g, ang in (("rx", 1.234), ("ry", 2.345), ("rz", 3.456)):
            qc = QC(n)
            getattr(qc, g)(ang, q)
            getattr(qc, g)(-ang, q)
            check(sim, qc, mode)
    for q0, q1 in ((0, 1), (0, n - 1), (n - 2, n - 1)):
        for g in ("cx", "swap"):
            qc = QC(n)
            getattr(qc, g)(q0, q1)
            getattr(qc, g)(q0, q1)
            sv, _ = sim.run_statevector(qc, mode=mode)
            assert sv[0] == 1 and np.count_nonzero(sv) == 1          # permutations: exact
    qc = QC(n)
    qc.ccx(0, 1, 2).ccx(0, 1, 2)
    sv, _ = sim.run_statevector(qc, mode=mode)
    assert sv[0] == 1 and np.count_nonzero(sv) == 1


@pytest.mark.parametrize("mode", MODES)
def test_all_qubit_sweep(sim, mode):
    n = 10
    for q in range(n):
        for g in ("h", "x", "s", "t"):
            qc = QC(n)
            getattr(qc, g)(q)
            check(sim, qc, mode)
        qc = QC(n)
        qc.rx(math.pi / 3, q)
        check(sim, qc, mode)
        qc = QC(n)
        qc.rz(math.pi / 5, q)
        check(sim, qc, mode)


@pytest.mark.parametrize("mode", MODES)
def test_qubit_pair_combinatorics(sim, mode):
    n = 8
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            qc = QC(n)
            qc.h(i).cx(i, j)
            check(sim, qc, mode)
            if i < j:
                qc = QC(n)
                qc.h(i).h(j).cz(i, j)
                check(sim, qc, mode)


@pytest.mark.parametrize("mode", MODES)
def test_norm_preservation(sim, mode):
    n, depth = 12, 200
    rng = np.random.default_rng(42)
    qc = QC(n)
    for _ in range(depth):
        g = rng.choice(["h", "rx", "rz", "cx", "cz", "s", "t"])
        if g in ("h", "s", "t"):
            getattr(qc, g)(int(rng.integers(n)))
        elif g in ("rx", "rz"):
            getattr(qc, g)(rng.uniform(0, 2 * np.pi), int(rng.integers(n)))
        else:
            q0, q1 = rng.choice(n, 2, replace=False)
            getattr(qc, g)(int(q0), int(q1))
    sv, _ = check(sim, qc, mode)
    assert abs(np.linalg.norm(sv.astype(np.complex128)) - 1) < 1e-6


@pytest.mark.parametrize("mode", MODES)
def test_gate_equivalences(sim, mode):
    n = 4

    def sv(qc):
        return sim.run_statevector(qc, mode=mode)[0]

    a, b = QC(n), QC(n)
    a.h(0).cx(0, 1)
    b.h(0).h(1).cz(0, 1).h(1)
    assert O.fidelity(sv(a), sv(b)) >= F_MIN                       # CX = H CZ H
    a, b = QC(n), QC(n)
    a.h(0).t(1).swap(0, 1)
    b.h(0).t(1).cx(0, 1).cx(1, 0).cx(0, 1)
    assert O.fidelity(sv(a), sv(b)) >= F_MIN                       # SWAP = 3 CX
    a, b = QC(n), QC(n)
    a.h(0).z(0).h(0)
    b.x(0)
    assert O.fidelity(sv(a), sv(b)) >= F_MIN                       # HZH = X
    a, b = QC(n), QC(n)
    a.h(0).s(0).s(0)
    b.h(0).z(0)
    assert O.fidelity(sv(a), sv(b)) >= F_MIN                       # S S = Z
    for q in (0, n - 1):
        a, b = QC(n), QC(n)
        a.h(q)
        for _ in range(8):
            a.t(q)
        b.h(q)
        assert O.fidelity(sv(a), sv(b)) >= F_MIN                   # T^8 = I


@pytest.mark.parametrize("mode", MODES)
def test_phase_accumulation(sim, mode):
    n = 8
    for q in (0, n // 2, n - 1):
        many, one = QC(n), QC(n)
        many.h(q)
        for _ in range(1000):
            many.rz(math.pi / 3 / 1000, q)
        one.h(q).rz(math.pi / 3, q)
        f = O.fidelity(sim.run_statevector(many, mode=mode)[0], sim.run_statevector(one, mode=mode)[0])
        assert f > 1 - 1e-6, (q, f)                                 # the reference asks 0.999
    for q in (0, n - 1):
        qc = QC(n)
        for _ in range(500):
            qc.rx(2 * math.pi / 500, q)                             # Rx(2 pi) = -I
        sv = sim.run_statevector(qc, mode=mode)[0]
        assert abs(abs(sv[0]) - 1) < 1e-5 and np.max(np.abs(sv[1:])) < 1e-5


@pytest.mark.parametrize("mode", MODES)
def test_entanglement_structures(sim, mode):
    n = 8
    ghz = QC(n)
    ghz.h(0)
    for i in range(n - 1):
        ghz.cx(i, i + 1)
    sv, _ = check(sim, ghz, mode)
    p = np.abs(sv.astype(np.complex128)) ** 2
    expect = np.zeros(1 << n)
    expect[0] = expect[-1] = 0.5
    assert 0.5 * np.sum(np.abs(p - expect)) < 1e-7
    cluster = QC(n)
    for i in range(n):
        cluster.h(i)
    for i in range(n - 1):
        cluster.cz(i, i + 1)
    check(sim, cluster, mode)
    rev = QC(n)
    rev.h(n - 1)
    for i in range(n - 2, -1, -1):
        rev.cx(i + 1, i)
    check(sim, rev, mode)


def test_measurement_statistics(sim):
    shots = 50000
    qc = QC(4)
    for i in range(4):
        qc.h(i)
    qc.measure_all()
    counts = sim.run(qc, shots=shots, seed=42).get_counts()
    idx = np.array([int(k, 16) for k in counts for _ in range(counts[k])])
    p, _, _ = O.chi_square_p(idx, np.full(16, 1 / 16), shots)
    assert p > 1e-3
    bell = QC(2)
    bell.h(0).cx(0, 1)
    bell.measure_all()
    counts = sim.run(bell, shots=shots, seed=43).get_counts()
    assert set(counts) <= {"0x0", "0x3"} and sum(counts.values()) == shots


@pytest.mark.parametrize("mode", MODES)
def test_multi_controlled(sim, mode):
    n = 6
    for c0 in range(n):
        for c1 in range(n):
            for t in range(n):
                if len({c0, c1, t}) < 3:
                    continue
                qc = QC(n)
                qc.x(c0).x(c1).ccx(c0, c1, t)
                sv, ref = check(sim, qc, mode)
                assert np.array_equal(sv, ref.astype(np.complex64))   # permutation: exact
    n = 8
    qc = QC(n)
    qc.x(0).ccx(0, 1, 2)                                           # one control set: no flip
    sv, _ = sim.run_statevector(qc, mode=mode)
    assert sv[1] == 1
    qc = QC(n)
    qc.x(0).x(1).x(2).ccx(0, 1, 2)                                 # target already 1
    sv, _ = sim.run_statevector(qc, mode=mode)
    assert sv[3] == 1


@pytest.mark.parametrize("mode", MODES)
def test_adversarial(sim, mode):
    n = 10
    rng = np.random.default_rng(10)
    circuits = []
    qc = QC(n)
    for i in range(n):
        qc.x(i)
    for i in range(n - 1):
        qc.cx(i, i + 1)
    circuits.append(qc)                                            # all ones, CX cascade
    qc = QC(n)
    for i in range(n):
        qc.h(i)
    for _ in range(5):
        for i in range(0, n - 1, 2):
            qc.cx(i, i + 1)
        for i in range(1, n - 1, 2):
            qc.cx(i + 1, i)
    circuits.append(qc)                                            # alternating directions
    qc = QC(n)
    qc.h(0).cx(0, n - 1)
    circuits.append(qc)                                            # max-stride CX
    qc = QC(n)
    qc.h(0)
    for _ in range(100):
        qc.t(0)
    circuits.append(qc)                                            # 100 T on one qubit
    qc = QC(n)
    qc.h(0).h(1)
    for _ in range(50):
        qc.cx(0, 1).cx(1, 0)
    circuits.append(qc)                                            # CX ping-pong
    qc = QC(n)
    for i in range(n):
        qc.h(i)
    for _ in range(20):
        for i in range(n):
            qc.rz(rng.uniform(0, 2 * np.pi), i)
        for i in range(n - 1):
            qc.cz(i, i + 1)
    circuits.append(qc)                                            # diagonal-only, deep
    qc = QC(n)
    qc.h(0).x(0).y(0).z(0).s(0).sdg(0).t(0).tdg(0).sx(0).sxdg(0)
    qc.rx(0.5, 0).ry(0.7, 0).rz(0.9, 0).p(1.1, 0).u(0.3, 0.4, 0.5, 0)
    circuits.append(qc)                                            # every 1Q gate on q0
    qc = QC(n)
    for _ in range(30):
        for i in range(n):
            qc.rx(0.1 * (i + 1), i)
        for i in range(n - 1):
            qc.cx(i, i + 1)
    circuits.append(qc)                                            # interleaved 1Q / 2Q
    qc = QC(n)
    qc.h(0)
    for i in range(1, n):
        qc.cx(0, i)
    circuits.append(qc)                                            # fan-out
    qc = QC(n)
    for i in range(n - 1):
        qc.h(i).cx(i, n - 1)
    circuits.append(qc)                                            # fan-in
    for qc in circuits:
        check(sim, qc, mode)


@pytest.mark.parametrize("n", [16, 20, 24])
def test_large_qubit_scaling(sim, n):
    ghz = QC(n)
    ghz.h(0)
    for i in range(n - 1):
        ghz.cx(i, i + 1)
    sv, _ = sim.run_statevector(ghz)
    amp = 1 / math.sqrt(2)
    assert abs(sv[0] - amp) < 1e-7 and abs(sv[-1] - amp) < 1e-7 and np.count_nonzero(sv) == 2
    uni = QC(n)
    for i in range(n):
        uni.h(i)
    sv, _ = sim.run_statevector(uni)
    assert np.max(np.abs(np.abs(sv) - 2 ** (-n / 2))) < 1e-6 * 2 ** (-n / 2) * 16
    qc = QC(n)
    qc.h(0).cx(0, n - 1)
    sv, _ = sim.run_statevector(qc)
    assert abs(sv[0] - amp) < 1e-7 and abs(sv[1 | (1 << (n - 1))] - amp) < 1e-7
    rng = np.random.default_rng(42)
    deep = QC(n)
    for _ in range(50):
        deep.rx(rng.uniform(0, 2 * np.pi), int(rng.integers(n)))
        q0, q1 = rng.choice(n, 2, replace=False)
        deep.cx(int(q0), int(q1))
    ref = ref_sv(deep)
    for mode in MODES:
        check(sim, deep, mode, ref=ref)


@pytest.mark.parametrize("mode", MODES)
def test_repeated_gate_stress(sim, mode):
    n = 8
    qc = QC(n)
    for _ in range(1000):
        qc.h(0)
    sv, _ = check(sim, qc, mode)
    assert abs(abs(sv[0]) - 1) < 1e-5
    qc = QC(n)
    for _ in range(999):
        qc.x(0)
    sv, _ = sim.run_statevector(qc, mode=mode)
    assert sv[1] == 1 and np.count_nonzero(sv) == 1
    a, b = QC(n), QC(n)
    a.h(0)
    for _ in range(1000):
        a.cx(0, 1)
    b.h(0)
    assert O.fidelity(sim.run_statevector(a, mode=mode)[0], sim.run_statevector(b, mode=mode)[0]) >= F_MIN
    a, b = QC(n), QC(n)
    a.h(0)
    for _ in range(10000):
        a.rz(0.001, 0)
    b.h(0).rz(10.0, 0)
    f = O.fidelity(sim.run_statevector(a, mode=mode)[0], sim.run_statevector(b, mode=mode)[0])
    assert f > 1 - 1e-6, f                                         # the reference asks 0.99


@pytest.mark.parametrize("family", ["GHZ", "QFT", "EfficientSU2", "DeepRz", "Toffoli", "QV"])
def test_ab_suite_accuracy(sim, family):
    """The reference's A/B accuracy phase (tests/ab_test.py:240-285 of the
    reference): each circuit family at 4-20 qubits against the FP64 oracle
    (Aer CPU there), fidelity >= 1 - 1e-7; at 4 and 8 qubits the 8192-shot
    counts of run() also pass chi-square against the FP64 distribution."""
    from blackwell_backend import workloads as W
    for n in (4, 8, 12, 16, 20):
        qc = W.AB_SUITE[family](n)
        for mode in ("auto", "plain", "blocks") if n >= 12 else ("auto", "plain"):
            check(sim, qc, mode)
        if n <= 8:
            m = qc.copy()
            m.measure_all()
            counts = sim.run(m, shots=8192, seed=n).get_counts()
            idx = np.array([int(k, 16) for k in counts for _ in range(counts[k])])
            p, _, _ = O.chi_square_p(idx, O.probs(ref_sv(qc)), 8192)
            assert p > 1e-3, (family, n, p)
