"""Seeded synthetic workloads (the benchmark configurations).

c1_rotation_ladder   12 qubits, 12 layers of [H all; Ry, Rz all (U[0,2pi));
                     CNOT q -> q+1]                              (564 gates)
su4_volume           n qubits, depth d: per layer a random permutation in
                     floor(n/2) pairs, each with a Haar 4x4 (QR of a complex
                     Gaussian, R-diagonal phase folded into Q; the recipe of
                     bwk/python/blackwell_kernels/qv8.py:55-73); row bit 0 of
                     the 4x4 <-> the first qubit of the pair
c3_rotation_mcx      n qubits, reps of [Ry, Rz all; CNOT ladder; 4 MCX with 3
                     random controls and 1 random target]
serving_requests     request stream: n = 10 + k, P(k) ~ 0.25 * 0.75^k,
                     k = 0..14, Poisson arrivals, su4_volume(n, depth=n)
heisenberg_chain     sum J_i Z_i Z_i+1 + K_i (X_i X_i+1 + Y_i Y_i+1) + h_i X_i,
                     J, K, h ~ U[-1, 1], as a SparsePauliOp
"""
import math

import numpy as np

from .circuit import QuantumCircuit
from .hamiltonian import SparsePauliOp


def haar_unitary(rng, dim):
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))[np.newaxis, :]


def haar_batch(rng, count, dim=4):
    re = rng.standard_normal((count, dim, dim))
    im = rng.standard_normal((count, dim, dim))
    return qr_phase_fixed((re + 1j * im) / math.sqrt(2))


def c1_rotation_ladder(n=12, layers=12, seed=1234):
    rng = np.random.default_rng(seed)
    qc = QuantumCircuit(n, name=f"rot_ladder_{n}q_{layers}l")
    for _ in range(layers):
        th = rng.uniform(0, 2 * math.pi, n)
        ph = rng.uniform(0, 2 * math.pi, n)
        for q in range(n):
            qc.h(q)
        for q in range(n):
            qc.ry(th[q], q)
            qc.rz(ph[q], q)
        for q in range(n - 1):
            qc.cx(q, q + 1)
    return qc


def su4_arrays(n, depth, seed):
    """(pairs int64 [depth*P, 2], unitaries complex128 [depth*P, 4, 4]),
    P = floor(n/2), gates in layer order.

    Draws, in this order from np.random.default_rng(seed): the depth
    permutations (one call, row-wise), then the real parts and then the
    imaginary parts of all depth*P complex Gaussian 4x4 matrices; the QR
    factorisations run as one batched call."""
    rng = np.random.default_rng(seed)
    P = n // 2
    G = depth * P
    perms = rng.permuted(np.tile(np.arange(n, dtype=np.int64), (depth, 1)), axis=1)
    pairs = perms[:, :2 * P].reshape(G, 2)
    if G == 0:
        return pairs, np.zeros((0, 4, 4), dtype=np.complex128)
    re = rng.standard_normal((G, 4, 4))
    im = rng.standard_normal((G, 4, 4))
    return pairs, qr_phase_fixed((re + 1j * im) / math.sqrt(2))


def qr_phase_fixed(z):
    """Q of the QR factorisation with a positive real R diagonal, for a batch
    [G, d, d] — i.e. Q * diag(R_ii / |R_ii|) of any QR — by modified
    Gram-Schmidt on the columns (vectorised over the batch)."""
    z = np.array(z, dtype=np.complex128)
    d = z.shape[-1]
    q = np.empty_like(z)
    for j in range(d):
        v = z[:, :, j]
        for k in range(j):
            qk = q[:, :, k]
            v = v - np.einsum("gi,gi->g", qk.conj(), v)[:, None] * qk
        q[:, :, j] = v / np.linalg.norm(v, axis=1)[:, None]
    return q


def su4_layers(n, depth, seed):
    """[(pairs [(q0, q1)], unitaries [P, 4, 4] complex128)] per layer (the
    same gates as su4_arrays)."""
    pairs, us = su4_arrays(n, depth, seed)
    P = n // 2
    return [([(int(a), int(b)) for a, b in pairs[k * P:(k + 1) * P]], us[k * P:(k + 1) * P])
            for k in range(depth)]


def su4_volume(n=28, depth=28, seed=2024):
    qc = QuantumCircuit(n, name=f"su4_volume_{n}q_d{depth}")
    for pairs, us in su4_layers(n, depth, seed):
        for (a, b), u in zip(pairs, us):
            qc.unitary(u, [a, b])
    return qc


def c3_rotation_mcx(n=32, reps=8, seed=77, n_mcx=4, n_controls=3):
    rng = np.random.default_rng(seed)
    qc = QuantumCircuit(n, name=f"rot_mcx_{n}q_{reps}r")
    for _ in range(reps):
        th = rng.uniform(0, 2 * math.pi, n)
        ph = rng.uniform(0, 2 * math.pi, n)
        for q in range(n):
            qc.ry(th[q], q)
            qc.rz(ph[q], q)
        for q in range(n - 1):
            qc.cx(q, q + 1)
        for _ in range(n_mcx):
            qs = rng.choice(n, n_controls + 1, replace=False)
            qc.mcx([int(c) for c in qs[:-1]], int(qs[-1]))
    return qc


def serving_qubit_counts(count, seed, n_min=10, n_max=24, p=0.25):
    rng = np.random.default_rng(seed)
    ks = np.arange(n_max - n_min + 1)
    w = p * (1 - p) ** ks
    w /= w.sum()
    return n_min + rng.choice(ks, size=count, p=w)


def serving_requests(count=4096, seed=99, rate=1000.0, shots=1024):
    """[(arrival_s, n_qubits, circuit_seed, shots)] with Poisson arrivals."""
    rng = np.random.default_rng(seed + 1)
    ns = serving_qubit_counts(count, seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    return [(float(arrivals[i]), int(ns[i]), int(seed * 100003 + i), shots) for i in range(count)]


def heisenberg_chain(n=22, seed=5):
    rng = np.random.default_rng(seed)
    J = rng.uniform(-1, 1, n - 1)
    K = rng.uniform(-1, 1, n - 1)
    h = rng.uniform(-1, 1, n)
    terms = []

    def lab(ops):
        s = ["I"] * n
        for q, ch in ops:
            s[n - 1 - q] = ch
        return "".join(s)

    for i in range(n - 1):
        terms.append((lab([(i, "Z"), (i + 1, "Z")]), J[i]))
        terms.append((lab([(i, "X"), (i + 1, "X")]), K[i]))
        terms.append((lab([(i, "Y"), (i + 1, "Y")]), K[i]))
    for i in range(n):
        terms.append((lab([(i, "X")]), h[i]))
    return SparsePauliOp.from_list(terms)


# ------------------------------------------------ the reference's A/B circuit suite
# The circuit families of the reference's performance table (tests/ab_test.py:43-118,
# README.md:121-152), built without qiskit's circuit library: the same gate
# structure and counts; angles and Haar draws come from this module's seeded
# generators.

def ghz(n):
    """H on qubit 0, then CNOT q -> q+1 (ab_test.py:43-49)."""
    qc = QuantumCircuit(n, name=f"ghz_{n}")
    qc.h(0)
    for i in range(n - 1):
        qc.cx(i, i + 1)
    return qc


def qft(n):
    """Textbook QFT as H, controlled phases and the final swaps: the gates of
    qiskit's QFT(n).decompose() used at ab_test.py:52-55 (n(n+1)/2 + n//2)."""
    qc = QuantumCircuit(n, name=f"qft_{n}")
    for j in reversed(range(n)):
        qc.h(j)
        for k in reversed(range(j)):
            qc.cp(math.pi / (1 << (j - k)), j, k)
    for i in range(n // 2):
        qc.swap(i, n - 1 - i)
    return qc


def efficient_su2(n, reps=2, seed=42):
    """RY and RZ layers on every qubit around `reps` reverse-linear CX
    entanglers, angles U[0, 2pi) in parameter order (ab_test.py:74-79)."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0, 2 * math.pi, 2 * n * (reps + 1))
    qc = QuantumCircuit(n, name=f"su2_{n}_r{reps}")
    k = 0
    for layer in range(reps + 1):
        for q in range(n):
            qc.ry(theta[k + q], q)
        for q in range(n):
            qc.rz(theta[k + n + q], q)
        k += 2 * n
        if layer < reps:
            for q in reversed(range(n - 1)):
                qc.cx(q, q + 1)
    return qc


def deep_rz(n, depth=50, seed=42):
    """Rz on every qubit, then CX on even and odd neighbour pairs, `depth`
    times (ab_test.py:105-118)."""
    rng = np.random.default_rng(seed)
    qc = QuantumCircuit(n, name=f"deep_rz_{n}_d{depth}")
    for _ in range(depth):
        for q in range(n):
            qc.rz(rng.uniform(0, 2 * math.pi), q)
        for q in range(0, n - 1, 2):
            qc.cx(q, q + 1)
        for q in range(1, n - 1, 2):
            qc.cx(q, q + 1)
    return qc


def toffoli_chain(n):
    """H on every qubit, then CCX(i, i+1, i+2) (ab_test.py:91-102)."""
    if n < 3:
        raise ValueError("a Toffoli chain needs at least 3 qubits")
    qc = QuantumCircuit(n, name=f"toffoli_{n}")
    for i in range(n):
        qc.h(i)
    for i in range(n - 2):
        qc.ccx(i, i + 1, i + 2)
    return qc


def quantum_volume(n, seed=42):
    """Depth-n square circuit of Haar SU(4) on seeded random pairs, the form
    of QuantumVolume(n, seed) at ab_test.py:82-88."""
    qc = su4_volume(n, n, seed)
    qc.name = f"qv_{n}_d{n}"
    return qc


AB_SUITE = {
    "GHZ": ghz,
    "QFT": qft,
    "EfficientSU2": efficient_su2,
    "DeepRz": deep_rz,
    "Toffoli": toffoli_chain,
    "QV": quantum_volume,
}
