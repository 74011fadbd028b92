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
    z = (re + 1j * im) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    return q * (d / np.abs(d))[:, np.newaxis, :]


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


def su4_layers(n, depth, seed):
    """[(pairs [(q0, q1)], unitaries [P, 4, 4] complex128)] per layer.

    Draw order per layer: the permutation, then the layer's complex Gaussian
    matrices (real parts, then imaginary parts); all QR factorisations run
    in one batched call afterwards (same matrices as one call per layer)."""
    rng = np.random.default_rng(seed)
    P = n // 2
    perms, zs = [], []
    for _ in range(depth):
        perms.append(rng.permutation(n))
        re = rng.standard_normal((P, 4, 4))
        im = rng.standard_normal((P, 4, 4))
        zs.append((re + 1j * im) / math.sqrt(2))
    if depth == 0 or P == 0:
        return [([], np.zeros((0, 4, 4), dtype=np.complex128)) for _ in range(depth)]
    q, r = np.linalg.qr(np.concatenate(zs))
    d = np.diagonal(r, axis1=-2, axis2=-1)
    us = (q * (d / np.abs(d))[:, np.newaxis, :]).reshape(depth, P, 4, 4)
    return [([(int(pm[2 * j]), int(pm[2 * j + 1])) for j in range(P)], us[k])
            for k, pm in enumerate(perms)]


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
