"""Pins of the Pauli-string and noise-channel oracle (CPU only).

Every check is independent of the oracle's formulas: explicit numpy.kron
products of the 2x2 Pauli matrices, closed forms on small states, Kraus
completeness, and the trajectory identity sum_b p_b |psi_b><psi_b| =
sum_k E_k rho E_k^+ that a correct stochastic unravelling of a channel
must satisfy.
"""
import math

import numpy as np
import pytest

from oracle import observables as OB


def rand_psi(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def test_apply_pauli_string_matches_kron_every_string_3q():
    n = 3
    psi = rand_psi(n, 1)
    for x in range(1 << n):
        for z in range(1 << n):
            want = OB.pauli_dense(n, x, z) @ psi
            assert np.array_equal(OB.apply_pauli_string(psi, x, z), want), (x, z)


def test_apply_pauli_string_exact_in_complex64():
    psi = rand_psi(6, 2).astype(np.complex64)
    for x, z in [(0b101101, 0b010111), (1, 1), (0, 0b111111), (0b100000, 0)]:
        out = OB.apply_pauli_string(psi, x, z)
        assert out.dtype == np.complex64
        # every output component is an input component up to sign
        src = psi[np.arange(psi.size) ^ x]
        a = np.sort(np.abs(np.concatenate([out.real, out.imag])))
        b = np.sort(np.abs(np.concatenate([src.real, src.imag])))
        assert np.array_equal(a, b)
        back = OB.apply_pauli_string(out, x, z)     # P^2 = I for a Hermitian string
        assert np.array_equal(back, psi)


def test_expectation_matches_kron_every_string_3q_and_random_5q():
    n = 3
    psi = rand_psi(n, 3) * 0.7          # unnormalised on purpose
    for x in range(1 << n):
        for z in range(1 << n):
            want = np.vdot(psi, OB.pauli_dense(n, x, z) @ psi)
            assert abs(want.imag) < 1e-14
            assert abs(OB.expectation_pauli_string(psi, x, z) - want.real) < 1e-14
    rng = np.random.default_rng(4)
    psi = rand_psi(5, 5)
    for _ in range(40):
        x, z = int(rng.integers(0, 32)), int(rng.integers(0, 32))
        want = np.vdot(psi, OB.pauli_dense(5, x, z) @ psi).real
        assert abs(OB.expectation_pauli_string(psi, x, z) - want) < 1e-13


def test_expectation_closed_forms():
    plus = np.array([1, 1], dtype=complex) / math.sqrt(2)
    plus_i = np.array([1, 1j], dtype=complex) / math.sqrt(2)
    one = np.array([0, 1], dtype=complex)
    assert abs(OB.expectation_pauli(plus, 0, 0, 1) - 1) < 1e-15        # <+|X|+> = 1
    assert abs(OB.expectation_pauli(plus, 0, 0, 2)) < 1e-15            # <+|Y|+> = 0
    assert abs(OB.expectation_pauli(plus_i, 0, 0, 2) - 1) < 1e-15      # <+i|Y|+i> = 1
    assert abs(OB.expectation_z(one, 1) + 1) < 1e-15                   # <1|Z|1> = -1
    bell = np.array([1, 0, 0, 1], dtype=complex) / math.sqrt(2)
    assert abs(OB.expectation_pauli_string(bell, 3, 0) - 1) < 1e-15    # XX
    assert abs(OB.expectation_pauli_string(bell, 3, 3) + 1) < 1e-15    # YY
    assert abs(OB.expectation_z(bell, 3) - 1) < 1e-15                  # ZZ
    assert abs(OB.expectation_z(bell, 1)) < 1e-15                      # ZI
    # the reference's argument form ignores a z bit on the X/Y qubit
    psi = rand_psi(4, 6)
    assert OB.expectation_pauli(psi, 0b1011, 1, 1) == OB.expectation_pauli(psi, 0b1001, 1, 1)
    assert OB.expectation_pauli(psi, 0b1011, 1, 2) == OB.expectation_pauli(psi, 0b1001, 1, 2)


def test_kraus_completeness():
    for kind in ("damping", "dephasing"):
        for g in (0.0, 0.1, 0.5, 1.0):
            e0, e1 = OB.kraus(kind, g)
            assert np.allclose(e0.conj().T @ e0 + e1.conj().T @ e1, np.eye(2), atol=1e-15)


def _outer(v):
    return np.outer(v, v.conj())


@pytest.mark.parametrize("kind,g", [("damping", 0.3), ("damping", 0.9), ("dephasing", 0.25),
                                    ("dephasing", 0.8)])
def test_damping_trajectories_average_to_the_channel(kind, g):
    n, t = 3, 1
    psi = rand_psi(n, 7)
    p1 = OB.branch_probability(psi, t, g)
    assert 0 < p1 < 1
    b1, s1 = OB.apply_damping_channel(kind, psi, t, g, 0.5 * p1)          # u inside [0, p1)
    b0, s0 = OB.apply_damping_channel(kind, psi, t, g, 0.5 * (1 + p1))    # u inside [p1, 1)
    assert (b0, b1) == (0, 1)
    assert abs(np.linalg.norm(s0) - 1) < 1e-14 and abs(np.linalg.norm(s1) - 1) < 1e-14
    avg = p1 * _outer(s1) + (1 - p1) * _outer(s0)
    want = OB.channel_on_density(OB.kraus(kind, g), _outer(psi), t, n)
    assert np.allclose(avg, want, atol=1e-14)


def test_depolarizing_trajectories_average_to_the_channel():
    n, t, p = 3, 2, 0.6
    psi = rand_psi(n, 8)
    # u at the middle of each branch's interval: [0,1-p), then three of width p/3
    mids = [(1 - p) / 2] + [1 - p + (k + 0.5) * p / 3 for k in range(3)]
    widths = [1 - p] + [p / 3] * 3
    avg = np.zeros((1 << n, 1 << n), dtype=complex)
    for b, (u, w) in enumerate(zip(mids, widths)):
        got_b, out = OB.apply_depolarizing(psi, t, p, u)
        assert got_b == b
        avg += w * _outer(out)
    paulis = [OB.I2, OB.PX, OB.PY, OB.PZ]
    ks = [math.sqrt(1 - p) * paulis[0]] + [math.sqrt(p / 3) * m for m in paulis[1:]]
    want = OB.channel_on_density(ks, _outer(psi), t, n)
    assert np.allclose(avg, want, atol=1e-14)


def test_damping_closed_forms():
    one = np.array([0, 1], dtype=complex)
    b, out = OB.apply_amplitude_damping(one, 0, 0.3, 0.1)      # u < gamma: decays to |0>
    assert b == 1 and np.allclose(out, [1, 0])
    b, out = OB.apply_amplitude_damping(one, 0, 0.3, 0.5)      # stays |1>
    assert b == 0 and np.allclose(out, [0, 1])
    plus = np.array([1, 1], dtype=complex) / math.sqrt(2)
    lam = 0.36
    b, out = OB.apply_dephasing(plus, 0, lam, 0.9)
    want = np.array([1, math.sqrt(1 - lam)]) / math.sqrt(2 - lam)
    assert b == 0 and np.allclose(out, want, atol=1e-15)
    b, out = OB.apply_dephasing(plus, 0, lam, 0.01)
    assert b == 1 and np.allclose(out, [0, 1])
