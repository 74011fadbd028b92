"""Pins of the FP64 oracle (CPU only).

Each check is independent of the oracle's own formulas: closed forms from
the reference's stress tests (tests/stress_test.py:134-836 of the
reference), brute-force index loops, numpy/scipy library routines and
published known answers.
"""
import cmath
import itertools
import math

import numpy as np
import pytest

from oracle import hamiltonian as OH
from oracle import statevector as O


def rand_psi(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


# ------------------------------------------------------------------ matrices
NAMES_0 = ["h", "x", "y", "z", "s", "sdg", "t", "tdg", "id", "sx", "sxdg"]


def test_named_gates_unitary():
    rng = np.random.default_rng(0)
    for nm in NAMES_0:
        u = O.gate_1q(nm)
        assert np.allclose(u.conj().T @ u, np.eye(2), atol=1e-15)
    for nm, k in [("rx", 1), ("ry", 1), ("rz", 1), ("p", 1), ("u1", 1), ("u2", 2), ("u3", 3), ("u", 3)]:
        u = O.gate_1q(nm, rng.uniform(-7, 7, k))
        assert np.allclose(u.conj().T @ u, np.eye(2), atol=1e-14)
    for nm, k in [("cx", 0), ("cz", 0), ("swap", 0), ("cp", 1), ("crz", 1), ("cu", 4)]:
        u = O.gate_2q_textbook(nm, rng.uniform(-7, 7, k))
        assert np.allclose(u.conj().T @ u, np.eye(4), atol=1e-14)


def test_gate_identities():
    g = O.gate_1q
    assert np.allclose(g("h") @ g("h"), np.eye(2))
    assert np.allclose(g("h") @ g("z") @ g("h"), g("x"))
    assert np.allclose(g("s") @ g("s"), g("z"))
    assert np.allclose(np.linalg.matrix_power(g("t"), 8), np.eye(2))
    assert np.allclose(g("t") @ g("tdg"), np.eye(2))
    assert np.allclose(g("sx") @ g("sx"), g("x"))
    assert np.allclose(g("rx", [2 * math.pi]), -np.eye(2))
    th = 0.731
    assert np.allclose(np.linalg.matrix_power(g("rz", [th / 1000]), 1000), g("rz", [th]))
    # u3(t, p, l) = e^{i(p+l)/2} rz(p) ry(t) rz(l)
    t, p, l = 0.3, -1.2, 2.2
    assert np.allclose(g("u3", [t, p, l]),
                       cmath.exp(0.5j * (p + l)) * g("rz", [p]) @ g("ry", [t]) @ g("rz", [l]))
    assert np.allclose(g("u2", [p, l]), g("u3", [math.pi / 2, p, l]))
    assert np.allclose(g("p", [0.4]), g("u1", [0.4]))
    # rx = H rz H
    assert np.allclose(g("rx", [0.9]), g("h") @ g("rz", [0.9]) @ g("h"))


def test_2q_textbook_definitions():
    cx = O.gate_2q_textbook("cx")
    # |c t> with c the MSB: |10> -> |11>
    e = np.eye(4)
    assert np.allclose(cx @ e[2], e[3]) and np.allclose(cx @ e[1], e[1])
    assert np.allclose(O.gate_2q_textbook("cz"), np.diag([1, 1, 1, -1]))
    assert np.allclose(O.swap_bits_4(cx), np.eye(4)[[0, 3, 2, 1]])


# ------------------------------------------------------------------ index math
def _brute_1q(psi, u, t):
    out = psi.copy()
    for i in range(psi.size):
        if not (i >> t) & 1:
            j = i | (1 << t)
            out[i] = u[0, 0] * psi[i] + u[0, 1] * psi[j]
            out[j] = u[1, 0] * psi[i] + u[1, 1] * psi[j]
    return out


def _brute_2q_kernel(psi, u, q0, q1):
    lo, hi = min(q0, q1), max(q0, q1)
    out = psi.copy()
    for i in range(psi.size):
        if (i >> lo) & 1 or (i >> hi) & 1:
            continue
        idx = [i, i | 1 << lo, i | 1 << hi, i | 1 << lo | 1 << hi]
        a = psi[idx]
        for r in range(4):
            out[idx[r]] = sum(u[r, c] * a[c] for c in range(4))
    return out


def test_apply_1q_bruteforce():
    rng = np.random.default_rng(1)
    for n in (1, 3, 5):
        for t in range(n):
            psi = rand_psi(n, t)
            u = O.gate_1q("u3", rng.uniform(-3, 3, 3))
            assert np.allclose(O.apply_1q(psi, u, t), _brute_1q(psi, u, t), atol=1e-14)


def test_apply_2q_kernel_bruteforce():
    rng = np.random.default_rng(2)
    n = 4
    for q0, q1 in itertools.permutations(range(n), 2):
        psi = rand_psi(n, q0 + 5 * q1)
        u = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
        assert np.allclose(O.apply_2q_kernel(psi, u, q0, q1), _brute_2q_kernel(psi, u, q0, q1),
                           atol=1e-13)


def test_controlled_bruteforce():
    rng = np.random.default_rng(3)
    n = 5
    for ctl, t in [([0], 1), ([4, 2], 0), ([1, 3, 0], 4), ([], 2)]:
        psi = rand_psi(n, t)
        u = O.gate_1q("u3", rng.uniform(-3, 3, 3))
        cm = sum(1 << c for c in ctl)
        ref = psi.copy()
        for i in range(psi.size):
            if (i >> t) & 1 == 0 and (i & cm) == cm:
                j = i | 1 << t
                ref[i] = u[0, 0] * psi[i] + u[0, 1] * psi[j]
                ref[j] = u[1, 0] * psi[i] + u[1, 1] * psi[j]
        assert np.allclose(O.apply_mc(psi, u, ctl, [t]), ref, atol=1e-14)


def test_permutations_bruteforce():
    n = 5
    psi = rand_psi(n, 9)
    for ctl, t in [([], 0), ([1], 3), ([0, 4], 2), ([1, 2, 3], 0)]:
        cm = sum(1 << c for c in ctl)
        ref = psi.copy()
        for i in range(psi.size):
            if (i & cm) == cm:
                ref[i] = psi[i ^ (1 << t)]
        assert np.array_equal(psi[O.mcx_perm(n, ctl, t)], ref)
    for ctl, a, b in [([], 0, 4), ([2], 1, 3)]:
        cm = sum(1 << c for c in ctl)
        ref = psi.copy()
        for i in range(psi.size):
            if (i & cm) == cm and ((i >> a) & 1) != ((i >> b) & 1):
                ref[i] = psi[i ^ (1 << a) ^ (1 << b)]
        assert np.array_equal(psi[O.mcswap_perm(n, ctl, a, b)], ref)


def test_diagonals_bruteforce():
    n = 4
    psi = rand_psi(n, 4)
    d = np.exp(1j * np.array([0.3, -1.1, 2.0, 0.7]))
    for q0, q1 in [(0, 1), (3, 1)]:
        lo, hi = min(q0, q1), max(q0, q1)
        ref = np.array([psi[i] * d[2 * ((i >> hi) & 1) + ((i >> lo) & 1)] for i in range(16)])
        assert np.allclose(O.apply_diag_2q(psi, d, q0, q1), ref)
    ref = np.array([psi[i] * (d[1] if (i >> 2) & 1 else 1) for i in range(16)])
    assert np.allclose(O.apply_diag_phase(psi, d[1], 2), ref)


def test_apply_diag_bruteforce_every_target():
    # reference rule (bwk/csrc/apply_diagonal_sm120.cu:43-59): each amplitude
    # is multiplied by diag[bit(idx, target)]; distinct d0, d1 catch a swap
    d = np.array([np.exp(0.4j) * 0.9, np.exp(-2.1j) * 1.1])
    for n in range(1, 6):
        psi = rand_psi(n, 40 + n)
        for t in range(n):
            ref = np.empty_like(psi)
            for i in range(psi.size):
                ref[i] = psi[i] * (d[1] if (i >> t) & 1 else d[0])
            assert np.allclose(O.apply_diag(psi, d, t), ref, rtol=1e-15, atol=0), (n, t)
            # a diagonal 2x2 through the generic matrix routine gives the same
            assert np.allclose(O.apply_diag(psi, d, t), O.apply_1q(psi, np.diag(d), t), atol=1e-15)


def _brute_mc2q(psi, g, controls, t0, t1):
    """The reference's controlled 4x4 (bwk/csrc/apply_mc2qgate_sm120.cu:88-123):
    groups enumerated with every affected bit zero, controls then set;
    a00, a01 (bit lo), a10 (bit hi), a11, row r = 2*bit(hi) + bit(lo)."""
    lo, hi = min(t0, t1), max(t0, t1)
    cm = sum(1 << c for c in controls)
    aff = cm | (1 << lo) | (1 << hi)
    out = psi.copy()
    for base in range(psi.size):
        if base & aff:
            continue
        i00 = base | cm
        idx = [i00, i00 | 1 << lo, i00 | 1 << hi, i00 | 1 << lo | 1 << hi]
        a = [psi[k] for k in idx]
        for r in range(4):
            out[idx[r]] = g[r][0] * a[0] + g[r][1] * a[1] + g[r][2] * a[2] + g[r][3] * a[3]
    return out


def test_apply_mc2q_kernel_bruteforce():
    rng = np.random.default_rng(12)
    for n in (3, 4, 5):
        for t0, t1 in itertools.permutations(range(n), 2):
            rest = [q for q in range(n) if q not in (t0, t1)]
            for k in range(0, min(3, len(rest)) + 1):
                for ctl in itertools.combinations(rest, k):
                    psi = rand_psi(n, 7 * t0 + t1 + 100 * k)
                    g = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
                    got = O.apply_mc2q_kernel(psi, g, list(ctl), t0, t1)
                    assert np.allclose(got, _brute_mc2q(psi, g, ctl, t0, t1), atol=1e-13), \
                        (n, t0, t1, ctl)
    # the argument order of the targets does not matter (row bit 0 = lower target)
    psi = rand_psi(5, 3)
    g = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
    assert np.array_equal(O.apply_mc2q_kernel(psi, g, [4], 1, 3), O.apply_mc2q_kernel(psi, g, [4], 3, 1))
    # a controlled SWAP matrix reproduces the Fredkin permutation
    sw = np.eye(4)[[0, 2, 1, 3]]
    assert np.allclose(O.apply_mc2q_kernel(psi, sw, [0, 4], 1, 3), psi[O.mcswap_perm(5, [0, 4], 1, 3)])


def test_probs_bruteforce():
    # reference rule (bwk/csrc/measure_sm120.cu:115-130): p_i = re^2 + im^2
    for n in (1, 4, 7):
        psi = rand_psi(n, 60 + n) * 1.3
        p = O.probs(psi)
        ref = np.array([(psi[i] * np.conj(psi[i])).real for i in range(psi.size)])
        assert np.allclose(p, ref, rtol=1e-15, atol=0)
        assert abs(p.sum() - 1.69) < 1e-12          # Parseval: the norm squared
    assert np.allclose(O.probs(np.array([0.6, 0.8j, -0.6 + 0.8j, 0])), [0.36, 0.64, 1.0, 0.0],
                       rtol=1e-15, atol=0)


# ------------------------------------------------------------------ closed forms
def test_ghz_uniform_bell():
    for n in (2, 5, 9):
        instrs = [("h", [0], [])] + [("cx", [q, q + 1], []) for q in range(n - 1)]
        psi = O.run_instructions(n, instrs)
        assert abs(psi[0] - 1 / math.sqrt(2)) < 1e-15 and abs(psi[-1] - 1 / math.sqrt(2)) < 1e-15
        assert np.sum(np.abs(psi) > 1e-12) == 2
        psi = O.run_instructions(n, [("h", [q], []) for q in range(n)])
        assert np.allclose(psi, 2 ** (-n / 2), atol=1e-15)


def test_circuit_equivalences():
    n = 4
    psi0 = rand_psi(n, 11)
    a = O.run_instructions(n, [("cx", [1, 3], [])], psi0)
    b = O.run_instructions(n, [("h", [3], []), ("cz", [1, 3], []), ("h", [3], [])], psi0)
    assert np.allclose(a, b)
    a = O.run_instructions(n, [("swap", [0, 2], [])], psi0)
    b = O.run_instructions(n, [("cx", [0, 2], []), ("cx", [2, 0], []), ("cx", [0, 2], [])], psi0)
    assert np.allclose(a, b)
    a = O.run_instructions(n, [("cswap", [3, 0, 1], [])], psi0)
    assert np.allclose(a, psi0[O.mcswap_perm(n, [3], 0, 1)])


def test_qft_is_dft_column():
    n = 5
    N = 1 << n
    for x in (0, 3, 17):
        instrs = [("x", [q], []) for q in range(n) if x >> q & 1]
        for j in reversed(range(n)):
            instrs.append(("h", [j], []))
            for k in reversed(range(j)):
                instrs.append(("cp", [k, j], [math.pi / 2 ** (j - k)]))
        for q in range(n // 2):
            instrs.append(("swap", [q, n - 1 - q], []))
        psi = O.run_instructions(n, instrs)
        ref = np.exp(2j * math.pi * x * np.arange(N) / N) / math.sqrt(N)
        assert np.allclose(psi, ref, atol=1e-13)


def test_c_oracle_matches_numpy_oracle():
    from oracle import build_ref
    lib = build_ref.load()
    rng = np.random.default_rng(5)
    n = 10
    psi = np.zeros(1 << n, dtype=np.complex128)
    lib.sv_init(psi.ctypes.data, n)
    ref = O.zero_state(n)
    import ctypes
    pd = ctypes.POINTER(ctypes.c_double)
    for k in range(30):
        if k % 3 == 0:
            t = int(rng.integers(n))
            u = O.gate_1q("u3", rng.uniform(-3, 3, 3))
            v = np.empty(8); v[0::2] = u.reshape(-1).real; v[1::2] = u.reshape(-1).imag
            lib.sv_apply_1q(psi.ctypes.data, n, v.ctypes.data_as(pd), t)
            ref = O.apply_1q(ref, u, t)
        elif k % 3 == 1:
            a, b = (int(x) for x in rng.choice(n, 2, replace=False))
            z = rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4))
            u = np.linalg.qr(z)[0]
            v = np.empty(32); v[0::2] = u.reshape(-1).real; v[1::2] = u.reshape(-1).imag
            lib.sv_apply_2q(psi.ctypes.data, n, v.ctypes.data_as(pd), a, b)
            ref = O.apply_matrix(ref, u, [a, b])
        else:
            qs = [int(x) for x in rng.choice(n, 3, replace=False)]
            arr = (ctypes.c_int * 2)(*qs[:2])
            lib.sv_apply_mcx(psi.ctypes.data, n, arr, 2, qs[2])
            ref = ref[O.mcx_perm(n, qs[:2], qs[2])]
    assert np.allclose(psi, ref, atol=1e-13)
    assert abs(lib.sv_norm2(psi.ctypes.data, n) - 1) < 1e-12


# ------------------------------------------------------------------ readout
def test_sampling_rule():
    psi = np.array([0, 0.6, 0, 0.8], dtype=np.complex128)   # p = [0, .36, 0, .64]
    idx = O.sample(psi, [0.0, 0.35, 0.3599999, 0.36, 0.9999])
    assert list(idx) == [1, 1, 1, 3, 3]
    rng = np.random.default_rng(0)
    bell = np.zeros(4, dtype=np.complex128)
    bell[[0, 3]] = 1 / math.sqrt(2)
    s = O.sample(bell, rng.random(50000))
    assert set(np.unique(s)) <= {0, 3}
    p, _, _ = O.chi_square_p(s, np.abs(bell) ** 2, 50000)
    assert p > 1e-3


def test_measure_rule():
    psi = np.array([0.6, 0.8], dtype=np.complex128)
    o, st = O.measure(psi, 0, 0.35)
    assert o == 0 and np.allclose(st, [1, 0])
    o, st = O.measure(psi, 0, 0.36)
    assert o == 1 and np.allclose(st, [0, 1])


def test_splitmix64_known_answer():
    # SplitMix64 seeded with 0: first output 0xE220A8397B1DCDAF (Steele,
    # Lea, Flood 2014 reference generator; widely published test vector)
    u = O.splitmix_uniforms(0, 2)
    assert u[0] == (0xE220A8397B1DCDAF >> 11) * 2.0 ** -53
    assert 0 <= u.min() and u.max() < 1


def test_fidelity_and_phase_alignment():
    a = rand_psi(6, 1)
    assert abs(O.fidelity(a, 1j * a) - 1) < 1e-15
    assert O.max_abs_diff_phase_aligned(a, np.exp(0.3j) * a) < 1e-15


# ------------------------------------------------------------------ Hamiltonians
def test_pauli_dense_elements():
    X, Y, Z = O.X, O.Y, O.Z
    m = OH.pauli_dense(["XZ"], [1.0])      # qubit 1 = X, qubit 0 = Z
    assert np.allclose(m, np.kron(X, Z))
    m = OH.pauli_dense(["IY", "ZI"], [0.5, -2.0])
    assert np.allclose(m, 0.5 * np.kron(np.eye(2), Y) - 2.0 * np.kron(Z, np.eye(2)))
    sp = OH.pauli_sparse(["XY", "ZZ", "YI"], [0.3, 1.1, -0.7]).toarray()
    assert np.allclose(sp, OH.pauli_dense(["XY", "ZZ", "YI"], [0.3, 1.1, -0.7]))


def test_chebyshev_coefficients_jacobi_anger():
    # e^{-i a x} = sum_k c_k T_k(x) with c_k = (2 - d_k0) (-i)^k J_k(a)
    for alpha in (0.5, 7.0, 40.0):
        c, K, tail = OH.chebyshev_coefficients(alpha, 1.0, precision=1e-12)
        x = np.linspace(-1, 1, 41)
        approx = np.polynomial.chebyshev.chebval(x, c)
        assert np.max(np.abs(approx - np.exp(-1j * alpha * x))) < 1e-10
        assert tail < 1e-12
    cpos, _, _ = OH.chebyshev_coefficients(3.0, 2.0)
    cneg, _, _ = OH.chebyshev_coefficients(-3.0, 2.0)
    assert np.allclose(cneg, np.conj(cpos))


def test_chebyshev_evolution_matches_expm():
    rng = np.random.default_rng(3)
    labels = ["".join(rng.choice(list("IXYZ"), 6)) for _ in range(10)]
    coeffs = rng.uniform(-1, 1, 10)
    h = OH.pauli_sparse(labels, coeffs)
    psi = rand_psi(6, 2)
    bound = float(np.sum(np.abs(coeffs)))
    for t in (0.5, 3.0):
        a = OH.chebyshev_evolve_fp64(h, psi, t, bound, 1e-13)
        b = OH.expm_reference(h, psi, t)
        assert np.max(np.abs(a - b)) < 1e-10
    # H = Z on one qubit: exact phases (the reference's analytic case)
    hz = OH.pauli_sparse(["Z"], [1.0])
    psi = np.array([0.6, 0.8j])
    out = OH.chebyshev_evolve_fp64(hz, psi, 1.3, 1.0, 1e-13)
    assert np.allclose(out, [0.6 * cmath.exp(-1.3j), 0.8j * cmath.exp(1.3j)], atol=1e-12)


def test_pauli_csr_grouped_matches_kronecker():
    import scipy.sparse as sp
    from blackwell_backend import workloads as W
    rng = np.random.default_rng(4)
    for n in (1, 3, 6):
        labels = ["".join(rng.choice(list("IXYZ"), n)) for _ in range(12)]
        coeffs = rng.uniform(-1, 1, 12) + 1j * rng.uniform(-1, 1, 12) * (n % 2)
        rp, ci, v = OH.pauli_csr_grouped(labels, coeffs, n)
        m = sp.csr_matrix((v, ci, rp), shape=(1 << n, 1 << n)).toarray()
        assert np.allclose(m, OH.pauli_dense(labels, coeffs), atol=1e-14)
    spo = W.heisenberg_chain(6, seed=1)
    rp, ci, v = OH.pauli_csr_grouped(spo.labels, spo.coeffs, 6)
    m = sp.csr_matrix((v, ci, rp), shape=(64, 64)).toarray()
    assert np.allclose(m, OH.pauli_dense(spo.labels, spo.coeffs), atol=1e-14)
    assert np.allclose(m, m.conj().T)
    # exact zeros removed: XX+YY cancel on rows where the two bits agree
    assert len(ci) < 64 * (1 + 5 + 6)
    assert np.all(v != 0)


def test_c_pauli_matvec_matches_kronecker():
    # (P v)[i] from the strings (oracle/cheb_fp64.c) against explicit np.kron
    # products of the 2x2 Paulis, with complex coefficients and every Pauli
    rng = np.random.default_rng(21)
    for n in (1, 3, 6):
        labels = ["".join(rng.choice(list("IXYZ"), n)) for _ in range(14)]
        coeffs = rng.uniform(-1, 1, 14) + 1j * rng.uniform(-1, 1, 14)
        v = np.stack([rand_psi(n, 70 + k) for k in range(3)])
        got = OH.matvec_c(labels, coeffs, n, v, scale=0.5)
        want = 0.5 * (OH.pauli_dense(labels, coeffs) @ v.T).T
        assert np.allclose(got, want, atol=1e-13), n
    # the single-qubit Paulis in closed form
    e0, e1 = np.array([1, 0], complex), np.array([0, 1], complex)
    assert np.allclose(OH.matvec_c(["Y"], [1.0], 1, e0), [[0, 1j]])
    assert np.allclose(OH.matvec_c(["Y"], [1.0], 1, e1), [[-1j, 0]])
    assert np.allclose(OH.matvec_c(["Z"], [1.0], 1, e1), [[0, -1]])


def test_c_chebyshev_matches_expm_multiply():
    from blackwell_backend import workloads as W
    for n, t in ((8, 1.0), (10, 5.0)):
        op = W.heisenberg_chain(n, seed=3)
        bound = float(np.sum(np.abs(op.coeffs)))
        v = np.stack([rand_psi(n, 80 + k) for k in range(3)])
        got, K = OH.chebyshev_evolve_c(op.labels, op.coeffs, n, v, t, bound)
        h = OH.pauli_sparse(op.labels, op.coeffs)
        for k in range(3):
            want = OH.expm_reference(h, v[k], t)
            assert np.max(np.abs(got[k] - want)) < 1e-11, (n, t, k)
        assert K > bound * t
    # H = Z on one qubit: exact phases
    out, _ = OH.chebyshev_evolve_c(["Z"], [1.0], 1, np.array([0.6, 0.8j]), 1.3, 1.0)
    assert np.allclose(out[0], [0.6 * cmath.exp(-1.3j), 0.8j * cmath.exp(1.3j)], atol=1e-12)


def test_constant_group_entries_match_oracle_csr():
    """The host table of row-independent group entries (X-only groups) used by
    the matrix-free Chebyshev term: every such group's entry is, in every
    row, the oracle CSR's value at column row ^ X mask rounded to complex64
    (absent where the sum is zero), and every group with a Y or Z term is
    marked NaN."""
    from blackwell_backend.hamiltonian import pauli_groups, constant_group_entries
    n = 5
    labels = ["IIIXX", "IIIXX", "IXIXI", "XIIII", "XIIIZ", "IIIII", "ZZIII", "IYYII", "IXXII",
              "XXXXX", "XXXXX"]
    coeffs = [0.3, 0.45, -0.2, 0.7, 0.1, 1.25, -0.4, 0.35, 0.35, 0.6, -0.6]
    xm, toff, sm, coef = pauli_groups(labels, coeffs)
    gc = constant_group_entries(toff, sm, coef)
    rp, ci, v = OH.pauli_csr_grouped(labels, coeffs, n)
    seen_const = seen_zero = seen_row = 0
    for g in range(len(xm)):
        if np.any(sm[toff[g]:toff[g + 1]] != 0):
            assert np.isnan(gc[g, 0])
            seen_row += 1
            continue
        want = complex(gc[g, 0], gc[g, 1])
        seen_const += 1
        seen_zero += want == 0
        for r in range(1 << n):
            cols = ci[rp[r]:rp[r + 1]]
            hit = np.nonzero(cols == (r ^ int(xm[g])))[0]
            if want == 0:
                assert hit.size == 0
            else:
                assert hit.size == 1 and np.complex64(v[rp[r] + hit[0]]) == np.complex64(want)
    assert seen_const >= 3 and seen_zero == 1 and seen_row >= 3
