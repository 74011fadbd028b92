"""Parity of the Pauli-string expectations, Pauli-string application and
noise channels of libqsim_gfx950.so against the FP64 oracle.

Bars:
  * expectations: FP64 sums of exact FP64 products of the same complex64
    amplitudes, so they agree with the oracle to 1e-12 (summation order
    only) and are bitwise reproducible run to run;
  * Pauli application and the depolarizing channel: bit-exact (components
    are moved and negated only);
  * amplitude damping / dephasing: the same branch as the oracle for u on
    either side of the branch probability, amplitudes within the
    compensated-gate bar (1e-6 relative), unit norm afterwards.
"""
import numpy as np
import pytest
import torch

from oracle import observables as OB
from helpers import rand_state, to_dev, close

pytestmark = pytest.mark.gpu


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


def _terms(n, rng, count):
    lim = 1 << n
    xs = [0, 0, 1, lim - 1, lim >> 1] + [int(rng.integers(0, lim)) for _ in range(count)]
    zs = [0, lim - 1, 1, lim - 1, 0] + [int(rng.integers(0, lim)) for _ in range(count)]
    if n >= 2:
        xs += [3, 3, 2]
        zs += [3, 0, 2]
    return xs, zs


@pytest.mark.parametrize("n", [1, 2, 5, 7, 8, 11, 14])
def test_expectation_paulis_match_oracle(bk, dev, n):
    rng = np.random.default_rng(10 + n)
    psi = (rand_state(n, n) * np.float32(0.9)).astype(np.complex64)
    st = to_dev(psi, dev)
    xs, zs = _terms(n, rng, 24)
    # groups larger than one pass (8 terms) sharing an x mask
    xs += [0] * 11 + [(1 << n) - 1] * 9
    zs += [int(rng.integers(0, 1 << n)) for _ in range(20)]
    got = bk.expectation_paulis(st, xs, zs).cpu().numpy()
    again = bk.expectation_paulis(st, xs, zs).cpu().numpy()
    p64 = psi.astype(np.complex128)
    want = np.array([OB.expectation_pauli_string(p64, x, z) for x, z in zip(xs, zs)])
    assert np.max(np.abs(got - want)) < 1e-12, (n, np.max(np.abs(got - want)))
    assert np.array_equal(got.view(np.uint64), again.view(np.uint64))
    # the state is read only
    assert np.array_equal(_bits(st.cpu().numpy()), _bits(psi))


@pytest.mark.parametrize("n", [1, 4, 9, 12])
def test_reference_expectation_forms(bk, dev, n):
    rng = np.random.default_rng(20 + n)
    psi = rand_state(n, 30 + n)
    st = to_dev(psi, dev)
    p64 = psi.astype(np.complex128)
    for _ in range(6):
        zm = int(rng.integers(0, 1 << n))
        assert abs(bk.expectation_z(st, zm) - OB.expectation_z(p64, zm)) < 1e-12
        q = int(rng.integers(0, n))
        for op in (0, 1, 2):
            got = bk.expectation_pauli(st, zm | (1 << q), q, op)
            want = OB.expectation_pauli(p64, zm | (1 << q), q, op)
            assert abs(got - want) < 1e-12, (n, zm, q, op)
    with pytest.raises(RuntimeError):
        bk.expectation_pauli(st, 0, 0, 3)
    with pytest.raises(RuntimeError):
        bk.expectation_z(st, 1 << n)
    with pytest.raises(RuntimeError):
        bk.expectation_paulis(st, [1 << n], [0])


def test_expectation_closed_forms(bk, dev):
    bell = torch.zeros(4, dtype=torch.complex64, device=dev)
    bell[0] = bell[3] = 2 ** -0.5
    xx, yy, zz = bk.expectation_paulis(bell, [3, 3, 0], [0, 3, 3]).tolist()
    assert abs(xx - 1) < 1e-7 and abs(yy + 1) < 1e-7 and abs(zz - 1) < 1e-7


@pytest.mark.parametrize("n", [1, 2, 3, 6, 7, 8, 9, 12])
def test_apply_pauli_string_bit_exact(bk, dev, n):
    rng = np.random.default_rng(40 + n)
    lim = 1 << n
    cases = [(0, lim - 1), (lim - 1, 0), (lim - 1, lim - 1), (1, 0), (0, 1), (lim >> 1, lim >> 1)]
    if n >= 8:
        hi2 = 0b11 << (n - 2)
        cases += [(0, hi2), (hi2, 0), ((1 << (n - 1)) | 0b101, (1 << (n - 2)) | 0b11), (0, lim - 128)]
    cases += [(int(rng.integers(0, lim)), int(rng.integers(0, lim))) for _ in range(6)]
    for x, z in cases:
        psi = rand_state(n, 1000 * n + x + 7 * z)
        st = to_dev(psi, dev)
        bk.apply_pauli_string(st, x, z)
        want = OB.apply_pauli_string(psi, x, z)
        assert np.array_equal(_bits(st.cpu().numpy()), _bits(want)), (n, x, z)


@pytest.mark.parametrize("n,t", [(1, 0), (3, 1), (8, 7), (10, 3), (12, 11)])
def test_depolarizing_bit_exact(bk, dev, n, t):
    p = 0.6
    for b, u in enumerate([0.2, 1 - p + p / 6, 1 - p + p / 2, 1 - p + 5 * p / 6]):
        psi = rand_state(n, 50 + 10 * n + b)
        st = to_dev(psi, dev)
        got_b = bk.apply_depolarizing(st, t, p, u)
        want_b, want = OB.apply_depolarizing(psi, t, p, u)
        assert got_b == want_b == b
        assert np.array_equal(_bits(st.cpu().numpy()), _bits(want)), (n, t, b)


@pytest.mark.parametrize("kind", ["damping", "dephasing"])
@pytest.mark.parametrize("n,t", [(1, 0), (3, 2), (9, 0), (9, 5), (12, 7), (12, 11)])
def test_damping_channels_match_oracle(bk, dev, kind, n, t):
    fn = bk.apply_amplitude_damping if kind == "damping" else bk.apply_dephasing
    g = 0.35
    for side in (0, 1):
        psi = rand_state(n, 70 + 10 * n + t + side)
        p64 = psi.astype(np.complex128)
        p1 = OB.branch_probability(p64, t, g)
        u = 0.5 * p1 if side else 0.5 * (1 + p1)
        st = to_dev(psi, dev)
        got_b = fn(st, t, g, u)
        want_b, want = OB.apply_damping_channel(kind, p64, t, g, u)
        assert got_b == want_b == side
        ok, err = close(want, st.cpu().numpy(), rtol=1e-6, atol=0)
        assert ok, (kind, n, t, side, err)
        assert abs(bk.norm2(st) - 1) < 1e-6
    with pytest.raises(RuntimeError):
        fn(st, t, 1.5, 0.1)


def test_channel_edge_probabilities(bk, dev):
    n = 5
    one = torch.zeros(1 << n, dtype=torch.complex64, device=dev)
    one[1 << 3] = 1                       # qubit 3 = 1
    assert bk.apply_amplitude_damping(one, 3, 1.0, 0.999) == 1      # gamma = 1 always decays
    got = one.cpu().numpy()
    assert got[0] == 1 and np.count_nonzero(got) == 1
    zero = bk.state_init(n, device=dev)
    assert bk.apply_amplitude_damping(zero, 3, 1.0, 0.0) == 0       # nothing to decay
    assert zero.cpu().numpy()[0] == 1
    assert bk.apply_depolarizing(zero, 0, 0.0, 0.999) == 0           # p = 0: identity
