"""FP64 CPU oracle for Pauli-string expectations, Pauli-string application
and the stochastic noise channels — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only to check; no product path uses it.

Conventions (qubit q = bit q of the amplitude index):
  * a Pauli string is (x, z): qubit q carries X if only x bit q is set, Z if
    only z bit q, Y if both; P = i^popcount(x & z) X^x Z^z, so that
    (P psi)_i = i^y (-1)^popcount((i ^ x) & z) psi_{i ^ x}, y = popcount(x & z);
  * expectation_z / expectation_pauli keep the reference's arguments
    (bwk/csrc/expectation_sm120.cu:180-239): a Z string, or a Z string times
    one X (op 1) or Y (op 2) on xy_qubit, where a z bit on xy_qubit is
    ignored (the reference takes the sign from the index with that bit clear,
    expectation_sm120.cu:127-129);
  * depolarizing: identity when u < 1 - p, else X, Y, Z on the three
    intervals of width p/3 that follow (bwk/csrc/noise_sm120.cu:250-281);
  * amplitude damping and dephasing: Kraus pairs E0 = diag(1, sqrt(1 - g))
    with E1 = [[0, sqrt(g)], [0, 0]] (damping, noise_sm120.cu:105-106) or
    E1 = diag(0, sqrt(g)) (dephasing, noise_sm120.cu:202-203); branch 1 when
    u < g * P(target = 1) (noise_sm120.cu:314, 358), the chosen branch
    renormalised to unit norm.  The reference's decay kernel also adds the
    renormalised |0> amplitude to the decayed one (noise_sm120.cu:113,
    193-194), which is not E1 psi and leaves the state off unit norm; the
    oracle follows the E1 the reference defines (noise_sm120.cu:106, 191).

How it is pinned (tests/test_oracle.py): Pauli application and expectations
against explicit numpy.kron products of the 2x2 Pauli matrices (every
string on 3 qubits); closed forms on |+>, |+i>, Bell states; Kraus
completeness; and the trajectory identity sum_b p_b |psi_b><psi_b| =
sum_k E_k rho E_k^+ for every channel, with p_b the width of branch b's
interval of u.
"""
import numpy as np

I2 = np.eye(2, dtype=np.complex128)
PX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
PY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
PZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)


def _popcount_parity(a):
    """popcount(a) & 1 elementwise for a uint64 array."""
    a = a.copy()
    p = np.zeros_like(a)
    while np.any(a):
        p ^= a & np.uint64(1)
        a >>= np.uint64(1)
    return p.astype(np.int64)


def pauli_dense(n, x, z):
    """2^n x 2^n matrix of the string (x, z) as a kron product (qubit 0 is
    the rightmost factor)."""
    m = np.ones((1, 1), dtype=np.complex128)
    for q in reversed(range(n)):
        xb, zb = (x >> q) & 1, (z >> q) & 1
        f = PY if (xb and zb) else PX if xb else PZ if zb else I2
        m = np.kron(m, f)
    return m


def apply_pauli_string(psi, x, z):
    """P psi with the components moved and negated only (exact in any
    precision; the dtype of psi is kept)."""
    psi = np.asarray(psi)
    i = np.arange(psi.size, dtype=np.uint64)
    src = psi[(i ^ np.uint64(x)).astype(np.int64)]
    k = (bin(x & z).count("1") + 2 * _popcount_parity((i ^ np.uint64(x)) & np.uint64(z))) % 4
    re, im = src.real.copy(), src.imag.copy()
    out_re = np.where(k == 0, re, np.where(k == 1, -im, np.where(k == 2, -re, im)))
    out_im = np.where(k == 0, im, np.where(k == 1, re, np.where(k == 2, -im, -re)))
    out = np.empty_like(psi)
    out.real, out.imag = out_re, out_im
    return out


def expectation_pauli_string(psi, x, z):
    """<psi|P|psi> (real for a Hermitian P), FP64, unnormalised."""
    psi = np.asarray(psi, dtype=np.complex128)
    return float(np.vdot(psi, apply_pauli_string(psi, x, z)).real)


def expectation_z(psi, z_mask):
    return expectation_pauli_string(psi, 0, int(z_mask))


def expectation_pauli(psi, z_mask, xy_qubit, op):
    """The reference's argument form (op 0 = I/Z, 1 = X, 2 = Y)."""
    if op not in (0, 1, 2):
        raise ValueError("op must be 0, 1 or 2")
    if op == 0:
        return expectation_z(psi, z_mask)
    bit = 1 << int(xy_qubit)
    z = (int(z_mask) & ~bit) | (bit if op == 2 else 0)
    return expectation_pauli_string(psi, bit, z)


# ------------------------------------------------------------------ noise
def depolarizing_branch(p, u):
    """0 = identity, 1/2/3 = X/Y/Z."""
    if u < 1.0 - p:
        return 0
    r, third = u - (1.0 - p), p / 3.0
    return 1 if r < third else (2 if r < 2.0 * third else 3)


def apply_depolarizing(psi, t, p, u):
    """(branch, new psi); the Pauli is applied exactly."""
    b = depolarizing_branch(p, u)
    if b == 0:
        return 0, np.array(psi, copy=True)
    bit = 1 << t
    x = 0 if b == 3 else bit
    z = 0 if b == 1 else bit
    return b, apply_pauli_string(psi, x, z)


def kraus(kind, g):
    """(E0, E1) of amplitude damping (kind 'damping') or dephasing."""
    e0 = np.diag([1.0, np.sqrt(1.0 - g)]).astype(np.complex128)
    if kind == "damping":
        e1 = np.array([[0, np.sqrt(g)], [0, 0]], dtype=np.complex128)
    elif kind == "dephasing":
        e1 = np.diag([0.0, np.sqrt(g)]).astype(np.complex128)
    else:
        raise ValueError(kind)
    return e0, e1


def _apply_2x2(psi, m, t):
    psi = np.asarray(psi, dtype=np.complex128)
    n = psi.size.bit_length() - 1
    v = psi.reshape((1 << (n - t - 1), 2, 1 << t))
    return np.einsum("ab,xbz->xaz", m, v).reshape(-1)


def branch_probability(psi, t, g):
    """g * P(qubit t = 1) of the normalised state."""
    psi = np.asarray(psi, dtype=np.complex128)
    p = np.abs(psi) ** 2
    idx = np.arange(psi.size)
    return g * float(p[(idx >> t) & 1 == 1].sum()) / float(p.sum())


def apply_damping_channel(kind, psi, t, g, u):
    """(branch, renormalised E_branch psi)."""
    e0, e1 = kraus(kind, g)
    b = 1 if u < branch_probability(psi, t, g) else 0
    out = _apply_2x2(psi, e1 if b else e0, t)
    nrm = np.linalg.norm(out)
    return b, (out / nrm if nrm > 0 else out)


def apply_amplitude_damping(psi, t, gamma, u):
    return apply_damping_channel("damping", psi, t, gamma, u)


def apply_dephasing(psi, t, lam, u):
    return apply_damping_channel("dephasing", psi, t, lam, u)


def channel_on_density(kraus_ops, rho, t, n):
    """sum_k E_k rho E_k^+ with E_k acting on qubit t of n."""
    out = np.zeros_like(rho)
    for e in kraus_ops:
        full = np.ones((1, 1), dtype=np.complex128)
        for q in reversed(range(n)):
            full = np.kron(full, e if q == t else I2)
        out += full @ rho @ full.conj().T
    return out
