"""FP64 CPU oracle for the SparsePauliOp -> CSR -> Chebyshev path —
TEST INFRASTRUCTURE ONLY (tests/, smoke() and bench.py's cpu_baseline leg).

* pauli_dense: brute-force Kronecker products of 2x2 Paulis in Qiskit label
  order (rightmost character = qubit 0), the definition behind
  SparsePauliOp.to_matrix used at blackwell_backend/hamiltonian.py:58.
* pauli_csr_grouped: the CSR layout the device builder writes (terms
  grouped by X mask, each group summed per row in FP64 in term order, exact
  zeros dropped, groups in ascending X-mask order within a row), computed
  with numpy; pinned against pauli_dense.  The reference gets its CSR from
  Qiskit's to_matrix(sparse=True) (hamiltonian.py:58), whose exact layout
  (explicit zeros, column order) only an import could pin; the operator it
  represents is what pauli_dense fixes.
* chebyshev_coefficients: restates blackwell_backend/chebyshev.py:26-78
  (c_k = (2 - delta_k0) (-i)^k J_k(alpha), K_est = alpha + 2 ln(1/eps) + 20,
  tail trimmed while the cumulative tail is below eps, K >= 2, conjugated for
  t < 0), with J_k from scipy.special.jv.
* chebyshev_evolve_fp64: the three-term recurrence of chebyshev.py:111-156
  in complex128.
* expm_reference: scipy.sparse.linalg.expm_multiply, the independent check.
* pauli_xz_groups / matvec_c / chebyshev_evolve_c: the same recurrence in C
  with OpenMP (oracle/cheb_fp64.c), with the matrix elements computed from
  the Pauli strings on the fly; fast enough for the 22-qubit chain with
  several right-hand sides.  Pinned against pauli_dense (matvec) and
  expm_multiply (evolution) in tests/test_oracle.py.
"""
import numpy as np
import scipy.sparse as sp
from scipy.special import jv

_P = {
    "I": np.eye(2, dtype=np.complex128),
    "X": np.array([[0, 1], [1, 0]], dtype=np.complex128),
    "Y": np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    "Z": np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def pauli_dense(labels, coeffs):
    """Dense matrix of sum_k coeffs[k] * P(labels[k])."""
    n = len(labels[0])
    m = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    for lab, c in zip(labels, coeffs):
        t = np.array([[1.0 + 0j]])
        for ch in lab:
            t = np.kron(t, _P[ch])
        m += c * t
    return m


def pauli_sparse(labels, coeffs):
    """Same operator as a scipy CSR matrix (built term by term from sparse
    Kronecker products; exact zeros removed)."""
    n = len(labels[0])
    m = sp.csr_matrix((1 << n, 1 << n), dtype=np.complex128)
    for lab, c in zip(labels, coeffs):
        t = sp.csr_matrix(np.array([[1.0 + 0j]]))
        for ch in lab:
            t = sp.kron(t, sp.csr_matrix(_P[ch]), format="csr")
        m = m + c * t
    m.eliminate_zeros()
    return m.tocsr()


def _parity(x):
    """popcount(x) & 1 for a uint64 array."""
    x = x.copy()
    s = 32
    while s:
        x ^= x >> np.uint64(s)
        s >>= 1
    return (x & np.uint64(1)).astype(np.int64)


def pauli_csr_grouped(labels, coeffs, n):
    """(row_ptr int32 [2^n+1], col_idx int32 [nnz], values complex128 [nnz])."""
    phase = (1 + 0j, -1j, -1 + 0j, 1j)
    groups = {}
    for lab, c in zip(labels, coeffs):
        xm = ym = zm = 0
        for pos, ch in enumerate(reversed(lab)):
            if ch in "XY":
                xm |= 1 << pos
            if ch == "Y":
                ym |= 1 << pos
            if ch == "Z":
                zm |= 1 << pos
        groups.setdefault(xm, []).append((complex(c) * phase[bin(ym).count("1") % 4], ym | zm))
    N = 1 << n
    rows = np.arange(N, dtype=np.uint64)
    masks = sorted(groups)
    vals = np.zeros((N, len(masks)), dtype=np.complex128)
    cols = np.zeros((N, len(masks)), dtype=np.int64)
    for j, xm in enumerate(masks):
        v = np.zeros(N, dtype=np.complex128)
        for coef, sm in groups[xm]:
            v += coef * (1 - 2 * _parity(rows & np.uint64(sm)))
        vals[:, j] = v
        cols[:, j] = (rows ^ np.uint64(xm)).astype(np.int64)
    keep = vals != 0
    row_ptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(keep.sum(axis=1), out=row_ptr[1:])
    return row_ptr.astype(np.int32), cols[keep].astype(np.int32), vals[keep]


def chebyshev_coefficients(t, spectral_bound, precision=1e-10):
    alpha = spectral_bound * abs(t)
    k_est = max(int(alpha + 2.0 * np.log(1.0 / precision) + 20), 10)
    ks = np.arange(k_est + 1)
    coeffs = np.where(ks > 0, 2.0, 1.0) * ((-1j) ** ks) * jv(ks, alpha)
    cumtail = np.cumsum(np.abs(coeffs)[::-1])[::-1]
    K = k_est
    for k in range(k_est, 0, -1):
        if cumtail[k] < precision:
            K = k
        else:
            break
    K = max(K, 2)
    coeffs = coeffs[:K + 1]
    if t < 0:
        coeffs = np.conj(coeffs)
    tail = float(cumtail[min(K, len(cumtail) - 1)])
    return coeffs, K, tail


def chebyshev_evolve_fp64(h, psi, t, spectral_bound, precision=1e-10):
    """exp(-iHt) psi by the Chebyshev recurrence in complex128 (H sparse or dense)."""
    coeffs, K, _ = chebyshev_coefficients(t, spectral_bound, precision)
    scale = 1.0 / spectral_bound
    t0 = np.asarray(psi, dtype=np.complex128)
    t1 = scale * (h @ t0)
    acc = coeffs[0] * t0 + coeffs[1] * t1
    prev, cur = t0, t1
    for k in range(2, K + 1):
        nxt = 2.0 * scale * (h @ cur) - prev
        acc = acc + coeffs[k] * nxt
        prev, cur = cur, nxt
    return acc


def expm_reference(h, psi, t):
    from scipy.sparse.linalg import expm_multiply
    return expm_multiply(-1j * t * sp.csr_matrix(h), np.asarray(psi, dtype=np.complex128))


def pauli_xz_groups(labels, coeffs):
    """Terms as (x mask, z mask, coefficient) grouped by x mask, in Qiskit
    label order (rightmost character = qubit 0; Y sets both masks).
    Returns (xm uint64 [ng], toff int32 [ng+1], zm uint64 [nt],
    coef float64 [nt, 2])."""
    groups = {}
    for lab, c in zip(labels, coeffs):
        x = z = 0
        for pos, ch in enumerate(reversed(lab)):
            if ch in "XY":
                x |= 1 << pos
            if ch in "ZY":
                z |= 1 << pos
        groups.setdefault(x, []).append((z, complex(c)))
    xs = sorted(groups)
    toff = [0]
    zm, cf = [], []
    for x in xs:
        for z, c in groups[x]:
            zm.append(z)
            cf.append((c.real, c.imag))
        toff.append(len(zm))
    return (np.array(xs, dtype=np.uint64), np.array(toff, dtype=np.int32),
            np.array(zm, dtype=np.uint64), np.array(cf, dtype=np.float64).reshape(-1, 2))


def _rhs_block(psi_rows):
    """(rows of vectors [R, N]) -> RHS-contiguous complex128 [N, R]."""
    v = np.atleast_2d(np.asarray(psi_rows, dtype=np.complex128))
    return np.ascontiguousarray(v.T), v.shape[0]


def matvec_c(labels, coeffs, n, psi_rows, scale=1.0, threads=0):
    """scale * H v for each row v of psi_rows, by oracle/cheb_fp64.c."""
    from oracle import build_ref
    lib = build_ref.load()
    lib.ph_set_threads(int(threads))
    xm, toff, zm, cf = pauli_xz_groups(labels, coeffs)
    v, r = _rhs_block(psi_rows)
    y = np.empty_like(v)
    lib.ph_matvec(n, len(xm), xm.ctypes.data, toff.ctypes.data, zm.ctypes.data, cf.ctypes.data, r,
                  v.ctypes.data, y.ctypes.data, float(scale))
    return np.ascontiguousarray(y.T)


def chebyshev_evolve_c(labels, coeffs, n, psi_rows, t, spectral_bound, precision=1e-13, threads=0):
    """exp(-iHt) v for each row v of psi_rows: the Chebyshev recurrence in
    complex128 on host cores (oracle/cheb_fp64.c).  Returns (rows, K)."""
    from oracle import build_ref
    lib = build_ref.load()
    lib.ph_set_threads(int(threads))
    c, K, _ = chebyshev_coefficients(t, spectral_bound, precision)
    cc = np.ascontiguousarray(np.stack([c.real, c.imag], axis=1))
    xm, toff, zm, cf = pauli_xz_groups(labels, coeffs)
    v, r = _rhs_block(psi_rows)
    out = np.empty_like(v)
    rc = lib.ph_cheb_evolve(n, len(xm), xm.ctypes.data, toff.ctypes.data, zm.ctypes.data, cf.ctypes.data,
                            r, 1.0 / float(spectral_bound), int(K), cc.ctypes.data, v.ctypes.data,
                            out.ctypes.data)
    if rc != 0:
        raise MemoryError("chebyshev_evolve_c: work vectors could not be allocated")
    return np.ascontiguousarray(out.T), K
