"""Device-resident CSR Hamiltonians for Chebyshev evolution
(blackwell_backend/hamiltonian.py:20-79).

The reference converts a qiskit SparsePauliOp with
spo.to_matrix(sparse=True).tocsr().  This module builds the same operator
without qiskit, straight from the Pauli strings: a term with label L
(rightmost character = qubit 0) and coefficient c has exactly one nonzero
per row, at column i ^ xmask (xmask = X or Y positions), with value
    c * (-i)^{#Y} * (-1)^{popcount(i & (ymask | zmask))}
(each Y contributes <i_q|Y|i_q^1> = -i (-1)^{i_q}, each Z (-1)^{i_q}).
Terms sharing an xmask are summed, exact zeros dropped, and each row holds
its entries in ascending xmask order.  The spectral bound is sum |c_k|
(hamiltonian.py:66).
"""
import numpy as np
import torch


class SparsePauliOp:
    """Sum of Pauli strings with complex coefficients (Qiskit label order).

    Accepts the subset of qiskit.quantum_info.SparsePauliOp used here:
    from_list([(label, coeff), ...]), .coeffs, .num_qubits, .paulis labels."""

    def __init__(self, labels, coeffs):
        labels = [str(l).upper() for l in labels]
        if not labels:
            raise ValueError("empty Pauli sum")
        n = len(labels[0])
        for l in labels:
            if len(l) != n or set(l) - set("IXYZ"):
                raise ValueError(f"bad Pauli label {l!r}")
        self.labels = labels
        self.coeffs = np.asarray(coeffs, dtype=np.complex128)
        self.num_qubits = n

    @classmethod
    def from_list(cls, pairs):
        return cls([p for p, _ in pairs], [c for _, c in pairs])

    def to_labels(self):
        return list(self.labels)


def _labels_and_coeffs(spo):
    """Labels/coeffs of this module's SparsePauliOp or of qiskit's."""
    if hasattr(spo, "labels"):
        return spo.labels, np.asarray(spo.coeffs, dtype=np.complex128), spo.num_qubits
    labels = spo.paulis.to_labels()
    return labels, np.asarray(spo.coeffs, dtype=np.complex128), spo.num_qubits


def _parity(x):
    """popcount(x) & 1 for an int64 array."""
    x = x.copy()
    s = 32
    while s:
        x ^= x >> s
        s >>= 1
    return x & 1


def pauli_to_csr_numpy(labels, coeffs, n):
    """(row_ptr int32 [N+1], col_idx int32 [nnz], values complex128 [nnz])."""
    groups = {}
    for lab, c in zip(labels, coeffs):
        xm = ym = zm = 0
        for pos, ch in enumerate(reversed(lab)):
            if ch == "X":
                xm |= 1 << pos
            elif ch == "Y":
                xm |= 1 << pos
                ym |= 1 << pos
            elif ch == "Z":
                zm |= 1 << pos
        groups.setdefault(xm, []).append((c * (-1j) ** bin(ym).count("1"), ym | zm))
    N = 1 << n
    rows = np.arange(N, dtype=np.int64)
    masks = sorted(groups)
    vals = np.zeros((N, len(masks)), dtype=np.complex128)
    cols = np.zeros((N, len(masks)), dtype=np.int64)
    for j, xm in enumerate(masks):
        v = np.zeros(N, dtype=np.complex128)
        for coef, sm in groups[xm]:
            sign = 1 - 2 * _parity(rows & sm)
            v += coef * sign
        vals[:, j] = v
        cols[:, j] = rows ^ xm
    keep = vals != 0
    counts = keep.sum(axis=1)
    row_ptr = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    return (row_ptr.astype(np.int32), cols[keep].astype(np.int32), vals[keep])


class HamiltonianCSR:
    """CSR Hamiltonian on a device: row_ptr/col_idx int32, values complex64."""

    def __init__(self, row_ptr, col_idx, values, n_qubits, spectral_bound):
        self.row_ptr = row_ptr
        self.col_idx = col_idx
        self.values = values
        self.n_qubits = int(n_qubits)
        self.n_rows = 1 << self.n_qubits
        self.nnz = int(col_idx.shape[0])
        self.spectral_bound = float(spectral_bound)

    @staticmethod
    def from_sparse_pauli_op(spo, device="cuda"):
        labels, coeffs, n = _labels_and_coeffs(spo)
        rp, ci, v = pauli_to_csr_numpy(labels, coeffs, n)
        return HamiltonianCSR(torch.from_numpy(rp).to(device), torch.from_numpy(ci).to(device),
                              torch.from_numpy(v.astype(np.complex64)).to(device), n,
                              float(np.sum(np.abs(coeffs))))

    def memory_bytes(self):
        return self.row_ptr.nbytes + self.col_idx.nbytes + self.values.nbytes

    def __repr__(self):
        return (f"HamiltonianCSR(n_qubits={self.n_qubits}, dim={self.n_rows}, nnz={self.nnz}, "
                f"spectral_bound={self.spectral_bound:.4f}, memory={self.memory_bytes()/1024:.1f} KB)")
