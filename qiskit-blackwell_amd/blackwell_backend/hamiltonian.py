"""Device-resident CSR Hamiltonians for Chebyshev evolution
(blackwell_backend/hamiltonian.py:20-79).

The reference converts a qiskit SparsePauliOp with
spo.to_matrix(sparse=True).tocsr() on the CPU and copies the CSR to the
device.  Here the CSR is built on the device, straight from the Pauli
strings: a term with label L (rightmost character = qubit 0) and
coefficient c has exactly one nonzero per row, at column i ^ xmask
(xmask = X or Y positions), with value
    c * (-i)^{#Y} * (-1)^{popcount(i & (ymask | zmask))}
(each Y contributes <i_q|Y|i_q^1> = -i (-1)^{i_q}, each Z (-1)^{i_q}).
The host groups the terms by xmask (pauli_groups); the library's
qsim_pauli_csr_count / qsim_pauli_csr_fill sum each group per row in FP64,
drop exact zeros and write the rows with their entries in ascending xmask
order.  The spectral bound is sum |c_k| (hamiltonian.py:66).
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


# (-i)^k for k = 0..3, exact
_MINUS_I_POW = (1 + 0j, -1j, -1 + 0j, 1j)


def pauli_groups(labels, coeffs):
    """Terms grouped by X mask, in the layout the device CSR builder takes:
    (xm uint64 [G] ascending, toff int32 [G+1], sm uint64 [T] (Y|Z masks),
    coef complex128 [T] with (-i)^#Y folded in), terms of a group in input
    order."""
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
        groups.setdefault(xm, []).append((complex(c) * _MINUS_I_POW[bin(ym).count("1") % 4], ym | zm))
    masks = sorted(groups)
    toff = np.zeros(len(masks) + 1, dtype=np.int32)
    sm, coef = [], []
    for g, xm in enumerate(masks):
        for cf, m in groups[xm]:
            coef.append(cf)
            sm.append(m)
        toff[g + 1] = len(sm)
    return (np.array(masks, dtype=np.uint64), toff, np.array(sm, dtype=np.uint64),
            np.array(coef, dtype=np.complex128))


def constant_group_entries(toff, sm, coef):
    """Per group, the entry rounded to complex64 (as FP64 pairs) when no term
    has a Y|Z mask, so that it is the same in every row: the FP64 sum of the
    coefficients in term order, as the device evaluates it; NaN where the
    entry depends on the row."""
    out = np.full((len(toff) - 1, 2), np.nan)
    for g in range(len(toff) - 1):
        t0, t1 = int(toff[g]), int(toff[g + 1])
        if np.any(sm[t0:t1] != 0):
            continue
        re = im = 0.0
        for t in range(t0, t1):
            re += float(coef[t].real)
            im += float(coef[t].imag)
        out[g] = (float(np.float32(re)), float(np.float32(im)))
    return out


class PauliTables:
    """pauli_groups() on a device: n_groups and the tensors xm (uint64 as
    int64), toff (int32), sm (uint64 as int64), coef (FP64 pairs), the
    layout qsim_pauli_csr_count and qsim_cheb_pauli_step take, plus gconst
    (constant_group_entries) for the matrix-free Chebyshev term."""

    def __init__(self, labels, coeffs, device):
        xm, toff, sm, coef = pauli_groups(labels, coeffs)
        dev = torch.device(device)
        self.n_groups = len(xm)
        self.xm = torch.from_numpy(xm.view(np.int64) if xm.size else np.zeros(1, np.int64)).to(dev)
        self.toff = torch.from_numpy(toff).to(dev)
        self.sm = torch.from_numpy(sm.view(np.int64) if sm.size else np.zeros(1, np.int64)).to(dev)
        self.coef = torch.from_numpy(coef.view(np.float64) if coef.size else np.zeros(2)).to(dev)
        gc = constant_group_entries(toff, sm, coef)
        self.gconst = torch.from_numpy(gc.reshape(-1) if gc.size else np.zeros(2)).to(dev)

    def args(self):
        """(n_groups, xm, toff, sm, coef) pointers for qsim_pauli_csr_count / _fill."""
        from blackwell_kernels import _lib
        return (self.n_groups, _lib.ptr(self.xm), _lib.ptr(self.toff), _lib.ptr(self.sm), _lib.ptr(self.coef))

    def step_args(self):
        """args() plus gconst, for qsim_cheb_pauli_step."""
        from blackwell_kernels import _lib
        return self.args() + (_lib.ptr(self.gconst),)


def pauli_to_csr_device(labels, coeffs, n, device, tables=None):
    """(row_ptr int32 [2^n+1], col_idx int32 [nnz], values complex64 [nnz])
    as device tensors, built by the library from pauli_groups (or from
    `tables`, a PauliTables of the same terms)."""
    import ctypes
    from blackwell_kernels import _lib
    pt = tables if tables is not None else PauliTables(labels, coeffs, device)
    dev = torch.device(device)
    ws = torch.empty(int(_lib.load().qsim_pauli_csr_workspace_bytes(n)), dtype=torch.uint8, device=dev)
    s = _lib.stream_handle(dev)
    args = (n,) + pt.args() + (_lib.ptr(ws),)
    nnz = ctypes.c_int64(0)
    _lib.call("qsim_pauli_csr_count", s, *args, ctypes.byref(nnz))
    row_ptr = torch.empty((1 << n) + 1, dtype=torch.int32, device=dev)
    col_idx = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
    values = torch.empty(max(nnz.value, 1), dtype=torch.complex64, device=dev)
    _lib.call("qsim_pauli_csr_fill", s, *args, nnz.value, _lib.ptr(row_ptr), _lib.ptr(col_idx),
              _lib.ptr(values))
    return row_ptr, col_idx[:nnz.value], values[:nnz.value]


class HamiltonianCSR:
    """CSR Hamiltonian on a device: row_ptr/col_idx int32, values complex64."""

    def __init__(self, row_ptr, col_idx, values, n_qubits, spectral_bound, pauli=None):
        # pauli: the PauliTables the CSR was built from, when it came from a
        # Pauli sum (single-state Chebyshev terms then run matrix-free)
        self.pauli = pauli
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
        pt = PauliTables(labels, coeffs, device)
        rp, ci, v = pauli_to_csr_device(labels, coeffs, n, device, tables=pt)
        return HamiltonianCSR(rp, ci, v, n, float(np.sum(np.abs(coeffs))), pauli=pt)

    def memory_bytes(self):
        return self.row_ptr.nbytes + self.col_idx.nbytes + self.values.nbytes

    def split(self, split_bit=None):
        """The CSR split at bit s (default n // 2) for the two-pass 32-RHS
        Chebyshev term (qsim_csr_split_count / _fill): (split_bit, row_ptr_lo,
        col_idx_lo, values_lo, row_ptr_hi, col_idx_hi, values_hi) device
        tensors, the high part's rows in the rotated order.  Built on the
        device once and cached."""
        import ctypes
        from blackwell_kernels import _lib
        s_bit = int(split_bit) if split_bit is not None else self.n_qubits // 2
        cache = getattr(self, "_split", None)
        if cache is not None and cache[0] == s_bit:
            return cache
        dev = self.row_ptr.device
        N = self.n_rows
        ws = torch.empty(int(_lib.load().qsim_csr_split_workspace_bytes(N)), dtype=torch.uint8, device=dev)
        st = _lib.stream_handle(dev)
        nlo, nhi = ctypes.c_int64(0), ctypes.c_int64(0)
        _lib.call("qsim_csr_split_count", st, _lib.ptr(self.row_ptr), _lib.ptr(self.col_idx), N, s_bit,
                  _lib.ptr(ws), ctypes.byref(nlo), ctypes.byref(nhi))
        rp_lo = torch.empty(N + 1, dtype=torch.int32, device=dev)
        rp_hi = torch.empty(N + 1, dtype=torch.int32, device=dev)
        ci_lo = torch.empty(max(nlo.value, 1), dtype=torch.int32, device=dev)
        ci_hi = torch.empty(max(nhi.value, 1), dtype=torch.int32, device=dev)
        v_lo = torch.empty(max(nlo.value, 1), dtype=torch.complex64, device=dev)
        v_hi = torch.empty(max(nhi.value, 1), dtype=torch.complex64, device=dev)
        _lib.call("qsim_csr_split_fill", st, _lib.ptr(self.row_ptr), _lib.ptr(self.col_idx),
                  _lib.ptr(self.values), N, s_bit, _lib.ptr(ws), _lib.ptr(rp_lo), _lib.ptr(ci_lo),
                  _lib.ptr(v_lo), _lib.ptr(rp_hi), _lib.ptr(ci_hi), _lib.ptr(v_hi))
        self._split = (s_bit, rp_lo, ci_lo[:nlo.value], v_lo[:nlo.value], rp_hi, ci_hi[:nhi.value],
                       v_hi[:nhi.value])
        return self._split

    def __repr__(self):
        return (f"HamiltonianCSR(n_qubits={self.n_qubits}, dim={self.n_rows}, nnz={self.nnz}, "
                f"spectral_bound={self.spectral_bound:.4f}, memory={self.memory_bytes()/1024:.1f} KB)")


def pauli_masks(labels):
    """(x masks, z masks) of Pauli labels (rightmost character = qubit 0):
    X sets x, Z sets z, Y sets both."""
    xs, zs = [], []
    for lab in labels:
        xm = zm = 0
        for pos, ch in enumerate(reversed(lab)):
            if ch in "XY":
                xm |= 1 << pos
            if ch in "ZY":
                zm |= 1 << pos
        xs.append(xm)
        zs.append(zm)
    return xs, zs


def expectation(state, spo):
    """<psi|H|psi> = sum_k c_k <psi|P_k|psi> for a SparsePauliOp (this
    module's or qiskit's) on a device state, unnormalised.  Each <P_k> is an
    FP64 device reduction; terms sharing an X mask share one pass over the
    state.  Returns a float when every coefficient is real, else a complex."""
    from blackwell_kernels import expectation_paulis
    labels, coeffs, n = _labels_and_coeffs(spo)
    if int(state.numel()) != 1 << n:
        raise ValueError(f"state has {state.numel()} amplitudes, operator acts on {n} qubits")
    xs, zs = pauli_masks(labels)
    vals = expectation_paulis(state, xs, zs).cpu().numpy()
    total = complex(np.sum(coeffs * vals))
    return total.real if not np.any(coeffs.imag) else total
