"""FP64 CPU oracle for the statevector path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only to check or time; no product path uses it.

A complex128 statevector engine restating the reference's semantics:
  * layout: amplitude index i has bit q = qubit q (qubit 0 = LSB)
    (bwk/csrc/apply_gate_sm120.cu:11-14);
  * gate matrices: blackwell_backend/simulator.py:120-242 (Qiskit definitions);
  * 2Q kernel index r = 2*bit(max(q0,q1)) + bit(min(q0,q1))
    (bwk/csrc/apply_gate2q_sm120.cu:11-20, 245-246);
  * multi-controlled gates: insert-zeros over sorted controls+targets,
    i0 = base | control_mask, i1 = i0 | 2^t (bwk/csrc/apply_mcgate_sm120.cu:64-91,
    apply_mc2qgate_sm120.cu:78-110);
  * diagonals: apply_diagonal_sm120.cu:57, 76-79, 101;
  * measurement rule: outcome 0 iff u < p0 (measure_sm120.cu:88-106), here
    with p0 normalised by the FP64 norm;
  * sampling: smallest i with cdf(i) > u * total (the reference's rule at
    sample_sm120.cu:61-68 is smallest i with cdf[i] >= u on an FP32 CDF).

How each function is pinned (tests/test_oracle.py): closed forms (GHZ,
uniform superposition, Bell, H^2 = I, HZH = X, S^2 = Z, T^8 = I,
CX = H CZ H, SWAP = 3 CX, Rx(2pi) = -I, N x Rz(t/N) = Rz(t), QFT|x> = DFT
column), brute-force index enumeration for n <= 8, and library routines
(numpy.kron products of 2x2 matrices, scipy expm).
"""
import cmath
import math

import numpy as np

# ------------------------------------------------------------------ gates
I2 = np.eye(2, dtype=np.complex128)
X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)
H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2)


def gate_1q(name, params=()):
    """2x2 matrix of a named 1Q gate (blackwell_backend/simulator.py:120-177)."""
    p = [float(x) for x in params]
    if name == "h":
        return H.copy()
    if name == "x":
        return X.copy()
    if name == "y":
        return Y.copy()
    if name == "z":
        return Z.copy()
    if name == "s":
        return np.diag([1, 1j]).astype(np.complex128)
    if name == "sdg":
        return np.diag([1, -1j]).astype(np.complex128)
    if name == "t":
        return np.diag([1, cmath.exp(1j * math.pi / 4)])
    if name == "tdg":
        return np.diag([1, cmath.exp(-1j * math.pi / 4)])
    if name == "id":
        return I2.copy()
    if name == "sx":
        return np.array([[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]])
    if name == "sxdg":
        return np.array([[0.5 - 0.5j, 0.5 + 0.5j], [0.5 + 0.5j, 0.5 - 0.5j]])
    if name == "rx":
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        return np.array([[c, -1j * s], [-1j * s, c]])
    if name == "ry":
        c, s = math.cos(p[0] / 2), math.sin(p[0] / 2)
        return np.array([[c, -s], [s, c]], dtype=np.complex128)
    if name == "rz":
        return np.diag([cmath.exp(-1j * p[0] / 2), cmath.exp(1j * p[0] / 2)])
    if name in ("p", "u1"):
        return np.diag([1, cmath.exp(1j * p[0])])
    if name == "u2":
        phi, lam = p
        r = 1 / math.sqrt(2)
        return np.array([[r, -cmath.exp(1j * lam) * r],
                         [cmath.exp(1j * phi) * r, cmath.exp(1j * (phi + lam)) * r]])
    if name in ("u3", "u"):
        th, phi, lam = p
        c, s = math.cos(th / 2), math.sin(th / 2)
        return np.array([[c, -cmath.exp(1j * lam) * s],
                         [cmath.exp(1j * phi) * s, cmath.exp(1j * (phi + lam)) * c]])
    raise ValueError(f"Unknown 1Q gate: {name}")


def gate_2q_textbook(name, params=()):
    """4x4 with the FIRST listed qubit as the row MSB (simulator.py:179-210)."""
    p = [float(x) for x in params]
    g = np.eye(4, dtype=np.complex128)
    if name == "cx":
        g[2:, 2:] = X
    elif name == "cz":
        g[3, 3] = -1
    elif name == "swap":
        g = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]
    elif name == "cp":
        g[3, 3] = cmath.exp(1j * p[0])
    elif name == "crz":
        g[2, 2] = cmath.exp(-1j * p[0] / 2)
        g[3, 3] = cmath.exp(1j * p[0] / 2)
    elif name == "cu":
        th, phi, lam, gam = p
        c, s = math.cos(th / 2), math.sin(th / 2)
        e = cmath.exp(1j * gam)
        g[2, 2] = e * c
        g[2, 3] = -e * cmath.exp(1j * lam) * s
        g[3, 2] = e * cmath.exp(1j * phi) * s
        g[3, 3] = e * cmath.exp(1j * (phi + lam)) * c
    else:
        raise ValueError(f"Unknown 2Q gate: {name}")
    return g


def swap_bits_4(g):
    """Exchange which qubit is row bit 0 and bit 1 (simulator.py:212-216)."""
    perm = [0, 2, 1, 3]
    return g[np.ix_(perm, perm)]


# ------------------------------------------------------------------ engine
def zero_state(n):
    psi = np.zeros(1 << n, dtype=np.complex128)
    psi[0] = 1.0
    return psi


def apply_matrix(psi, u, qubits):
    """Apply a 2^k x 2^k matrix; row/column bit j <-> qubits[j]."""
    n = int(psi.size).bit_length() - 1
    k = len(qubits)
    t = psi.reshape([2] * n)
    axes = [n - 1 - q for q in qubits]
    ut = np.asarray(u, dtype=np.complex128).reshape([2] * (2 * k))
    in_axes = [k + (k - 1 - j) for j in range(k)]
    res = np.tensordot(ut, t, axes=(in_axes, axes))
    res = np.moveaxis(res, [k - 1 - j for j in range(k)], axes)
    return np.ascontiguousarray(res).reshape(-1)


def apply_1q(psi, g, t):
    return apply_matrix(psi, g, [t])


def apply_2q_kernel(psi, g4, q0, q1):
    """The 2Q kernel convention: row r = 2*bit(max) + bit(min)."""
    lo, hi = min(q0, q1), max(q0, q1)
    return apply_matrix(psi, g4, [lo, hi])


def controlled(u, n_controls):
    """Matrix of u controlled by n_controls qubits placed ABOVE u's bits."""
    d = u.shape[0]
    full = np.eye(d << n_controls, dtype=np.complex128)
    full[-d:, -d:] = u
    return full


def apply_mc(psi, g, controls, targets):
    """Controlled gate: g acts on targets (row bit j <-> targets[j]) where all
    control bits are 1."""
    return apply_matrix(psi, controlled(np.asarray(g), len(controls)), list(targets) + list(controls))


def apply_mc2q_kernel(psi, g4, controls, t0, t1):
    lo, hi = min(t0, t1), max(t0, t1)
    return apply_mc(psi, g4, controls, [lo, hi])


def apply_diag(psi, d, t):
    n = psi.size
    idx = np.arange(n)
    return psi * np.asarray(d, dtype=np.complex128)[(idx >> t) & 1]


def apply_diag_phase(psi, ph, t):
    idx = np.arange(psi.size)
    out = psi.copy()
    sel = ((idx >> t) & 1) == 1
    out[sel] *= ph
    return out


def apply_diag_2q(psi, d, q0, q1):
    lo, hi = min(q0, q1), max(q0, q1)
    idx = np.arange(psi.size)
    return psi * np.asarray(d, dtype=np.complex128)[(((idx >> hi) & 1) << 1) | ((idx >> lo) & 1)]


def mcx_perm(n, controls, target):
    """Index permutation of a multi-controlled X: out[i] = in[perm[i]]."""
    idx = np.arange(1 << n)
    cm = 0
    for c in controls:
        cm |= 1 << c
    hit = (idx & cm) == cm
    return np.where(hit, idx ^ (1 << target), idx)


def mcswap_perm(n, controls, q0, q1):
    idx = np.arange(1 << n)
    cm = 0
    for c in controls:
        cm |= 1 << c
    b0 = (idx >> q0) & 1
    b1 = (idx >> q1) & 1
    hit = ((idx & cm) == cm) & (b0 != b1)
    return np.where(hit, idx ^ ((1 << q0) | (1 << q1)), idx)


# ------------------------------------------------------------------ circuits
def run_instructions(n, instrs, psi=None):
    """Simulate [(name, qubits, params), ...] with Qiskit semantics in FP64.

    2Q named gates take (first, second) with the first listed qubit as the
    textbook MSB (control for cx/cz/cp/crz/cu); ccx/mcx/c3x/c4x: controls then
    target; cswap: control then the two swapped qubits."""
    psi = zero_state(n) if psi is None else np.asarray(psi, dtype=np.complex128).copy()
    for name, qs, params in instrs:
        name = name.lower()
        if name in ("barrier", "measure"):
            continue
        if len(qs) == 1 and name not in ("mcx",):
            psi = apply_1q(psi, gate_1q(name, params), qs[0])
        elif name in ("cx", "cz", "swap", "cp", "crz", "cu"):
            psi = apply_matrix(psi, gate_2q_textbook(name, params), [qs[1], qs[0]])
        elif name in ("ccx", "toffoli", "mcx", "c3x", "c4x"):
            psi = psi[mcx_perm(n, qs[:-1], qs[-1])]
        elif name == "cswap":
            psi = psi[mcswap_perm(n, [qs[0]], qs[1], qs[2])]
        elif name == "unitary":
            psi = apply_matrix(psi, params[0], list(qs))
        else:
            raise ValueError(f"Unsupported instruction: {name}")
    return psi


# ------------------------------------------------------------------ readout
def probs(psi):
    return (psi.real ** 2 + psi.imag ** 2)


def measure(psi, t, u):
    """(outcome, collapsed state): 0 iff u * total < p0."""
    p = probs(psi)
    idx = np.arange(psi.size)
    b = (idx >> t) & 1
    tot = p.sum()
    p0 = p[b == 0].sum()
    o = 0 if u * tot < p0 else 1
    po = p0 if o == 0 else tot - p0
    sc = 1.0 / math.sqrt(po) if po > 0 else 0.0
    out = np.where(b == o, psi * sc, 0)
    return o, out


def sample(psi, uniforms):
    """Smallest i with cdf(i) > u * total, FP64 CDF."""
    cdf = np.cumsum(probs(psi))
    x = np.asarray(uniforms, dtype=np.float64) * cdf[-1]
    return np.minimum(np.searchsorted(cdf, x, side="right"), psi.size - 1)


def splitmix_uniforms(seed, count):
    """u_k = (SplitMix64(seed + (k+1) * 0x9E3779B97F4A7C15) >> 11) * 2^-53."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def fidelity(a, b):
    """|<a|b>|^2 of the normalised complex128 up-casts."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    a = a / np.linalg.norm(a)
    b = b / np.linalg.norm(b)
    return float(abs(np.vdot(a, b)) ** 2)


def max_abs_diff_phase_aligned(ref, got):
    """max |ref - e^{i phi} got| with phi aligning the global phase."""
    ref = np.asarray(ref, dtype=np.complex128)
    got = np.asarray(got, dtype=np.complex128)
    ov = np.vdot(got, ref)
    ph = ov / abs(ov) if abs(ov) > 0 else 1.0
    return float(np.max(np.abs(ref - ph * got)))


def chi_square_p(counts_idx, p, n_shots, min_expected=5.0):
    """Pearson chi-square p-value of sampled indices against probabilities p,
    merging outcomes with expected count < min_expected into one bin."""
    from scipy.stats import chi2
    obs = np.bincount(np.asarray(counts_idx, dtype=np.int64), minlength=p.size).astype(np.float64)
    exp = p / p.sum() * n_shots
    big = exp >= min_expected
    o = list(obs[big])
    e = list(exp[big])
    if (~big).any():
        o.append(obs[~big].sum())
        e.append(exp[~big].sum())
    o = np.array(o)
    e = np.array(e)
    keep = e > 0
    stat = float((((o - e) ** 2)[keep] / e[keep]).sum())
    dof = max(int(keep.sum()) - 1, 1)
    return float(chi2.sf(stat, dof)), stat, dof
