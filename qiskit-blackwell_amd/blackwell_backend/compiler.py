"""Circuit compiler: instruction list -> ops -> execution plans.

Stage 1, parse (blackwell_backend/simulator.py:261-306): every named gate
becomes one op with its matrix in FP64 (gate definitions of
simulator.py:120-242).  Two-qubit controlled gates become controlled
one-qubit ops (cz, cp: controlled phase; crz: controlled diagonal; cu:
controlled 2x2; cx/ccx/mcx: controlled X; cswap: controlled SWAP), so the
permutation gates stay pure data movement (bit-exact) and diagonal ones
touch only the control=1 half.

Stage 2, 1Q-run fusion: consecutive one-qubit ops on the same qubit,
between two multi-qubit ops that touch it, are multiplied into one 2x2 in
FP64 on the host.  This covers the reference's phases 2-3.5 (diagonal
merge, same-axis rotation merge in FP64, diagonal absorption, same-qubit
matrix products: simulator.py:308-377) without their adjacency limit and
without complex64 device matmuls.  A fused product that is diagonal becomes
a diagonal op (phase-only when d0 = 1, the reference's rule at
simulator.py:388); the identity is dropped; an exact X stays a permutation.

Stage 3, planning:
  * plan_plain  - one kernel launch per op (the per-gate baseline);
  * plan_blocks - dense k-qubit blocks (k = 5 on FP32 MFMA) built by a
                  dependency-aware greedy scan, matrices multiplied in FP64;
  * plan_tiles  - LDS tile programs: ops packed into passes whose qubits
                  fit a 2^m-amplitude tile, one HBM round trip per pass.
"""
import cmath
import math

import numpy as np

_NATIVE_1Q = {"h", "x", "y", "z", "s", "sdg", "t", "tdg", "rx", "ry", "rz", "p", "u1", "u2",
              "u3", "u", "id", "sx", "sxdg"}
_DIAG_1Q = {"z", "s", "sdg", "t", "tdg", "rz", "p", "u1", "id"}
NATIVE_NAMES = _NATIVE_1Q | {"cx", "cz", "swap", "cp", "crz", "cu", "ccx", "toffoli", "mcx",
                             "c3x", "c4x", "cswap", "unitary", "measure", "reset", "barrier"}

_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)


def gate_1q(name, params=()):
    """2x2 FP64 matrix of a named 1Q gate (simulator.py:120-177)."""
    p = [float(v) for v in params]
    if name == "h":
        return np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2)
    if name == "x":
        return _X.copy()
    if name == "y":
        return np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
    if name == "z":
        return np.diag([1.0, -1.0]).astype(np.complex128)
    if name == "s":
        return np.diag([1, 1j]).astype(np.complex128)
    if name == "sdg":
        return np.diag([1, -1j]).astype(np.complex128)
    if name == "t":
        return np.diag([1, cmath.exp(1j * math.pi / 4)])
    if name == "tdg":
        return np.diag([1, cmath.exp(-1j * math.pi / 4)])
    if name == "id":
        return np.eye(2, dtype=np.complex128)
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


class Op:
    """One operation.  kinds:
      m1    dense 2x2 on qubits[0]                       mat (2,2)
      d1    diagonal (d0, d1) on qubits[0]               mat (2,)
      ph    phase on the bit=1 half of qubits[0]         mat (1,)
      m2    dense 4x4, row bit j <-> qubits[j]           mat (4,4)
      mk    dense 2^k block, row bit j <-> qubits[j]     mat (2^k,2^k)
      x     X on qubits[0]           (permutation)
      swap  SWAP of qubits[0], qubits[1] (permutation)
      measure / reset on qubits[0] (clbit for measure)
    Any unitary kind may carry controls (all must be 1)."""
    __slots__ = ("kind", "qubits", "controls", "mat", "clbit")

    def __init__(self, kind, qubits, mat=None, controls=(), clbit=None):
        self.kind = kind
        self.qubits = tuple(int(q) for q in qubits)
        self.controls = tuple(int(c) for c in controls)
        self.mat = mat
        self.clbit = clbit

    @property
    def all_qubits(self):
        return self.qubits + self.controls

    def __repr__(self):
        c = f" ctl={self.controls}" if self.controls else ""
        return f"Op({self.kind} {self.qubits}{c})"

    def target_matrix(self):
        """Matrix on self.qubits (row bit j <-> qubits[j]), controls excluded."""
        k = self.kind
        if k in ("m1", "m2", "mk"):
            return self.mat
        if k == "d1":
            return np.diag(self.mat)
        if k == "ph":
            return np.diag([1.0, self.mat[0]])
        if k == "x":
            return _X.copy()
        if k == "swap":
            return np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]
        raise ValueError(f"{k} has no matrix")

    def full_matrix(self):
        """Matrix on all_qubits (targets first, then controls)."""
        u = self.target_matrix()
        if not self.controls:
            return u
        d = u.shape[0]
        full = np.eye(d << len(self.controls), dtype=np.complex128)
        full[-d:, -d:] = u
        return full


def _textbook_2q_to_op(name, qs, params):
    """cx/cz/swap/cp/crz/cu with (first, second) qubits (simulator.py:182-208)."""
    c, t = qs
    p = [float(v) for v in params]
    if name == "cx":
        return Op("x", [t], controls=[c])
    if name == "cz":
        return Op("ph", [t], np.array([-1.0 + 0j]), controls=[c])
    if name == "swap":
        return Op("swap", [c, t])
    if name == "cp":
        return Op("ph", [t], np.array([cmath.exp(1j * p[0])]), controls=[c])
    if name == "crz":
        return Op("d1", [t], np.array([cmath.exp(-1j * p[0] / 2), cmath.exp(1j * p[0] / 2)]),
                  controls=[c])
    if name == "cu":
        th, phi, lam, gam = p
        cs, sn = math.cos(th / 2), math.sin(th / 2)
        e = cmath.exp(1j * gam)
        u = np.array([[e * cs, -e * cmath.exp(1j * lam) * sn],
                      [e * cmath.exp(1j * phi) * sn, e * cmath.exp(1j * (phi + lam)) * cs]])
        return Op("m1", [t], u, controls=[c])
    raise ValueError(f"Unsupported instruction: {name}")


def parse(instrs):
    """[(name, qubits, clbits, params)] -> [Op]."""
    ops = []
    for name, qs, cs, params in instrs:
        if name == "barrier":
            continue
        if name == "measure":
            ops.append(Op("measure", [qs[0]], clbit=cs[0] if cs else None))
        elif name == "reset":
            ops.append(Op("reset", [qs[0]]))
        elif name in _NATIVE_1Q:
            if name == "x":
                ops.append(Op("x", [qs[0]]))
            elif name in _DIAG_1Q:
                ops.append(Op("d1", [qs[0]], np.diagonal(gate_1q(name, params)).copy()))
            else:
                ops.append(Op("m1", [qs[0]], gate_1q(name, params)))
        elif name in ("cx", "cz", "swap", "cp", "crz", "cu"):
            ops.append(_textbook_2q_to_op(name, qs, params))
        elif name in ("ccx", "toffoli", "mcx", "c3x", "c4x"):
            ops.append(Op("x", [qs[-1]], controls=qs[:-1]))
        elif name == "cswap":
            ops.append(Op("swap", [qs[1], qs[2]], controls=[qs[0]]))
        elif name == "unitary":
            m = np.asarray(params[0], dtype=np.complex128)
            k = len(qs)
            ops.append(Op("m1" if k == 1 else ("m2" if k == 2 else "mk"), qs, m))
        else:
            raise ValueError(f"Unsupported instruction: {name}")
    return ops


# ------------------------------------------------------------------ stage 2
def _classify_1q(q, m, tol=1e-14):
    """Op for a fused 2x2 product, or None for the identity.  Off-diagonal
    or identity deviations below tol are FP64 rounding of products such as
    H.H and are dropped (far below the FP32 resolution of the kernels)."""
    off = abs(m[0, 1]) + abs(m[1, 0])
    if off <= tol:
        d0, d1 = m[0, 0], m[1, 1]
        if abs(d0 - 1) <= tol and abs(d1 - 1) <= tol:
            return None
        if abs(d0 - 1.0) < 1e-10:
            return Op("ph", [q], np.array([d1]))
        return Op("d1", [q], np.array([d0, d1]))
    if m[0, 0] == 0 and m[1, 1] == 0 and m[0, 1] == 1 and m[1, 0] == 1:
        return Op("x", [q])
    return Op("m1", [q], m)


def fuse_1q_runs(ops):
    """Multiply runs of uncontrolled 1Q ops per qubit in FP64 (stage 2)."""
    pending = {}      # qubit -> FP64 2x2 product of the run so far
    out = []

    def flush(q):
        m = pending.pop(q, None)
        if m is not None:
            op = _classify_1q(q, m)
            if op is not None:
                out.append(op)

    for op in ops:
        if op.kind in ("m1", "d1", "ph", "x") and not op.controls and len(op.qubits) == 1:
            q = op.qubits[0]
            u = op.target_matrix()
            pending[q] = u @ pending[q] if q in pending else u.copy()
            continue
        for q in op.all_qubits:
            flush(q)
        out.append(op)
    for q in sorted(pending):
        flush(q)
    return out


def compile_ops(instrs, fuse=True):
    ops = parse(instrs)
    return fuse_1q_runs(ops) if fuse else ops


# ------------------------------------------------------------------ stage 3
def kernel_2q_matrix(op):
    """4x4 in the 2Q kernel's convention (row bit 0 <-> min qubit)."""
    m = op.target_matrix()
    if op.qubits[0] < op.qubits[1]:
        return m
    perm = [0, 2, 1, 3]
    return m[np.ix_(perm, perm)]


def embed(u, qubits, block_qubits):
    """Matrix u on `qubits` (row bit j <-> qubits[j]) expanded onto
    block_qubits (row bit b <-> block_qubits[b])."""
    k = len(block_qubits)
    pos = [block_qubits.index(q) for q in qubits]
    kq = len(qubits)
    d = 1 << k
    rows = np.arange(d)
    sub = np.zeros(d, dtype=np.int64)
    for j, b in enumerate(pos):
        sub |= ((rows >> b) & 1) << j
    rest_mask = d - 1
    for b in pos:
        rest_mask &= ~(1 << b)
    full = np.zeros((d, d), dtype=np.complex128)
    same_rest = (rows[:, None] & rest_mask) == (rows[None, :] & rest_mask)
    full[same_rest] = np.asarray(u)[sub[:, None], sub[None, :]][same_rest]
    return full


class Block:
    """Dense block on `qubits` (ascending) holding the product of `ops`."""
    __slots__ = ("qubits", "ops", "mat")

    def __init__(self, qubits, ops, mat):
        self.qubits = qubits
        self.ops = ops
        self.mat = mat


def _mask(qs):
    m = 0
    for q in qs:
        m |= 1 << q
    return m


def plan_blocks(ops, n, k=5, effort=1):
    """Pack unitary ops into dense blocks of min(k, n) qubits with the
    library's host planner (qsim_plan_blocks_search, FP64): an op joins a
    block if its qubits lie in the block's qubit set and none of them was
    touched by an earlier op left out; effort 1 (the default) also searches
    the block qubit sets around each seed op, effort 0 is the greedy scan
    alone.  measure/reset split the
    list into independently planned segments.  Returns a list of Block | Op
    (ops wider than k pass through)."""
    out = []
    seg = []
    for op in ops:
        if op.kind in ("measure", "reset"):
            out.extend(_plan_segment(seg, n, k, effort))
            seg = []
            out.append(op)
        else:
            seg.append(op)
    out.extend(_plan_segment(seg, n, k, effort))
    return out


def plan_blocks_arrays(n, op_nq, op_qubits, op_mats, k=5, effort=1):
    """The native planner on flat arrays: op_nq int32 [m], op_qubits int32
    [m, 8] (-1 padded), op_mats complex128 (each op's 2^nq x 2^nq matrix,
    row bit j <-> op_qubits[j], concatenated); effort as plan_blocks (the
    serving lane plans gate payloads with the same default, so a payload and
    its circuit get the same blocks).  Returns (refs, block_qubits
    int32 [B, 8], block_mats complex128 [B, D, D]): refs[e] >= 0 is block
    refs[e], refs[e] < 0 is op -1 - refs[e] passed through.  Pure host code;
    the library call releases the GIL, so plans can be built on worker
    threads."""
    import ctypes
    from blackwell_kernels import _lib
    m = len(op_nq)
    nq = np.ascontiguousarray(op_nq, dtype=np.int32)
    qb = np.ascontiguousarray(op_qubits, dtype=np.int32).reshape(m, 8)
    mv = np.ascontiguousarray(op_mats, dtype=np.complex128).reshape(-1)
    if mv.size == 0:
        mv = np.zeros(1, dtype=np.complex128)
    D = 1 << min(k, n)
    refs = np.zeros(max(m, 1), dtype=np.int32)
    bqs = np.zeros((max(m, 1), 8), dtype=np.int32)
    bm = np.empty((max(m, 1), D, D), dtype=np.complex128)
    n_out = ctypes.c_int(0)
    vp = lambda a: ctypes.c_void_p(a.ctypes.data)
    _lib.call("qsim_plan_blocks_search", n, k, effort, m, vp(nq), vp(qb), vp(mv.view(np.float64)),
              vp(refs), vp(bqs), vp(bm), ctypes.byref(n_out))
    refs = refs[:n_out.value]
    nb = int((refs >= 0).sum())
    return refs, bqs[:nb], bm[:nb]


def _plan_segment(ops, n, k, effort=1):
    if not ops:
        return []
    m = len(ops)
    nq = np.zeros(m, dtype=np.int32)
    qb = np.full((m, 8), -1, dtype=np.int32)
    mats = []
    for i, op in enumerate(ops):
        qs = op.all_qubits
        nq[i] = len(qs)
        qb[i, :len(qs)] = qs
        mats.append(np.asarray(op.full_matrix(), dtype=np.complex128).reshape(-1))
    refs, bqs, bm = plan_blocks_arrays(n, nq, qb, np.concatenate(mats), k, effort)
    res = []
    for e in refs:
        if e >= 0:
            qs = [int(q) for q in bqs[e] if q >= 0]
            res.append(Block(qs, None, bm[e]))
        else:
            res.append(ops[-1 - int(e)])
    return res


def apply_to_block(mat, u, qubits, block_qubits):
    """u (row bit j <-> qubits[j]) applied on the left of the block matrix
    mat (row bit b <-> block_qubits[b]), by contracting the matching axes."""
    k = len(block_qubits)
    kq = len(qubits)
    d = 1 << k
    t = mat.reshape([2] * k + [d])            # axis a <-> block bit k-1-a
    axes = [k - 1 - block_qubits.index(q) for q in qubits]
    ut = np.asarray(u).reshape([2] * (2 * kq))
    res = np.tensordot(ut, t, axes=([kq + (kq - 1 - j) for j in range(kq)], axes))
    res = np.moveaxis(res, [kq - 1 - j for j in range(kq)], axes)
    return res.reshape(d, d)


class TilePass:
    """One LDS tile program: `mask` (state qubits of the tile) and its ops."""
    __slots__ = ("mask", "ops")

    def __init__(self, mask, ops):
        self.mask = mask
        self.ops = ops


def plan_tiles(ops, n, m=13, low=4):
    """Pack ops into tile passes of at most m qubits that always include
    qubits 0..low-1 (contiguous 2^low-amplitude runs in HBM)."""
    m = min(m, n)
    low = min(low, m)
    base = (1 << low) - 1
    remaining = list(ops)
    out = []
    while remaining:
        if remaining[0].kind in ("measure", "reset"):
            out.append(remaining.pop(0))
            continue
        mask = base
        chosen, keep, blocked = [], [], set()
        for op in remaining:
            qs = set(op.all_qubits)
            if op.kind in ("measure", "reset") or (qs & blocked):
                blocked |= qs
                keep.append(op)
                continue
            add = 0
            for q in qs:
                add |= 1 << q
            if bin(mask | add).count("1") <= m:
                mask |= add
                chosen.append(op)
            else:
                blocked |= qs
                keep.append(op)
        remaining = keep
        if not chosen:
            out.append(remaining.pop(0))
            continue
        q = 0
        while bin(mask).count("1") < m and q < n:
            mask |= 1 << q
            q += 1
        out.append(TilePass(mask, chosen))
    return out
