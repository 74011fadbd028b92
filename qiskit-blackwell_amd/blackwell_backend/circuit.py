"""Minimal circuit front-end with the slice of Qiskit's QuantumCircuit
surface that BlackwellSimulator reads: `data` (instructions with
`.operation.name`, `.operation.params`, `.qubits`, `.clbits`), `qubits`,
`clbits`, `num_qubits`, `num_clbits`, `name`.

Qiskit is not a dependency: when it is installed, real qiskit circuits are
read through the same attributes; this class covers building circuits
without it (tests, benchmarks, serving workloads).
"""
import itertools

import numpy as np


class Bit:
    __slots__ = ("index", "kind")

    def __init__(self, index, kind):
        self.index = index
        self.kind = kind

    def __repr__(self):
        return f"{self.kind}({self.index})"


class _Frozen:
    """Attributes are set once in __init__; array parameters are stored as
    read-only copies.  Instructions never change after they are built, so a
    circuit's content is identified by the identities of its instructions
    (see fingerprint)."""
    __slots__ = ()

    def __setattr__(self, name, value):
        raise AttributeError(f"{type(self).__name__} is immutable")


def _frozen_param(p):
    if isinstance(p, np.ndarray):
        p = np.array(p, copy=True)
        p.setflags(write=False)
    return p


class Operation(_Frozen):
    __slots__ = ("name", "params", "num_qubits")

    def __init__(self, name, params=(), num_qubits=None):
        object.__setattr__(self, "name", name)
        object.__setattr__(self, "params", tuple(_frozen_param(p) for p in params))
        object.__setattr__(self, "num_qubits", num_qubits)


class CircuitInstruction(_Frozen):
    __slots__ = ("operation", "qubits", "clbits")

    def __init__(self, operation, qubits, clbits=()):
        object.__setattr__(self, "operation", operation)
        object.__setattr__(self, "qubits", tuple(qubits))
        object.__setattr__(self, "clbits", tuple(clbits))


_ids = itertools.count()


class QuantumCircuit:
    """Circuit of named gates on integer qubits 0..n-1 (qubit 0 = LSB)."""

    def __init__(self, num_qubits, num_clbits=0, name=None):
        self.qubits = [Bit(i, "Qubit") for i in range(int(num_qubits))]
        self.clbits = [Bit(i, "Clbit") for i in range(int(num_clbits))]
        self.name = name or f"circuit-{next(_ids)}"
        self.data = []
        # e^{i global_phase} multiplies the whole state (Qiskit's attribute)
        self.global_phase = 0.0

    @property
    def num_qubits(self):
        return len(self.qubits)

    @property
    def num_clbits(self):
        return len(self.clbits)

    def _q(self, q):
        if isinstance(q, Bit):
            return q
        q = int(q)
        if q < 0 or q >= len(self.qubits):
            raise IndexError(f"qubit {q} out of range for {len(self.qubits)} qubits")
        return self.qubits[q]

    def _c(self, c):
        return c if isinstance(c, Bit) else self.clbits[int(c)]

    def append(self, operation, qargs, cargs=()):
        if isinstance(operation, str):
            operation = Operation(operation)
        qs = [self._q(q) for q in qargs]
        if len({id(q) for q in qs}) != len(qs):
            raise ValueError("duplicate qubit arguments")
        self.data.append(CircuitInstruction(operation, qs, [self._c(c) for c in cargs]))
        return self

    def _g(self, name, qubits, params=()):
        return self.append(Operation(name, params, len(qubits)), qubits)

    # 1Q gates
    def h(self, q): return self._g("h", [q])
    def x(self, q): return self._g("x", [q])
    def y(self, q): return self._g("y", [q])
    def z(self, q): return self._g("z", [q])
    def s(self, q): return self._g("s", [q])
    def sdg(self, q): return self._g("sdg", [q])
    def t(self, q): return self._g("t", [q])
    def tdg(self, q): return self._g("tdg", [q])
    def id(self, q): return self._g("id", [q])
    def sx(self, q): return self._g("sx", [q])
    def sxdg(self, q): return self._g("sxdg", [q])
    def rx(self, theta, q): return self._g("rx", [q], [theta])
    def ry(self, theta, q): return self._g("ry", [q], [theta])
    def rz(self, phi, q): return self._g("rz", [q], [phi])
    def p(self, lam, q): return self._g("p", [q], [lam])
    def u1(self, lam, q): return self._g("u1", [q], [lam])
    def u2(self, phi, lam, q): return self._g("u2", [q], [phi, lam])
    def u3(self, theta, phi, lam, q): return self._g("u3", [q], [theta, phi, lam])
    def u(self, theta, phi, lam, q): return self._g("u", [q], [theta, phi, lam])

    # 2Q / multi-qubit gates (first qubit = control where there is one)
    def cx(self, c, t): return self._g("cx", [c, t])
    def cz(self, c, t): return self._g("cz", [c, t])
    def swap(self, a, b): return self._g("swap", [a, b])
    def cp(self, lam, c, t): return self._g("cp", [c, t], [lam])
    def crz(self, theta, c, t): return self._g("crz", [c, t], [theta])
    def cu(self, theta, phi, lam, gamma, c, t): return self._g("cu", [c, t], [theta, phi, lam, gamma])
    def ccx(self, c1, c2, t): return self._g("ccx", [c1, c2, t])
    def cswap(self, c, a, b): return self._g("cswap", [c, a, b])

    def mcx(self, controls, target):
        return self._g("mcx", list(controls) + [target])

    def unitary(self, matrix, qubits, label=None):
        """Dense matrix on `qubits` (row bit j <-> qubits[j])."""
        m = np.asarray(matrix, dtype=np.complex128)
        if m.shape != (1 << len(qubits), 1 << len(qubits)):
            raise ValueError("unitary shape does not match the qubit count")
        return self.append(Operation("unitary", [m], len(qubits)), qubits)

    # non-unitary
    def measure(self, q, c): return self.append(Operation("measure", (), 1), [q], [c])
    def reset(self, q): return self.append(Operation("reset", (), 1), [q])

    def barrier(self, *qs):
        qs = qs or range(self.num_qubits)
        return self.append(Operation("barrier", (), len(list(qs))), list(qs))

    def measure_all(self):
        start = len(self.clbits)
        self.clbits += [Bit(start + i, "Clbit") for i in range(self.num_qubits)]
        for i in range(self.num_qubits):
            self.measure(i, start + i)
        return self

    def copy(self, name=None):
        c = QuantumCircuit(self.num_qubits, self.num_clbits, name or self.name)
        c.global_phase = self.global_phase
        for ins in self.data:
            c.data.append(CircuitInstruction(
                ins.operation,
                [c.qubits[q.index] for q in ins.qubits],
                [c.clbits[b.index] for b in ins.clbits]))
        return c

    def size(self):
        return sum(1 for d in self.data if d.operation.name not in ("barrier",))

    def __len__(self):
        return len(self.data)

    def __repr__(self):
        return f"QuantumCircuit({self.num_qubits} qubits, {len(self.data)} instructions, name={self.name!r})"


def fingerprint(circuit):
    """Cheap identity of the content of one of this module's circuits
    (instructions are immutable, so their identities and order determine the
    content); None for circuits of other classes."""
    if not isinstance(circuit, QuantumCircuit):
        return None
    return (circuit.num_qubits, circuit.num_clbits, len(circuit.data),
            hash(tuple(map(id, circuit.data))), float(circuit.global_phase))


def global_phase(circuit):
    """The circuit's global phase in radians (0 when it has none)."""
    try:
        return float(getattr(circuit, "global_phase", 0.0) or 0.0)
    except TypeError as e:
        raise ValueError("the circuit's global phase has unbound parameters") from e


def instructions(circuit):
    """[(name, [qubit indices], [clbit indices], params)] of any circuit with
    Qiskit's data layout (this module's or qiskit's)."""
    qidx = {id(q): i for i, q in enumerate(circuit.qubits)}
    cidx = {id(c): i for i, c in enumerate(circuit.clbits)}
    out = []
    for ins in circuit.data:
        op = ins.operation
        qs = [qidx[id(q)] if id(q) in qidx else circuit.qubits.index(q) for q in ins.qubits]
        cs = [cidx[id(c)] if id(c) in cidx else circuit.clbits.index(c) for c in ins.clbits]
        out.append((op.name.lower(), qs, cs, list(op.params)))
    return out
