"""BlackwellSimulator on MI355X: the reference's Qiskit backend surface
(blackwell_backend/simulator.py:86-723) over the gfx950 library.

run / run_statevector / run_and_time / run_hamiltonian keep the reference's
signatures and results: hex counts over all qubits with measurements stripped
before execution (simulator.py:555-604), the complex64 statevector
(:606-613), timing stats (:615-644), and the Chebyshev entry (:649-723).

Differences by design:
  * circuits are compiled by blackwell_backend.compiler (FP64 host fusion)
    into plans of prepared launches (blackwell_backend.executor);
  * the default execution mode is "auto": the whole circuit in one LDS tile
    program when the state fits (n <= 13), 5-qubit FP32-MFMA blocks above;
    mode="plain" runs one kernel per gate, mode="blocks/<k>" fuses into
    k-qubit MFMA blocks (k = 4, 5, 6) and mode="tune" times the modes and
    fusion widths once per circuit and keeps the fastest;
  * sampling uses FP64 uniforms and an FP64 inverse CDF; counts are tallied
    with numpy instead of a Python loop over shots;
  * a circuit's global_phase multiplies the statevector (the reference
    drops it), as the initial amplitude e^{i phase};
  * qiskit is optional: with it installed, circuits with non-native gates
    are first transpiled to the native basis at optimisation level 1 as the
    reference does (simulator.py:489-498), and qiskit Result objects are
    returned.
"""
import importlib
import os
import weakref
import sys
import time

import numpy as np
import torch

from . import compiler as C
from .circuit import fingerprint, global_phase, instructions
from .executor import CompiledCircuit
from .result import make_result

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_DEFAULT_KERNEL_PATH = _PKG_ROOT


def _load_kernels(kernel_path=None):
    """Import blackwell_kernels from kernel_path | $BLACKWELL_KERNEL_PATH |
    this package's directory (simulator.py:46-63).  The requested directory
    goes to the front of sys.path, and a module already imported from another
    directory is replaced, so two builds can be loaded one after the other
    (the reference keeps an existing sys.path entry where it is, so a path
    listed behind another build's did not win, and compares directories as
    substrings)."""
    path = os.path.abspath(kernel_path or os.environ.get("BLACKWELL_KERNEL_PATH", _DEFAULT_KERNEL_PATH))
    while path in sys.path:
        sys.path.remove(path)
    sys.path.insert(0, path)
    existing = sys.modules.get("blackwell_kernels")
    if existing is not None:
        ex_dir = (getattr(existing, "__path__", [None]) or [None])[0] or ""
        if os.path.dirname(os.path.abspath(ex_dir)) != path:
            for key in list(sys.modules):
                if key == "blackwell_kernels" or key.startswith("blackwell_kernels."):
                    del sys.modules[key]
    bk = importlib.import_module("blackwell_kernels")
    bk.load_library()
    return bk


def counts_from_samples(samples):
    """{hex(index): count} over all qubits (simulator.py:580-586)."""
    s = np.asarray(samples, dtype=np.int64)
    # a histogram is cheaper than a sort while its length stays near the shot count
    if s.size and int(s.max()) < max(4 * s.size, 1 << 12):
        bc = np.bincount(s)
        vals = np.flatnonzero(bc)
        cnt = bc[vals]
    else:
        vals, cnt = np.unique(s, return_counts=True)
    return dict(zip(map(hex, vals.tolist()), cnt.tolist()))


_QISKIT_QC = []


def _qiskit_circuit_class():
    """qiskit.QuantumCircuit, or None when qiskit is not installed (probed once)."""
    if not _QISKIT_QC:
        try:
            from qiskit import QuantumCircuit as QC
            _QISKIT_QC.append(QC)
        except ImportError:
            _QISKIT_QC.append(None)
    return _QISKIT_QC[0]


class BlackwellSimulator:
    """Statevector simulator on the MI355X gfx950 kernels.

        sim = BlackwellSimulator()
        result = sim.run(circuit, shots=1024)
    """

    MODES = ("auto", "plain", "blocks", "tiles", "whole", "tune")
    # largest circuit run() replays as a HIP graph (the graph owns a state
    # buffer of 8 * 2^n bytes per cached circuit)
    GRAPH_MAX_QUBITS = 20

    def __init__(self, device="cuda", kernel_path=None, mode="auto", use_graphs=True):
        if mode not in self.MODES and not str(mode).startswith(("tiles/", "blocks/")):
            raise ValueError(f"mode must be one of {self.MODES}, 'tiles/<m>/<low>' or 'blocks/<k>'")
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.kernel_path = kernel_path
        self.bk = _load_kernels(kernel_path)
        self.mode = mode
        self.use_graphs = bool(use_graphs)
        self._native_1q = set(C._NATIVE_1Q)
        self._native_2q = {"cx", "cz", "swap", "cp", "crz", "cu"}
        self._native_mcx = {"ccx"}
        self._native_other = {"measure", "barrier", "reset"}
        self._compile_cache = {}
        # per circuit object: (fingerprint, cache key, measurement-free copy)
        self._memo = weakref.WeakKeyDictionary()
        self._pm = None

    # ---------------------------------------------------------------- compile
    def _get_pass_manager(self):
        if self._pm is None:
            from qiskit.transpiler.preset_passmanagers import generate_preset_pass_manager
            basis = sorted(self._native_1q | self._native_2q | self._native_mcx)
            self._pm = generate_preset_pass_manager(optimization_level=1, basis_gates=basis)
        return self._pm

    @staticmethod
    def _circuit_cache_key(circuit):
        parts = []
        for name, qs, cs, params in instructions(circuit):
            pv = []
            for p in params:
                if isinstance(p, np.ndarray):
                    pv.append(hash(p.tobytes()))
                    continue
                try:
                    pv.append(float(p))
                except (TypeError, ValueError):
                    pv.append(hash(str(p)))
            parts.append((name, tuple(pv), tuple(qs), tuple(cs)))
        return (circuit.num_qubits, circuit.num_clbits, hash(tuple(parts)), global_phase(circuit))

    def _memo_entry(self, circuit):
        """Memo of one of this module's circuits while its content is unchanged."""
        fp = fingerprint(circuit)
        if fp is None:
            return None
        try:
            ent = self._memo.get(circuit)
        except TypeError:
            return None
        if ent is None or ent[0] != fp:
            ent = [fp, None, None]
            self._memo[circuit] = ent
        return ent

    def _key(self, circuit):
        ent = self._memo_entry(circuit)
        if ent is None:
            return self._circuit_cache_key(circuit)
        if ent[1] is None:
            ent[1] = self._circuit_cache_key(circuit)
        return ent[1]

    def compile(self, circuit):
        """CompiledCircuit for `circuit` (memoised like simulator.py:517-527)."""
        key = self._key(circuit)
        cc = self._compile_cache.get(key)
        if cc is not None:
            return cc
        instrs = instructions(circuit)
        if any(name not in C.NATIVE_NAMES for name, *_ in instrs):
            try:
                pm = self._get_pass_manager()
            except ImportError as e:
                bad = sorted({nm for nm, *_ in instrs if nm not in C.NATIVE_NAMES})
                raise ValueError(f"Unsupported instruction(s) {bad} (qiskit is not installed "
                                 "to transpile them)") from e
            circuit = pm.run(circuit)
            instrs = instructions(circuit)
        ops = C.compile_ops(instrs)
        cc = CompiledCircuit(circuit.num_qubits, circuit.num_clbits, ops, global_phase(circuit), bk=self.bk)
        self._compile_cache[key] = cc
        return cc

    # ---------------------------------------------------------------- execute
    def _execute_circuit(self, circuit, mode=None, graphed=False):
        """(state, clbits, elapsed_s) — state stays on the device.

        graphed: callers that consume the state before the next execution
        (run, run_statevector, run_and_time) get circuits of at most
        GRAPH_MAX_QUBITS qubits without mid-circuit measurement replayed from a
        HIP graph; the returned state is then the graph's own buffer."""
        cc = self.compile(circuit)
        mode = cc.resolve_mode(mode or self.mode, self.device)
        if graphed and self.use_graphs and cc.n_qubits <= self.GRAPH_MAX_QUBITS:
            g = cc.graphed(mode, self.device)
            if g is not None:
                torch.cuda.synchronize(self.device)
                t0 = time.perf_counter()
                state = g.replay()
                torch.cuda.synchronize(self.device)
                return state, [0] * cc.n_clbits, time.perf_counter() - t0
        plan = cc.plan(mode, self.device)
        state = cc.init_state(torch.empty(1 << cc.n_qubits, dtype=torch.complex64, device=self.device))
        clbits = [0] * cc.n_clbits
        torch.cuda.synchronize(self.device)
        t0 = time.perf_counter()
        plan.run(state, rand=lambda: float(torch.rand(1, dtype=torch.float64).item()),
                 clbits=clbits)
        torch.cuda.synchronize(self.device)
        return state, clbits, time.perf_counter() - t0

    def _strip_measurements(self, circuit):
        """Copy of `circuit` without measure/reset (simulator.py:546-553)."""
        ent = self._memo_entry(circuit)
        if ent is not None and ent[2] is not None:
            return ent[2]
        QC = _qiskit_circuit_class()
        if QC is not None and type(circuit).__module__.startswith("qiskit"):
            new = QC(circuit.num_qubits, circuit.num_clbits, name=circuit.name)
            new.global_phase = circuit.global_phase
            for ins in circuit.data:
                if ins.operation.name.lower() not in ("measure", "reset"):
                    new.append(ins.operation, ins.qubits, ins.clbits)
            return new
        new = circuit.copy()
        new.data = [ins for ins in new.data if ins.operation.name.lower() not in ("measure", "reset")]
        if ent is not None:
            ent[2] = new
        return new

    def run(self, circuit, shots=1024, seed=None, mode=None):
        """Qiskit-style Result with hex counts over all qubits."""
        if seed is not None:
            torch.manual_seed(seed)
            np.random.seed(seed)
        clean = self._strip_measurements(circuit)
        state, _, elapsed = self._execute_circuit(clean, mode, graphed=True)
        samples = self.bk.sample(state, shots)
        counts = counts_from_samples(samples.cpu().numpy())
        n = int(state.numel()).bit_length() - 1
        return make_result(counts, shots, {"name": circuit.name, "n_qubits": n}, elapsed)

    def run_statevector(self, circuit, mode=None):
        """(numpy complex64 statevector, elapsed_s)."""
        state, _, elapsed = self._execute_circuit(circuit, mode, graphed=True)
        return state.cpu().numpy(), elapsed

    def run_and_time(self, circuit, shots=1024, seed=None, warmup=2, repeats=5, mode=None):
        if seed is not None:
            torch.manual_seed(seed)
            np.random.seed(seed)
        for _ in range(warmup):
            self._execute_circuit(circuit, mode, graphed=True)
        times = []
        for _ in range(repeats):
            _, _, el = self._execute_circuit(circuit, mode, graphed=True)
            times.append(el)
        result = self.run(circuit, shots=shots, seed=seed, mode=mode)
        stats = {"times_s": times, "median_s": float(np.median(times)),
                 "min_s": float(np.min(times)), "mean_s": float(np.mean(times))}
        return result, stats

    # ---------------------------------------------------------------- Chebyshev
    def run_hamiltonian(self, hamiltonian, time_val, initial_state=None, precision=1e-10,
                        shots=None):
        """exp(-iHt)|psi>; (statevector, info) or (Result, info) with shots."""
        from .hamiltonian import HamiltonianCSR
        from .chebyshev import chebyshev_evolve
        if isinstance(hamiltonian, HamiltonianCSR):
            h = hamiltonian
        else:
            h = HamiltonianCSR.from_sparse_pauli_op(hamiltonian, device=self.device)
        if initial_state is None:
            st = self.bk.state_init(h.n_qubits, device=self.device)
        else:
            st = torch.as_tensor(initial_state)
            if st.dtype != torch.complex64:
                st = st.to(torch.complex64)
            st = st.to(self.device).contiguous()
        out, info = chebyshev_evolve(st, h, time_val, precision=precision, bk=self.bk)
        if shots is None:
            return out.cpu().numpy(), info
        samples = self.bk.sample(out, shots)
        counts = counts_from_samples(samples.cpu().numpy())
        res = make_result(counts, shots, {"name": "chebyshev_evolution", "n_qubits": h.n_qubits},
                          info["elapsed_s"], metadata=info)
        return res, info
