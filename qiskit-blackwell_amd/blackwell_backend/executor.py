"""Runs compiled circuits on the gfx950 library.

A CompiledCircuit owns the FP64 op list of one circuit and builds, per
execution mode and device, a list of prepared launches: ctypes argument
tuples computed once, with every matrix that must live on the device
uploaded once in a single tensor.  Executing a plan is then a tight loop of
library calls on torch's current stream (the reference rebuilds small
device tensors per op: blackwell_backend/simulator.py:348-404).

Modes:
  plain  - one kernel per op (per-gate baseline);
  blocks - 5-qubit dense blocks on FP32 MFMA (n >= 10), other entries plain;
  blocks/<k> - the same with k-qubit blocks, k = 4, 5, 6 (k = 4 needs
           n >= 11 and k = 6 n >= 13 for their MFMA kernels);
  tiles  - LDS tile programs, many ops per HBM pass;
  auto   - tiles with the whole state in one tile when n <= 13, else blocks;
  tune   - every applicable mode timed once on the device, fastest kept
           (the fusion width is one of the choices).
"""
import ctypes
import importlib

import numpy as np
import torch

from blackwell_kernels import _lib
from blackwell_kernels._lib import OP_DTYPE
from . import compiler as C

WHOLE_TILE_MAX = 13     # n <= 13: the whole state is one LDS tile (64 KiB)
BLOCK_MIN_QUBITS = 10   # the MFMA block kernel needs 2^10 amplitudes
# auto-mode cost model for 10 <= n <= 13 (MI355X measurements): one op of the
# one-workgroup LDS program costs ~2 us at 12 qubits and scales with the
# amplitudes per thread; one fused block costs ~5 us (launch-bound at this size)
WHOLE_OP_US_12Q = 2.0
BLOCK_US = 5.0
# plan steps a HIP graph can hold: library launches on the current stream with
# device-resident arguments only (no host uploads, no host-side decisions)
GRAPH_SAFE_STEPS = ("blocks", "mfma", "tile", "mcgate")
# fusion widths of the MFMA block kernels and the smallest state each runs on
BLOCK_WIDTHS = {4: 11, 5: BLOCK_MIN_QUBITS, 6: 13}
TUNE_REPS = 3
# Tile programs above 13 qubits: a tile of m qubits is one workgroup, so m
# trades ops fused per HBM pass against workgroups per pass (2^(n-m)).
# Measured on MI355X (A/B circuit families, 16 and 20 qubits): 10-qubit
# tiles beat 13-qubit ones by 2-3x there (13-qubit tiles give 8-128
# workgroups); larger states can afford larger tiles.
AUTO_TILES_MAX_QUBITS = 20   # auto considers tile programs up to this size
# auto-mode cost model for 14 <= n <= 20 (same measurements): a tile pass
# ~5 us plus ~0.85 us per op at 16 qubits (x1.13 per qubit; a dense 4x4 op
# counts 1.4, a wider dense op 2); a fused block ~8.2 us at 16 qubits (x1.07
# per qubit)
TILE_PASS_US = 5.0
TILE_OP_US_16Q = 0.85
TILE_OP_WEIGHT = {"m2": 1.4, "mk": 2.0}
BLOCK_US_16Q = 8.2


def block_width(mode):
    """Fusion width of a block mode ("blocks" is 5, "blocks/<k>" is k)."""
    return 5 if mode == "blocks" else int(mode.split("/")[1])


def block_modes(n):
    """The MFMA block modes that can run on n qubits ("blocks" is width 5)."""
    return [f"blocks/{k}" if k != 5 else "blocks" for k, nmin in sorted(BLOCK_WIDTHS.items()) if n >= nmin]


def default_tile_shape(n):
    """(tile qubits, always-included low qubits) of "tiles" for n qubits."""
    if n <= WHOLE_TILE_MAX:
        return n, min(4, n)
    m = 10 if n <= 20 else (11 if n <= 22 else (12 if n <= 24 else 13))
    return m, 4


def tune_tile_shapes(n):
    """Tile shapes the tuner times for n > 13 qubits."""
    m, low = default_tile_shape(n)
    shapes = [(m, low), (m, 3), (m - 1, low), (min(m + 1, WHOLE_TILE_MAX), low), (WHOLE_TILE_MAX, 4)]
    out = []
    for sh in shapes:
        if sh not in out and 1 <= sh[1] <= sh[0] <= WHOLE_TILE_MAX:
            out.append(sh)
    return out


def _d8(m):
    m = np.asarray(m, dtype=np.complex128).reshape(-1)
    v = np.empty(2 * m.size, dtype=np.float64)
    v[0::2] = m.real
    v[1::2] = m.imag
    return _lib.f64_array(v)


def _local(mask, q):
    return bin(mask & ((1 << q) - 1)).count("1")


def tile_ops(ops, mask):
    """[Op] -> (OP_DTYPE array with tile-local qubits, complex128 coefs)."""
    arr = np.zeros(len(ops), dtype=OP_DTYPE)
    coefs = []
    off = 0
    for i, op in enumerate(ops):
        loc = [_local(mask, q) for q in op.qubits]
        cm = 0
        for c in op.controls:
            cm |= 1 << _local(mask, c)
        arr[i]["cmask"] = cm
        arr[i]["nq"] = len(loc)
        for j, q in enumerate(loc):
            arr[i]["q"][j] = q
        k = op.kind
        if k == "m1":
            arr[i]["kind"] = _lib.OP_M1
            c = op.mat.reshape(-1)
        elif k == "d1":
            arr[i]["kind"] = _lib.OP_D1
            c = op.mat.reshape(-1)
        elif k == "ph":
            arr[i]["kind"] = _lib.OP_PH
            c = op.mat.reshape(-1)
        elif k == "m2":
            arr[i]["kind"] = _lib.OP_M2
            c = op.mat.reshape(-1)
        elif k == "mk":
            arr[i]["kind"] = _lib.OP_MK
            c = op.mat.reshape(-1)
        elif k == "x":
            arr[i]["kind"] = _lib.OP_X
            c = np.zeros(0, dtype=np.complex128)
        elif k == "swap":
            arr[i]["kind"] = _lib.OP_SWAP
            c = np.zeros(0, dtype=np.complex128)
        else:
            raise ValueError(f"op {k} cannot run inside a tile program")
        arr[i]["coef"] = off
        coefs.append(np.asarray(c, dtype=np.complex128))
        off += c.size
    cf = np.concatenate(coefs) if coefs else np.zeros(0, dtype=np.complex128)
    return arr, cf


class _Dev:
    """Device copies of a plan's matrices/op tables (one upload per plan)."""

    def __init__(self, device, mats=None, ops=None, coefs=None):
        self.mats = None if mats is None else torch.from_numpy(mats).to(device)
        self.ops = None if ops is None else torch.from_numpy(ops.view(np.uint8).copy()).to(device)
        self.coefs = None if coefs is None else torch.from_numpy(coefs).to(device)


class CompiledCircuit:
    def __init__(self, n_qubits, n_clbits, ops, global_phase=0.0, bk=None):
        """bk: the blackwell_kernels module the plans launch through (the
        simulator's, which its kernel_path selects); None = the one imported."""
        self.bk = bk if bk is not None else importlib.import_module("blackwell_kernels")
        self.n_qubits = n_qubits
        self.n_clbits = n_clbits
        self.ops = ops
        # the circuit's global phase enters as the initial amplitude e^{i phi}
        # (the reference drops it); 0 keeps the plain |0...0>
        phi = float(global_phase)
        self.phase0 = None if phi == 0.0 else complex(np.cos(phi), np.sin(phi))
        self._plans = {}
        self._blocks = None
        self._tiles = None
        self._auto = None
        self._graphs = {}
        self._tuned = {}
        self.has_stochastic = any(o.kind in ("measure", "reset") for o in ops)

    def init_state(self, state):
        """|0...0> times e^{i global_phase} into `state` (device launches only,
        so a graph can hold them)."""
        kl = self.bk._lib
        kl.call("qsim_state_init", kl.stream_handle(state.device), ctypes.c_void_p(state.data_ptr()),
                self.n_qubits)
        if self.phase0 is not None:
            state[:1].fill_(self.phase0)
        return state

    # ---------------------------------------------------------------- graphs
    def graphed(self, mode, device):
        """GraphedRun of |0...0> + the plan for `mode` (built once), or None
        when the plan has a step a graph cannot hold (mid-circuit
        measure/reset, host-uploaded op tables)."""
        key = (mode, str(device))
        if key not in self._graphs:
            plan = self.plan(mode, device)
            self._graphs[key] = (GraphedRun(plan, self, device)
                                 if not self.has_stochastic and plan.graph_safe() else None)
        return self._graphs[key]

    # ---------------------------------------------------------------- tuning
    def tuned_mode(self, device):
        """The fastest execution mode for this circuit on `device`, measured
        once: every applicable mode runs once to warm up and TUNE_REPS times
        timed with HIP events on the launch stream, from |0...0>.  Above 13
        qubits the candidates include LDS tile programs of several tile
        shapes ("tiles/<tile qubits>/<low qubits>")."""
        key = str(device)
        if key in self._tuned:
            return self._tuned[key]
        n = self.n_qubits
        cands = ["plain"]
        if n <= WHOLE_TILE_MAX:
            cands.append("whole")
        else:
            cands += [f"tiles/{m}/{low}" for m, low in tune_tile_shapes(n)]
        cands += block_modes(n)
        best = self._time_modes(cands, device)
        self._tuned[key] = best
        return best

    def tuned_block_width(self, device):
        """(mode, {width: ms}): the fusion width of the MFMA block plan
        measured fastest for this circuit on `device` (same timing rule as
        tuned_mode, candidates blocks/4, blocks/5, blocks/6)."""
        key = ("width", str(device))
        if key not in self._tuned:
            times = {}
            best = self._time_modes(block_modes(self.n_qubits), device, times)
            self._tuned[key] = (best, {block_width(m): t for m, t in times.items()})
        return self._tuned[key]

    def _time_modes(self, cands, device, times=None):
        n = self.n_qubits
        if not cands:
            return "plain"
        state = torch.empty(1 << n, dtype=torch.complex64, device=device)
        stream = torch.cuda.current_stream(device)
        best, best_ms = None, None
        for mode in cands:
            plan = self.plan(mode, device)
            reps = []
            for r in range(TUNE_REPS + 1):
                self.init_state(state)
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record(stream)
                plan.run(state)
                b.record(stream)
                b.synchronize()
                if r:
                    reps.append(a.elapsed_time(b))
            ms = float(np.median(reps))
            if times is not None:
                times[mode] = ms
            if best_ms is None or ms < best_ms:
                best, best_ms = mode, ms
        return best

    def block_entries(self, k=5):
        """plan_blocks of the ops with k-qubit blocks (computed once per
        compiled circuit and width)."""
        if self._blocks is None:
            self._blocks = {}
        if k not in self._blocks:
            self._blocks[k] = C.plan_blocks(self.ops, self.n_qubits, k=k)
        return self._blocks[k]

    # ---------------------------------------------------------------- plans
    def plan(self, mode, device):
        key = (mode, str(device))
        p = self._plans.get(key)
        if p is None:
            p = self._build(mode, device)
            self._plans[key] = p
        return p

    def tile_entries(self):
        """plan_tiles of the ops at the default tile shape (computed once)."""
        if self._tiles is None:
            m, low = default_tile_shape(self.n_qubits)
            self._tiles = C.plan_tiles(self.ops, self.n_qubits, m=m, low=low)
        return self._tiles

    def resolve_mode(self, mode, device=None):
        """auto: the whole state as one LDS program below 10 qubits; between
        10 and 13 qubits that or MFMA blocks, and between 14 and 20 qubits
        tile programs or MFMA blocks, whichever the cost model says is cheaper
        (deep circuits favour blocks: one block absorbs many ops); MFMA blocks
        above.  tune: the mode tuned_mode measured fastest on `device`."""
        if mode == "tune":
            if device is None:
                raise ValueError("mode 'tune' needs the device to measure on")
            return self.tuned_mode(device)
        if mode != "auto":
            return mode
        if self._auto is None:
            self._auto = self._auto_mode()
        return self._auto

    def _auto_mode(self):
        n = self.n_qubits
        if n < BLOCK_MIN_QUBITS:
            return "whole"
        if n > AUTO_TILES_MAX_QUBITS:
            return "blocks"
        if n > WHOLE_TILE_MAX:
            passes = sum(1 for e in self.tile_entries() if isinstance(e, C.TilePass))
            others = len(self.tile_entries()) - passes
            w_ops = sum(TILE_OP_WEIGHT.get(o.kind, 1.0) for o in self.ops)
            t_tiles = (passes * TILE_PASS_US + w_ops * TILE_OP_US_16Q * 1.13 ** (n - 16)
                       + others * BLOCK_US_16Q)
            t_blocks = len(self.block_entries()) * BLOCK_US_16Q * 1.07 ** (n - 16)
            return "tiles" if t_tiles < t_blocks else "blocks"
        t_whole = len(self.ops) * WHOLE_OP_US_12Q * 2.0 ** (n - 12)
        t_blocks = len(self.block_entries()) * BLOCK_US
        return "blocks" if t_blocks < t_whole else "whole"

    def _build(self, mode, device):
        n = self.n_qubits
        if mode == "plain":
            return self._build_plain(self.ops, device)
        if mode == "whole":
            mask = (1 << n) - 1
            unitary = [o for o in self.ops if o.kind not in ("measure", "reset")]
            if len(unitary) == len(self.ops):
                return self._build_tiles([C.TilePass(mask, self.ops)] if self.ops else [], device)
            mode = "tiles"
        if mode == "tiles":
            return self._build_tiles(self.tile_entries(), device)
        if mode.startswith("tiles/"):
            try:
                m, low = (int(x) for x in mode.split("/")[1:])
            except ValueError:
                raise ValueError(f"tile mode must be 'tiles/<m>/<low>' (got {mode!r})") from None
            if not (1 <= low <= m <= WHOLE_TILE_MAX):
                raise ValueError(f"tile shape needs 1 <= low <= m <= {WHOLE_TILE_MAX} (got {mode!r})")
            return self._build_tiles(C.plan_tiles(self.ops, n, m=min(m, n), low=low), device)
        if mode == "blocks":
            if n < 10:
                return self._build_tiles(C.plan_tiles(self.ops, n, m=n), device)
            return self._build_blocks(self.block_entries(), device)
        if mode.startswith("blocks/"):
            try:
                k = int(mode.split("/")[1])
            except ValueError:
                raise ValueError(f"block mode must be 'blocks/<k>' (got {mode!r})") from None
            if k not in BLOCK_WIDTHS or n < BLOCK_WIDTHS[k]:
                raise ValueError(f"{mode!r} needs k in {sorted(BLOCK_WIDTHS)} and n >= "
                                 f"{BLOCK_WIDTHS.get(k, '?')} (n = {n})")
            return self._build_blocks(self.block_entries(k), device)
        raise ValueError(f"unknown execution mode {mode!r}")

    def _build_plain(self, ops, device):
        n = self.n_qubits
        mats = []
        steps = []
        for op in ops:
            k, q = op.kind, op.qubits
            if k in ("measure", "reset"):
                steps.append((k, op))
                continue
            if op.controls and k in ("m1", "d1", "ph"):
                mats.append(op.target_matrix().astype(np.complex64).reshape(-1))
                steps.append(("mcgate", (len(mats) - 1, _lib.i64_array(op.controls),
                                         len(op.controls), q[0])))
            elif k == "m1":
                steps.append(("qsim_apply_gate_h", (n, _d8(op.mat), q[0])))
            elif k == "d1":
                steps.append(("qsim_apply_diagonal_h", (n, _d8(op.mat), q[0])))
            elif k == "ph":
                ph = complex(op.mat[0])
                steps.append(("qsim_apply_diagonal_phase", (n, ph.real, ph.imag, q[0])))
            elif k == "m2" and not op.controls:
                steps.append(("qsim_apply_gate_2q_h", (n, _d8(C.kernel_2q_matrix(op)),
                                                       q[0], q[1])))
            elif k == "x":
                steps.append(("qsim_apply_mcx", (n, _lib.i64_array(op.controls),
                                                 len(op.controls), q[0])))
            elif k == "swap":
                steps.append(("qsim_apply_mcswap", (n, _lib.i64_array(op.controls),
                                                    len(op.controls), q[0], q[1])))
            elif k in ("m2", "mk"):
                if op.controls:
                    raise ValueError("controlled multi-qubit dense ops are not supported")
                mats.append(op.mat.astype(np.complex64).reshape(-1))
                steps.append(("block", (len(mats) - 1, list(q))))
            else:
                raise ValueError(f"cannot execute op {op}")
        return _Plan(self, steps, mats, None, None, device)

    def _build_blocks(self, entries, device):
        n = self.n_qubits
        mats = []
        steps = []
        run = []                  # consecutive 5-qubit blocks -> one qsim_apply_blocks call

        def flush():
            if run:
                qs = np.full((len(run), 8), -1, dtype=np.int32)
                for i, b in enumerate(run):
                    qs[i, :5] = b.qubits
                steps.append(("blocks", (len(mats), len(run), qs)))
                mats.extend(b.mat.astype(np.complex64).reshape(-1) for b in run)
                run.clear()

        for e in entries:
            if isinstance(e, C.Block) and len(e.qubits) == 5:
                run.append(e)
                continue
            flush()
            k = len(e.qubits) if isinstance(e, C.Block) else 0
            if isinstance(e, C.Block) and k in BLOCK_WIDTHS and n >= BLOCK_WIDTHS[k]:
                # one MFMA launch with the block on the device (graph-safe)
                mats.append(e.mat.astype(np.complex64).reshape(-1))
                steps.append(("mfma", (len(mats) - 1, _lib.int_array(e.qubits), k)))
            elif isinstance(e, C.Block):
                mats.append(e.mat.astype(np.complex64).reshape(-1))
                steps.append(("block", (len(mats) - 1, list(e.qubits))))
            else:
                steps.append(("sub", self._build_plain([e], device)))
        flush()
        return _Plan(self, steps, mats, None, None, device)

    def _build_tiles(self, entries, device):
        all_ops, all_cf, steps = [], [], []
        op_off, cf_off = 0, 0
        for e in entries:
            if isinstance(e, C.TilePass):
                arr, cf = tile_ops(e.ops, e.mask)
                arr["coef"] += cf_off
                all_ops.append(arr)
                all_cf.append(cf)
                steps.append(("tile", (e.mask, op_off, len(arr))))
                op_off += len(arr)
                cf_off += cf.size
            else:
                steps.append(("sub", self._build_plain([e], device)))
        ops = np.concatenate(all_ops) if all_ops else np.zeros(0, dtype=OP_DTYPE)
        cf = np.concatenate(all_cf) if all_cf else np.zeros(0, dtype=np.complex128)
        if cf.size == 0:
            cf = np.zeros(1, dtype=np.complex128)
        return _Plan(self, steps, [], ops, cf.astype(np.complex64), device)


class _Plan:
    def __init__(self, cc, steps, mats, ops, coefs, device):
        self.cc = cc
        self.steps = steps
        self.n_launches = 0
        mats_arr = None
        self.mat_offsets = []
        if mats:
            off = 0
            for m in mats:
                self.mat_offsets.append(off)
                off += m.size
            mats_arr = np.concatenate(mats).astype(np.complex64)
        self.dev = _Dev(device, mats_arr, ops, coefs)
        for kind, p in steps:
            if kind == "sub":
                self.n_launches += p.n_launches
            elif kind == "blocks":
                self.n_launches += p[1]
            else:
                self.n_launches += 1

    def graph_safe(self):
        for kind, a in self.steps:
            if kind == "sub":
                if not a.graph_safe():
                    return False
            elif not (kind.startswith("qsim_") or kind in GRAPH_SAFE_STEPS):
                return False
        return True

    def _mat_ptr(self, idx):
        return ctypes.c_void_p(self.dev.mats.data_ptr() + 8 * self.mat_offsets[idx])

    def run(self, state, rand=None, clbits=None, step_timer=None):
        """Apply the plan to `state` (complex64 [2^n] on the plan's device).
        rand() supplies uniforms for measure/reset.  step_timer(kind, args),
        when given, returns a context manager entered around each step (for
        per-kernel-class timing with events on the launch stream)."""
        n = self.cc.n_qubits
        kl = self.cc.bk._lib
        s = kl.stream_handle(state.device)
        sp = ctypes.c_void_p(state.data_ptr())
        lib = kl.load()
        for kind, a in self.steps:
            if step_timer is not None and kind != "sub":
                with step_timer(kind, a):
                    self._step(kind, a, n, s, sp, lib, state, rand, clbits, None)
            else:
                self._step(kind, a, n, s, sp, lib, state, rand, clbits, step_timer)
        return state

    def _step(self, kind, a, n, s, sp, lib, state, rand, clbits, step_timer):
        bk = self.cc.bk
        call = bk._lib.call
        if kind.startswith("qsim_"):
            rc = getattr(lib, kind)(s, sp, *a)
            if rc:
                call(kind, s, sp, *a)  # re-raise with the message
        elif kind == "blocks":
            first, count, qs = a
            call("qsim_apply_blocks", s, sp, n, self._mat_ptr(first),
                 ctypes.c_void_p(qs.ctypes.data), count)
        elif kind == "tile":
            mask, off, cnt = a
            call("qsim_tile_program", s, sp, n, mask,
                 ctypes.c_void_p(self.dev.ops.data_ptr() + 16 * off), cnt,
                 ctypes.c_void_p(self.dev.coefs.data_ptr()))
        elif kind == "mcgate":
            idx, ctl, nc, t = a
            call("qsim_apply_mcgate", s, sp, n, self._mat_ptr(idx), ctl, nc, t)
        elif kind == "mfma":
            idx, qs, k = a
            call("qsim_apply_block_mfma", s, sp, n, self._mat_ptr(idx), qs, k)
        elif kind == "block":
            idx, qs = a
            k = len(qs)
            blk = self.dev.mats[self.mat_offsets[idx]:self.mat_offsets[idx] + (1 << (2 * k))]
            bk.apply_block(state, blk.reshape(1 << k, 1 << k), qs)
        elif kind == "sub":
            a.run(state, rand, clbits, step_timer)
        elif kind in ("measure", "reset"):
            op = a
            u = rand() if rand is not None else float(torch.rand(1).item())
            o = bk.measure_qubit(state, op.qubits[0], u)
            if kind == "measure":
                if clbits is not None and op.clbit is not None and op.clbit < len(clbits):
                    clbits[op.clbit] = o
            elif o == 1:
                bk.apply_mcx(state, [], op.qubits[0])
        else:
            raise RuntimeError(f"unknown step {kind}")


class GraphedRun:
    """A HIP graph (torch.cuda.CUDAGraph) of qsim_state_init + one plan over a
    state buffer the graph owns.  replay() relaunches the whole circuit with
    one host call; the buffer holds the result until the next replay.  The
    plan runs once eagerly on a side stream before capture, so every lazy
    allocation (workspaces, plan tables) exists before the graph records
    pointers."""

    def __init__(self, plan, cc, device):
        self.cc = cc
        self.buf = torch.empty(1 << cc.n_qubits, dtype=torch.complex64, device=device)
        side = torch.cuda.Stream(device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            self._body(plan)
        torch.cuda.current_stream(device).wait_stream(side)
        torch.cuda.synchronize(device)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body(plan)
        torch.cuda.synchronize(device)

    def _body(self, plan):
        self.cc.init_state(self.buf)
        plan.run(self.buf)

    def replay(self):
        self.graph.replay()
        return self.buf
