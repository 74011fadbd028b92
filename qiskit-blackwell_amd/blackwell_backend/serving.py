"""Continuous-batching circuit server on one MI355X.

Requests (circuit, shots, seed) arrive over time.  The server admits them as
capacity frees up, keeps every state resident on the device from the first
gate to the last shot, and returns seeded samples.  Two lanes run
concurrently on separate HIP streams:

  * LDS lane (n <= 14): many circuits per launch of qsim_run_instances, one
    workgroup per circuit with its whole state in LDS, its own gate table,
    and in-kernel sampling.  Nothing but the samples touches HBM.
  * HBM lane (n > 14): each circuit in its own slab of HBM, run as fused
    5-qubit blocks on FP32 MFMA (one qsim_apply_blocks call per circuit),
    then sampled with qsim_sample.  Block plans are built ahead of time on a
    pool of host threads (the native planner releases the GIL), so planning
    overlaps the device work of the circuits before it.

Admission is by statevector bytes against an HBM budget (default 80 % of
the free device memory; tests inject a small one).  A request that arrives
while others run is admitted at the next scheduling point if its bytes fit;
LDS-lane requests are grouped into a launch as soon as the previous LDS
launch has been handed to the device, so their work overlaps with the
running batches and the HBM lane.

Repeated circuits on the HBM lane replay as HIP graphs: a request that
carries a circuit (no gate payload) and whose compiled circuit the server
has already run once is executed by replaying a captured graph of state
init + its block plan (one host call instead of one launch per block), into
the graph's own buffer, which is sampled before the next replay on the
same stream.  Gate-payload requests are unique by construction and run as
batched block steps instead.

Sampling in both lanes uses the same rule and the same uniforms for a given
seed: u_k = SplitMix64(seed + (k+1) * golden) >> 11 scaled by 2^-53 and
index = smallest i with cdf(i) > u_k * ||psi||^2.  Each instance's arithmetic
is independent of the batch it runs in (one workgroup per instance, fixed
thread count), so a request's samples are bitwise the same alone and in a
mix.

The reference has no serving loop; its batched precedent is fixed-n and
single-launch (bwk/csrc/qv8_sim_sm120.cu:58-151).
"""
import collections
import ctypes
import time
import weakref

import numpy as np
import torch

from blackwell_kernels import _lib
from blackwell_kernels._lib import call, BLOCK_REC_DTYPE, OP_DTYPE, INSTANCE_DTYPE, SAMPLE_REC_DTYPE
from . import compiler as C
from .circuit import instructions
from .executor import CompiledCircuit, tile_ops
from . import workloads as W

LDS_MAX_QUBITS = 14
# HBM-lane states up to this size are sampled together (qsim_sample_batch)
SAMPLE_BATCH_MAX_QUBITS = 20
# state bytes per HBM-lane batch: the 256 MB MALL (Infinity Cache) of MI355X,
# which keeps a batch's states resident between its block steps
HBM_BATCH_BYTES = 1 << 28
GOLDEN = 0x9E3779B97F4A7C15


def splitmix_uniforms(seed, count):
    """Host copy of the in-kernel uniforms (qsim_uniforms_splitmix draws the
    same numbers on the device)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


class Request:
    """One circuit to serve.  Either `circuit` (any circuit with Qiskit's
    data layout) or a (`program`, `n_qubits`) pair from a program builder.
    With keep_state the final statevector (complex64) is copied back into
    .state beside the samples."""
    __slots__ = ("rid", "arrival", "n_qubits", "shots", "seed", "circuit", "program", "gates",
                 "t_admit", "t_done", "samples", "lane", "keep_state", "state")

    def __init__(self, rid, arrival, n_qubits, shots, seed, circuit=None, program=None, gates=None,
                 keep_state=False):
        self.rid = rid
        self.arrival = float(arrival)
        self.n_qubits = int(n_qubits)
        self.shots = int(shots)
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.circuit = circuit
        self.program = program
        # (pairs int [G, 2], unitaries complex128 [G, 4, 4]): a circuit of
        # two-qubit gates as the client sends it (row bit j <-> pairs[g, j])
        self.gates = gates
        self.t_admit = None
        self.t_done = None
        self.samples = None
        self.lane = None
        self.keep_state = bool(keep_state)
        self.state = None

    def state_bytes(self):
        return 8 << self.n_qubits

    def counts(self):
        v, c = np.unique(self.samples, return_counts=True)
        return {hex(int(a)): int(b) for a, b in zip(v, c)}


# ---------------------------------------------------------------- programs
def gates_program(pairs, us):
    """Whole-state LDS program (OP_DTYPE ops, complex128 coefs) of a list of
    two-qubit gates, built without per-gate objects."""
    g = len(pairs)
    ops = np.zeros(g, dtype=OP_DTYPE)
    ops["kind"] = _lib.OP_M2
    ops["nq"] = 2
    ops["q"][:, 0] = pairs[:, 0]
    ops["q"][:, 1] = pairs[:, 1]
    ops["coef"] = np.arange(g, dtype=np.uint32) * 16
    return ops, np.asarray(us, dtype=np.complex128).reshape(-1)


def su4_program(n, depth, seed):
    """gates_program of workloads.su4_volume(n, depth, seed)."""
    return gates_program(*W.su4_arrays(n, depth, seed))


def circuit_program(circuit):
    """Whole-state LDS program of an arbitrary circuit (via the compiler)."""
    ops = C.compile_ops(instructions(circuit))
    if any(o.kind in ("measure", "reset") for o in ops):
        raise ValueError("served circuits may not contain mid-circuit measure/reset")
    n = circuit.num_qubits
    return tile_ops(ops, (1 << n) - 1)


def su4_request(rid, arrival, n, shots, seed, keep_state=False):
    """Request for the random-SU(4) volume circuit of depth n, carrying its
    gates (built here, on the client side, before it is submitted)."""
    return Request(rid, arrival, n, shots, seed, program=("su4", n, n, seed),
                   gates=W.su4_arrays(n, n, seed), keep_state=keep_state)


def _materialise(req):
    """(ops, coefs) for the LDS lane."""
    if req.gates is not None:
        return gates_program(*req.gates)
    p = req.program
    if isinstance(p, tuple) and p and isinstance(p[0], str) and p[0] == "su4":
        return su4_program(p[1], p[2], p[3])
    if p is not None:
        return p
    return circuit_program(req.circuit)


def _circuit_of(req):
    if req.circuit is not None:
        return req.circuit
    p = req.program
    if isinstance(p, tuple) and p and isinstance(p[0], str) and p[0] == "su4":
        return W.su4_volume(p[1], p[2], p[3])
    raise ValueError("request has neither a circuit nor a rebuildable program")


def _batch_programs(reqs):
    """Per-request op and coefficient counts, and the batch's concatenated
    ops (coefficient offsets global) and complex64 coefficients.  Requests
    that all carry gate payloads are stacked without per-request programs;
    the result is the same as concatenating their gates_program()s."""
    if all(r.gates is not None for r in reqs):
        pairs = [np.asarray(r.gates[0]) for r in reqs]
        n_ops = np.array([len(pp) for pp in pairs], dtype=np.int64)
        total = int(n_ops.sum())
        ops = np.zeros(total, dtype=OP_DTYPE)
        if total:
            pq = np.concatenate(pairs)
            ops["kind"] = _lib.OP_M2
            ops["nq"] = 2
            ops["q"][:, 0] = pq[:, 0]
            ops["q"][:, 1] = pq[:, 1]
            ops["coef"] = np.arange(total, dtype=np.uint32) * 16
            cf = np.concatenate([np.asarray(r.gates[1]).reshape(-1) for r in reqs]).astype(np.complex64)
        else:
            cf = np.zeros(0, np.complex64)
        return n_ops, 16 * n_ops, ops, cf
    progs = [_materialise(r) for r in reqs]
    n_ops = np.array([len(o) for o, _ in progs], dtype=np.int64)
    n_cf = np.array([len(c) for _, c in progs], dtype=np.int64)
    ops = np.concatenate([o for o, _ in progs])       # a new array: the programs stay as built
    ops["coef"] += np.repeat(np.concatenate([[0], np.cumsum(n_cf)[:-1]]), n_ops).astype(np.uint32)
    cf = np.concatenate([np.asarray(c, dtype=np.complex64) for _, c in progs])
    return n_ops, n_cf, ops, cf


# ---------------------------------------------------------------- LDS lane
class _LdsBatch:
    def __init__(self, reqs, device, stream):
        self.reqs = reqs
        self.device = device
        self.stream = stream
        self.event = None

    def launch(self):
        # the batch's records are assembled column-wise: per-request Python
        # work is only the (ops, coefs) of each circuit
        reqs = self.reqs
        n_ops, n_cf, ops_all, blob_cf = _batch_programs(reqs)
        shots = np.array([r.shots for r in reqs], dtype=np.int64)
        nq = np.array([r.n_qubits for r in reqs], dtype=np.int64)
        want = np.array([r.keep_state for r in reqs], dtype=bool)
        op_off = np.concatenate([[0], np.cumsum(n_ops)])
        cf_off = np.concatenate([[0], np.cumsum(n_cf)])
        shot_offs = np.concatenate([[0], np.cumsum(shots)])
        st_sizes = np.where(want, np.left_shift(1, nq), 0)
        st_offs = np.concatenate([[0], np.cumsum(st_sizes)])
        inst = np.zeros(len(reqs), dtype=INSTANCE_DTYPE)
        inst["seed"] = np.array([r.seed for r in reqs], dtype=np.uint64)
        inst["shot_off"] = shot_offs[:-1]
        inst["op_off"] = op_off[:-1]
        inst["n_ops"] = n_ops
        inst["shots"] = shots
        inst["n_qubits"] = nq
        inst["init"] = 1
        inst["want"] = want
        inst["state_off"] = st_offs[:-1]
        self.state_off = [int(o) for o in st_offs[:-1]]
        op_off, cf_off, shot_off, st_off = int(op_off[-1]), int(cf_off[-1]), int(shot_offs[-1]), int(st_offs[-1])
        max_n = max(1, int(nq.max()))
        blob_ops = ops_all.view(np.uint8)
        if not cf_off:
            blob_cf = np.zeros(1, np.complex64)
        self.shot_offs = shot_offs
        with torch.cuda.stream(self.stream):
            self.d_inst = torch.from_numpy(inst.view(np.uint8).copy()).to(self.device)
            self.d_ops = torch.from_numpy(blob_ops.copy()).to(self.device)
            self.d_cf = torch.from_numpy(blob_cf).to(self.device)
            self.d_out = torch.empty(max(shot_off, 1), dtype=torch.int64, device=self.device)
            # final amplitudes only for keep_state requests (want = 1)
            self.d_states = (torch.empty(st_off, dtype=torch.complex64, device=self.device)
                             if st_off else self.d_out)
            self.h_inst, self.h_ops = inst, blob_ops       # validated host copies of the records
            call("qsim_run_instances", ctypes.c_void_p(self.stream.cuda_stream),
                 ctypes.c_void_p(self.d_states.data_ptr()), st_off,
                 ctypes.c_void_p(inst.ctypes.data), ctypes.c_void_p(self.d_inst.data_ptr()),
                 len(self.reqs), ctypes.c_void_p(blob_ops.ctypes.data),
                 ctypes.c_void_p(self.d_ops.data_ptr()), op_off,
                 ctypes.c_void_p(self.d_cf.data_ptr()), int(blob_cf.size),
                 ctypes.c_void_p(self.d_out.data_ptr()), shot_off, max_n)
            self.h_out = torch.empty(self.d_out.numel(), dtype=torch.int64, pin_memory=True)
            self.h_out.copy_(self.d_out, non_blocking=True)
            self.h_states = None
            if st_off:
                self.h_states = torch.empty(st_off, dtype=torch.complex64, pin_memory=True)
                self.h_states.copy_(self.d_states, non_blocking=True)
            self.event = torch.cuda.Event()
            self.event.record(self.stream)

    def done(self):
        return self.event is not None and self.event.query()

    def finish(self, now):
        self.event.synchronize()
        # one copy out of the pinned buffer; each request's samples are a view of it
        out = self.h_out.numpy().copy()
        for r, so, off in zip(self.reqs, self.state_off, self.shot_offs[:-1]):
            r.samples = out[int(off):int(off) + r.shots]
            if r.keep_state:
                r.state = self.h_states.numpy()[so:so + (1 << r.n_qubits)].copy()
            r.t_done = now
        self.d_states = self.h_states = None


# ---------------------------------------------------------------- HBM lane
class _PinnedPool:
    """Reusable pinned host buffers for the block-matrix uploads: a buffer
    goes back to the pool when the job that used it has finished on the
    device, so no page-locking happens on the launch path."""

    def __init__(self):
        self.free = []

    def acquire(self, nbytes):
        for i, b in enumerate(self.free):
            if b.numel() >= nbytes:
                return self.free.pop(i)
        return torch.empty(max(int(nbytes), 1 << 20), dtype=torch.uint8, pin_memory=True)

    def release(self, buf):
        if buf is not None:
            self.free.append(buf)


class _HbmJob:
    def __init__(self, req, sim, stream, pool=None, graphs=None):
        self.req = req
        self.reqs = [req]
        self.sim = sim
        self.stream = stream
        self.graphs = graphs          # _GraphCache, or None: no graph replay
        self.replayed = False
        self.pool = pool if pool is not None else _PinnedPool()
        self.staging = None
        self.event = None
        self.prepared = None      # future of prepare() when planned ahead
        self.blocks = None        # (block qubits int32 [B, 8], matrices complex64 [B, 32, 32])

    def prepare(self):
        """Host-only block planning for gate-payload requests (thread safe,
        touches no device state)."""
        r = self.req
        if r.gates is not None:
            pairs, us = r.gates
            g = len(pairs)
            qb = np.full((g, 8), -1, dtype=np.int32)
            qb[:, :2] = pairs
            refs, bqs, bm = C.plan_blocks_arrays(r.n_qubits, np.full(g, 2, np.int32), qb, us, k=5)
            assert (refs >= 0).all()          # two-qubit gates always fit a 5-qubit block
            self.blocks = (np.ascontiguousarray(bqs), np.ascontiguousarray(bm, dtype=np.complex64))
        return self

    def launch(self):
        r = self.req
        dev = self.sim.device
        if self.prepared is not None:
            self.prepared.result()
        elif r.gates is not None:
            self.prepare()
        sh = ctypes.c_void_p(self.stream.cuda_stream)
        with torch.cuda.stream(self.stream):
            if self.blocks is not None:
                self.state = torch.empty(1 << r.n_qubits, dtype=torch.complex64, device=dev)
                call("qsim_state_init", sh, ctypes.c_void_p(self.state.data_ptr()), r.n_qubits)
                bqs, bm = self.blocks
                raw = bm.view(np.uint8).reshape(-1)
                self.staging = self.pool.acquire(raw.size)
                self.staging[:raw.size].numpy()[:] = raw
                self.d_mats = torch.empty(raw.size, dtype=torch.uint8, device=dev)
                self.d_mats.copy_(self.staging[:raw.size], non_blocking=True)
                call("qsim_apply_blocks", sh, ctypes.c_void_p(self.state.data_ptr()), r.n_qubits,
                     ctypes.c_void_p(self.d_mats.data_ptr()), ctypes.c_void_p(bqs.ctypes.data),
                     len(bqs))
            else:
                cc = self.sim.compile(_circuit_of(r))
                g = self.graphs.get(cc, dev) if self.graphs is not None else None
                if g is not None:
                    self.state = g.replay()       # graph buffer, read before the next replay
                    self.replayed = True
                else:
                    self.state = torch.empty(1 << r.n_qubits, dtype=torch.complex64, device=dev)
                    cc.init_state(self.state)
                    cc.plan(cc.resolve_mode("blocks"), dev).run(self.state)
            from blackwell_kernels import sample_with_uniforms, uniforms_splitmix
            u = uniforms_splitmix(r.seed, r.shots, dev)
            self.d_out = sample_with_uniforms(self.state, u)
            self.h_out = torch.empty(r.shots, dtype=torch.int64, pin_memory=True)
            self.h_out.copy_(self.d_out, non_blocking=True)
            self.h_state = None
            if r.keep_state:
                self.h_state = torch.empty(1 << r.n_qubits, dtype=torch.complex64, pin_memory=True)
                self.h_state.copy_(self.state, non_blocking=True)
            self.event = torch.cuda.Event()
            self.event.record(self.stream)

    def done(self):
        return self.event is not None and self.event.query()

    def finish(self, now):
        self.event.synchronize()
        self.req.samples = self.h_out.numpy().copy()
        if self.h_state is not None:
            self.req.state = self.h_state.numpy().copy()
            self.h_state = None
        self.req.t_done = now
        self.state = self.d_mats = None
        self.pool.release(self.staging)
        self.staging = None


class _HbmBatch:
    """HBM-lane requests with gate payloads, run together: their block plans
    execute as batched block steps (one qsim_apply_block_batch launch per
    step applies block s of every request that has one), then each state is
    sampled with its own seeded device uniforms.  Block matrices and step
    records go to the device in one pinned copy.  Each request's arithmetic
    is that of qsim_apply_blocks, so its samples do not depend on the batch."""

    def __init__(self, jobs, stream, pool):
        self.jobs = jobs
        self.reqs = [j.req for j in jobs]
        self.stream = stream
        self.pool = pool
        self.staging = None
        self.event = None

    def launch(self):
        from blackwell_kernels import sample_with_uniforms, uniforms_splitmix
        from blackwell_kernels.cuquantum import (block_step_records, run_block_steps, run_sample_batch,
                                                 sample_records)
        for j in self.jobs:
            if j.prepared is not None:
                j.prepared.result()
            else:
                j.prepare()
        dev = self.jobs[0].sim.device
        sh = ctypes.c_void_p(self.stream.cuda_stream)
        ns = [r.n_qubits for r in self.reqs]
        mats = [j.blocks[1] for j in self.jobs]
        moff = np.concatenate([[0], np.cumsum([m.nbytes for m in mats])]).astype(np.int64)
        n_rec = sum(len(m) for m in mats)
        small = [i for i, n in enumerate(ns) if n <= SAMPLE_BATCH_MAX_QUBITS]
        rec_at = int(moff[-1])
        srec_at = rec_at + n_rec * BLOCK_REC_DTYPE.itemsize
        total = srec_at + len(small) * SAMPLE_REC_DTYPE.itemsize
        with torch.cuda.stream(self.stream):
            soff = np.concatenate([[0], np.cumsum([1 << n for n in ns])]).astype(np.int64)
            shot_off = np.concatenate([[0], np.cumsum([r.shots for r in self.reqs])]).astype(np.int64)
            self.slab = torch.empty(int(soff[-1]), dtype=torch.complex64, device=dev)
            self.d_buf = torch.empty(max(total, 1), dtype=torch.uint8, device=dev)
            self.d_out = torch.empty(max(int(shot_off[-1]), 1), dtype=torch.int64, device=dev)
            base, obase = self.slab.data_ptr(), self.d_out.data_ptr()
            recs, off = block_step_records([base + 8 * int(o) for o in soff[:-1]], ns,
                                           [self.d_buf.data_ptr() + int(o) for o in moff[:-1]],
                                           [j.blocks[0] for j in self.jobs])
            srecs, leaves = sample_records([base + 8 * int(soff[i]) for i in small], [ns[i] for i in small],
                                           [obase + 8 * int(shot_off[i]) for i in small],
                                           [self.reqs[i].seed for i in small],
                                           [self.reqs[i].shots for i in small])
            self.staging = self.pool.acquire(total)
            host = self.staging.numpy()
            for m, o in zip(mats, moff[:-1]):
                host[int(o):int(o) + m.nbytes] = m.view(np.uint8).reshape(-1)
            host[rec_at:srec_at] = recs.view(np.uint8)
            host[srec_at:total] = srecs.view(np.uint8)
            self.d_buf[:total].copy_(self.staging[:total], non_blocking=True)
            self.recs, self.srecs = recs, srecs     # validated host copies of the records
            for n, o in zip(ns, soff[:-1]):
                call("qsim_state_init", sh, ctypes.c_void_p(base + 8 * int(o)), n)
            run_block_steps(sh, recs, self.d_buf[rec_at:], off)
            if small:
                run_sample_batch(sh, srecs, self.d_buf[srec_at:], leaves, dev)
            for i, r in enumerate(self.reqs):
                if r.n_qubits <= SAMPLE_BATCH_MAX_QUBITS:
                    continue
                st = self.slab[int(soff[i]):int(soff[i]) + (1 << r.n_qubits)]
                d = sample_with_uniforms(st, uniforms_splitmix(r.seed, r.shots, dev))
                self.d_out[int(shot_off[i]):int(shot_off[i]) + r.shots].copy_(d)
            self.h_out = torch.empty(self.d_out.numel(), dtype=torch.int64, pin_memory=True)
            self.h_out.copy_(self.d_out, non_blocking=True)
            self.h_states = {}
            for i, r in enumerate(self.reqs):
                if r.keep_state:
                    h = torch.empty(1 << r.n_qubits, dtype=torch.complex64, pin_memory=True)
                    h.copy_(self.slab[int(soff[i]):int(soff[i]) + (1 << r.n_qubits)], non_blocking=True)
                    self.h_states[i] = h
            self.shot_off = shot_off
            self.event = torch.cuda.Event()
            self.event.record(self.stream)

    def done(self):
        return self.event is not None and self.event.query()

    def finish(self, now):
        self.event.synchronize()
        out = self.h_out.numpy().copy()
        for i, (r, so) in enumerate(zip(self.reqs, self.shot_off[:-1])):
            r.samples = out[int(so):int(so) + r.shots]
            if i in self.h_states:
                r.state = self.h_states[i].numpy().copy()
            r.t_done = now
        self.h_states = {}
        self.slab = self.d_buf = self.d_out = None
        self.pool.release(self.staging)
        self.staging = None


class _GraphCache:
    """HIP graphs of compiled circuits seen more than once on the HBM lane
    (at most `max_graphs`, least recently used dropped; states of at most
    `max_qubits`).  A circuit's first run is eager and only counted, so a
    stream of distinct circuits never pays for a capture."""

    def __init__(self, max_graphs=8, max_qubits=24):
        self.max_graphs = int(max_graphs)
        self.max_qubits = int(max_qubits)
        # run counts die with their circuits, so a long stream of distinct
        # circuits leaves nothing behind
        self.seen = weakref.WeakKeyDictionary()
        # id(cc) -> graph; the graph holds cc, so the id stays cc's while cached
        self.graphs = collections.OrderedDict()
        self.captures = 0
        self.replays = 0

    def get(self, cc, device):
        if cc.n_qubits > self.max_qubits or cc.has_stochastic or self.max_graphs <= 0:
            return None
        self.seen[cc] = self.seen.get(cc, 0) + 1
        key = id(cc)
        g = self.graphs.get(key)
        if g is not None and g.cc is not cc:
            g = None
        if g is None:
            if self.seen[cc] < 2:
                return None
            mode = cc.resolve_mode("blocks")
            g = cc.graphed(mode, device)
            if g is None:
                return None
            self.captures += 1
            self.graphs[key] = g
            if len(self.graphs) > self.max_graphs:
                # a dropped graph's private memory goes back to the allocator:
                # let its replays finish first
                if torch.device(device).type == "cuda":
                    torch.cuda.synchronize(device)
                while len(self.graphs) > self.max_graphs:
                    _, old = self.graphs.popitem(last=False)
                    old.cc._graphs.clear()
        else:
            self.graphs.move_to_end(key)
        self.replays += 1
        return g


# ---------------------------------------------------------------- server
class Server:
    """Continuous-batching server.

    budget_bytes     HBM budget for resident states (None: 80 % of free memory)
    max_batch        most requests per LDS launch
    max_hbm_jobs     HBM-lane launch units (batches or single circuits) in flight
    max_hbm_batch    most HBM-lane circuits per batch
    max_hbm_batch_bytes  most state bytes per HBM-lane batch (a larger
                     circuit runs alone); the default is the 256 MB MALL of
                     MI355X, so a batch stays cache-resident across its block
                     steps
    planner_threads  host threads that build HBM-lane block plans ahead
    plan_ahead       most queued HBM-lane circuits planned ahead of admission
    graph_repeats    replay repeated HBM-lane circuits as HIP graphs (at most
                     this many graphs kept; 0 turns it off)
    """

    def __init__(self, sim=None, budget_bytes=None, max_batch=2048, max_hbm_jobs=4,
                 device="cuda:0", lds_factory=None, hbm_factory=None, planner_threads=8,
                 plan_ahead=256, max_hbm_batch=64, hbm_batch_factory=None,
                 max_hbm_batch_bytes=HBM_BATCH_BYTES, graph_repeats=8):
        """lds_factory(reqs) / hbm_factory(req) build lane jobs with launch(),
        done() and finish(now); hbm_batch_factory(jobs) runs several HBM-lane
        jobs as one unit (None: one job per unit).  The defaults run on the
        GPU; tests inject host-only fakes to drive the scheduling logic."""
        fake = lds_factory is not None and hbm_factory is not None
        if sim is None and not fake:
            from .simulator import BlackwellSimulator
            sim = BlackwellSimulator(device)
        self.sim = sim
        self.device = sim.device if sim is not None else None
        if budget_bytes is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            budget_bytes = int(0.8 * free)
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.max_hbm_jobs = int(max_hbm_jobs)
        self.planner_threads = int(planner_threads)
        self.plan_ahead = int(plan_ahead)
        self.max_hbm_batch = int(max_hbm_batch)
        self.max_hbm_batch_bytes = int(max_hbm_batch_bytes)
        self.hbm_batch_factory = hbm_batch_factory
        self.in_use = 0
        self.peak_in_use = 0
        if fake:
            self.lds_factory, self.hbm_factory = lds_factory, hbm_factory
        else:
            self.lds_stream = torch.cuda.Stream(self.device)
            self.hbm_stream = torch.cuda.Stream(self.device)
            self.lds_factory = lambda reqs: _LdsBatch(reqs, self.device, self.lds_stream)
            self.pinned = _PinnedPool()
            self.graphs = _GraphCache(graph_repeats)
            self.hbm_factory = lambda req: _HbmJob(req, self.sim, self.hbm_stream, self.pinned,
                                                   self.graphs)
            if hbm_batch_factory is None:
                self.hbm_batch_factory = lambda jobs: _HbmBatch(jobs, self.hbm_stream, self.pinned)
        self.rejected = []
        self.launches = 0

    def lane_of(self, req):
        return "lds" if req.n_qubits <= LDS_MAX_QUBITS else "hbm"

    def admit_bytes(self, req):
        # LDS-lane circuits never hold an HBM state; their sample buffer is tiny
        return 8 * req.shots if self.lane_of(req) == "lds" else req.state_bytes() + 8 * req.shots + (1 << 20)

    def serve(self, requests, realtime=True, timeout_s=600.0):
        """Serve requests (arrival times relative to the start).  Returns
        stats; every request gets .samples, .t_admit, .t_done (seconds from
        the start)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        self.rejected = [r for r in reqs if self.admit_bytes(r) > self.budget]
        rej = {id(r) for r in self.rejected}
        pending = collections.deque(r for r in reqs if id(r) not in rej)
        pool = None
        if self.planner_threads > 0:
            from concurrent.futures import ThreadPoolExecutor
            pool = ThreadPoolExecutor(self.planner_threads, thread_name_prefix="qsim-plan")
        try:
            return self._serve(reqs, pending, pool, realtime, timeout_s)
        finally:
            if pool is not None:
                pool.shutdown(wait=True, cancel_futures=True)

    def _serve(self, reqs, pending, pool, realtime, timeout_s):
        queue_lds, queue_hbm = collections.deque(), collections.deque()
        lds_inflight, hbm_inflight = [], []
        n_planned = 0             # queue_hbm[:n_planned] have plans under way
        t0 = time.perf_counter()
        clock = (lambda: time.perf_counter() - t0) if realtime else None
        done = 0
        total = len(pending)
        while done < total:
            now = clock() if realtime else float("inf")
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                r.lane = self.lane_of(r)
                if r.lane == "lds":
                    queue_lds.append(r)
                else:
                    queue_hbm.append(self.hbm_factory(r))
            # plan queued HBM-lane circuits ahead, on the host threads
            while pool is not None and n_planned < min(len(queue_hbm), self.plan_ahead):
                j = queue_hbm[n_planned]
                if hasattr(j, "prepare"):
                    j.prepared = pool.submit(j.prepare)
                n_planned += 1
            # retire finished work, freeing budget
            for b in [b for b in lds_inflight if b.done()]:
                b.finish(clock() if realtime else time.perf_counter() - t0)
                lds_inflight.remove(b)
                for r in b.reqs:
                    self.in_use -= self.admit_bytes(r)
                done += len(b.reqs)
            for j in [j for j in hbm_inflight if j.done()]:
                j.finish(clock() if realtime else time.perf_counter() - t0)
                hbm_inflight.remove(j)
                for r in j.reqs:
                    self.in_use -= self.admit_bytes(r)
                done += len(j.reqs)
            launched = False
            # LDS lane: one launch in flight at a time, everything queued goes in
            if queue_lds and len(lds_inflight) < 2:
                batch = []
                while queue_lds and len(batch) < self.max_batch:
                    r = queue_lds[0]
                    need = self.admit_bytes(r)
                    if self.in_use + need > self.budget:
                        break
                    queue_lds.popleft()
                    self.in_use += need
                    r.t_admit = time.perf_counter() - t0
                    batch.append(r)
                if batch:
                    b = self.lds_factory(batch)
                    b.launch()
                    self.launches += 1
                    self.peak_in_use = max(self.peak_in_use, self.in_use)
                    lds_inflight.append(b)
                    launched = True
            # HBM lane: ready jobs (planned, within budget) go out together;
            # a job without a gate payload runs alone
            while queue_hbm and len(hbm_inflight) < self.max_hbm_jobs:
                group, gbytes = [], 0
                limit = self.max_hbm_batch if self.hbm_batch_factory is not None else 1
                while queue_hbm and len(group) < limit:
                    j = queue_hbm[0]
                    fut = getattr(j, "prepared", None)
                    if fut is not None and not fut.done():
                        break             # its plan is still being built
                    batchable = getattr(j.reqs[0], "gates", None) is not None
                    if group and not batchable:
                        break
                    r = j.reqs[0]
                    if group and gbytes + r.state_bytes() > self.max_hbm_batch_bytes:
                        break
                    need = self.admit_bytes(r)
                    if self.in_use + need > self.budget:
                        break
                    gbytes += r.state_bytes()
                    queue_hbm.popleft()
                    n_planned = max(0, n_planned - 1)
                    self.in_use += need
                    r.t_admit = time.perf_counter() - t0
                    group.append(j)
                    if not batchable:
                        break
                if not group:
                    break
                unit = (self.hbm_batch_factory(group)
                        if self.hbm_batch_factory is not None and
                        getattr(group[0].reqs[0], "gates", None) is not None else group[0])
                unit.launch()
                self.launches += 1
                self.peak_in_use = max(self.peak_in_use, self.in_use)
                hbm_inflight.append(unit)
                launched = True
            if time.perf_counter() - t0 > timeout_s:
                raise TimeoutError(f"serving did not finish within {timeout_s} s ({done}/{total} done)")
            if not launched:
                if lds_inflight or hbm_inflight or queue_hbm:
                    time.sleep(0.0002)
                elif pending and realtime:
                    time.sleep(max(0.0, min(0.001, pending[0].arrival - clock())))
        elapsed = time.perf_counter() - t0
        return summarize(reqs, elapsed, self.rejected)


def summarize(reqs, elapsed, rejected=()):
    served = [r for r in reqs if r.t_done is not None]
    lat = np.array([r.t_done - r.arrival for r in served]) if served else np.zeros(1)
    first = min((r.arrival for r in served), default=0.0)
    last = max((r.t_done for r in served), default=elapsed)
    span = max(last - first, 1e-9)
    return {
        "served": len(served),
        "rejected": len(rejected),
        "circuits_per_s": len(served) / span,
        "span_s": span,
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(np.mean(lat)),
    }


def _plain_gates(state, req, stream):
    """One qsim_apply_gate_2q_h launch per gate of a gate-payload request."""
    pairs, us = req.gates
    g = len(pairs)
    km = np.array(us, dtype=np.complex128)
    # the 2Q kernel's row index is 2*bit(max qubit) + bit(min qubit)
    swap = pairs[:, 0] > pairs[:, 1]
    perm = [0, 2, 1, 3]
    km[swap] = km[swap][:, perm][:, :, perm]
    d = np.ascontiguousarray(km).view(np.float64).reshape(g, 32)
    fn = _lib.load().qsim_apply_gate_2q_h
    sp = ctypes.c_void_p(state.data_ptr())
    base = d.ctypes.data
    pd = ctypes.POINTER(ctypes.c_double)
    n = req.n_qubits
    for i in range(g):
        rc = fn(stream, sp, n, ctypes.cast(base + 256 * i, pd), int(pairs[i, 0]), int(pairs[i, 1]))
        if rc:
            raise _lib.QsimError("qsim_apply_gate_2q_h: " + _lib.load().qsim_last_error().decode())


def serve_plain(requests, sim, realtime=True):
    """The plain baseline: one circuit at a time, in arrival order, through
    the per-gate kernels (one launch per gate, matrices by value), then
    qsim_sample with the same uniforms, synchronising on each result."""
    from blackwell_kernels import sample_with_uniforms, uniforms_splitmix
    reqs = sorted(requests, key=lambda r: r.arrival)
    dev = sim.device
    t0 = time.perf_counter()
    for r in reqs:
        if realtime:
            wait = r.arrival - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
        r.t_admit = time.perf_counter() - t0
        sh = _lib.stream_handle(dev)
        st = torch.empty(1 << r.n_qubits, dtype=torch.complex64, device=dev)
        call("qsim_state_init", sh, ctypes.c_void_p(st.data_ptr()), r.n_qubits)
        if r.gates is not None:
            _plain_gates(st, r, sh)
        else:
            cc = sim.compile(_circuit_of(r))
            cc.plan("plain", dev).run(st)
        u = uniforms_splitmix(r.seed, r.shots, dev)
        r.samples = sample_with_uniforms(st, u).cpu().numpy()
        if r.keep_state:
            r.state = st.cpu().numpy()
        r.t_done = time.perf_counter() - t0
        r.lane = "plain"
    return summarize(reqs, time.perf_counter() - t0)
