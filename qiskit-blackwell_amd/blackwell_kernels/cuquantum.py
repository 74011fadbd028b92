"""Kernel API with the names, arguments and error behaviour of the
reference's blackwell_kernels module (bwk/python/blackwell_kernels/cuquantum.py),
backed by the gfx950 library libqsim_gfx950.so through its C ABI.

States are 1-D contiguous complex64 tensors on a ROCm device, modified in
place and returned.  Every call launches on torch's current stream of the
state's device.  Argument errors raise RuntimeError (as the reference's
TORCH_CHECK does) or ValueError where the reference raises it.
"""
import math

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_handle


# ------------------------------------------------------------------ checks
def _check_state(state, name="state"):
    if not isinstance(state, torch.Tensor):
        raise RuntimeError(f"{name} must be a torch.Tensor")
    if not state.is_cuda:
        raise RuntimeError(f"{name} must be CUDA")
    if state.dtype != torch.complex64:
        raise RuntimeError(f"{name} must be complex64")
    if not state.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def n_qubits_of(state):
    _check_state(state)
    n = state.numel()
    q = n.bit_length() - 1
    if n < 2 or (1 << q) != n:
        raise RuntimeError(f"state size {n} is not a power of two >= 2")
    return q


def _batched_dims(states):
    _check_state(states, "states")
    if states.dim() != 2:
        raise RuntimeError("states must be [N, 2^n]")
    per = states.shape[1]
    q = per.bit_length() - 1
    if per < 2 or (1 << q) != per:
        raise RuntimeError(f"state size {per} is not a power of two >= 2")
    return states.shape[0], q


def _dev_matrix(gate, shape, name, device):
    """Device complex64 contiguous tensor of the given shape (uploads host data)."""
    if isinstance(gate, torch.Tensor):
        if tuple(gate.shape) != shape and gate.numel() != int(np.prod(shape)):
            raise RuntimeError(f"{name} must be {'x'.join(map(str, shape))}")
        if gate.dtype != torch.complex64:
            if not gate.is_complex():
                raise RuntimeError(f"{name} must be complex64")
            gate = gate.to(torch.complex64)
        if gate.device != device:
            gate = gate.to(device)
        return gate.contiguous()
    arr = np.asarray(gate, dtype=np.complex64)
    if arr.size != int(np.prod(shape)):
        raise RuntimeError(f"{name} must be {'x'.join(map(str, shape))}")
    return torch.from_numpy(arr.reshape(shape).copy()).to(device)


# ------------------------------------------------------------------ gates
def apply_gate(state, gate, target_qubit):
    """2x2 gate on target_qubit (cuquantum.py:40-51)."""
    n = n_qubits_of(state)
    g = _dev_matrix(gate, (2, 2), "gate", state.device)
    call("qsim_apply_gate", stream_handle(state.device), ptr(state), n, ptr(g), int(target_qubit))
    return state


def apply_gate_2q(state, gate, qubit0, qubit1):
    """4x4 gate; row index 2*bit(max) + bit(min) (cuquantum.py:54-66)."""
    n = n_qubits_of(state)
    g = _dev_matrix(gate, (4, 4), "gate", state.device)
    call("qsim_apply_gate_2q", stream_handle(state.device), ptr(state), n, ptr(g),
         int(qubit0), int(qubit1))
    return state


def apply_gates_fused(state, gates, targets):
    """N 2x2 gates in order through LDS tiles (cuquantum.py:69-80)."""
    n = n_qubits_of(state)
    tl = [int(t) for t in (targets.tolist() if isinstance(targets, torch.Tensor) else targets)]
    g = _dev_matrix(gates, (len(tl), 2, 2), "gates", state.device)
    call("qsim_apply_gates_fused", stream_handle(state.device), ptr(state), n, ptr(g),
         _lib.i32_array(tl), len(tl))
    return state


def apply_gates_reg_tiled(state, gates, targets):
    """N 2x2 gates with targets < 8, register-resident (cuquantum.py:83-98).

    Targets may live on the device (as in the reference); a target outside
    0..min(n,8)-1 is skipped by the kernel.  Host targets are checked here.
    """
    n = n_qubits_of(state)
    if isinstance(targets, torch.Tensor) and targets.is_cuda:
        t = targets.to(torch.int32).contiguous()
    else:
        tl = [int(x) for x in (targets.tolist() if isinstance(targets, torch.Tensor) else targets)]
        lim = min(n, 8)
        for x in tl:
            if x < 0 or x >= lim:
                raise RuntimeError(f"reg-tiled targets must be in 0..{lim - 1} (got {x})")
        t = torch.tensor(tl, dtype=torch.int32, device=state.device)
    g = _dev_matrix(gates, (t.numel(), 2, 2), "gates", state.device)
    call("qsim_apply_gates_reg_tiled", stream_handle(state.device), ptr(state), n, ptr(g),
         ptr(t), t.numel())
    return state


def apply_diagonal(state, diag, target_qubit):
    """state[i] *= diag[bit(i, t)] (cuquantum.py:174-185)."""
    n = n_qubits_of(state)
    d = _dev_matrix(diag, (2,), "diag", state.device)
    call("qsim_apply_diagonal", stream_handle(state.device), ptr(state), n, ptr(d), int(target_qubit))
    return state


def apply_diagonal_phase(state, phase, target_qubit):
    """Multiply the bit=1 half by a complex phase (cuquantum.py:188-195)."""
    n = n_qubits_of(state)
    phase = complex(phase)
    call("qsim_apply_diagonal_phase", stream_handle(state.device), ptr(state), n,
         phase.real, phase.imag, int(target_qubit))
    return state


def apply_diagonal_2q(state, diag, qubit0, qubit1):
    """state[i] *= diag[2*bit(max) + bit(min)] (cuquantum.py:198-210)."""
    n = n_qubits_of(state)
    d = _dev_matrix(diag, (4,), "diag", state.device)
    call("qsim_apply_diagonal_2q", stream_handle(state.device), ptr(state), n, ptr(d),
         int(qubit0), int(qubit1))
    return state


def _controls_list(controls):
    if isinstance(controls, torch.Tensor):
        return [int(c) for c in controls.to(torch.int64).cpu().tolist()]
    return [int(c) for c in controls]


def apply_mcgate(state, gate, controls, target_qubit):
    """Controlled 2x2 gate, 0-7 controls (cuquantum.py:297-317)."""
    n = n_qubits_of(state)
    cl = _controls_list(controls)
    g = _dev_matrix(gate, (2, 2), "gate", state.device)
    call("qsim_apply_mcgate", stream_handle(state.device), ptr(state), n, ptr(g),
         _lib.i64_array(cl), len(cl), int(target_qubit))
    return state


def apply_mc2qgate(state, gate, controls, target0, target1):
    """Controlled 4x4 gate, 0-7 controls (cuquantum.py:320-337)."""
    n = n_qubits_of(state)
    cl = _controls_list(controls)
    g = _dev_matrix(gate, (4, 4), "gate", state.device)
    call("qsim_apply_mc2qgate", stream_handle(state.device), ptr(state), n, ptr(g),
         _lib.i64_array(cl), len(cl), int(target0), int(target1))
    return state


def apply_mcx(state, controls, target_qubit):
    """Bit-exact multi-controlled X (X, CNOT, Toffoli, MCX)."""
    n = n_qubits_of(state)
    cl = _controls_list(controls)
    call("qsim_apply_mcx", stream_handle(state.device), ptr(state), n, _lib.i64_array(cl), len(cl),
         int(target_qubit))
    return state


def apply_mcswap(state, controls, qubit0, qubit1):
    """Bit-exact multi-controlled SWAP (SWAP, Fredkin)."""
    n = n_qubits_of(state)
    cl = _controls_list(controls)
    call("qsim_apply_mcswap", stream_handle(state.device), ptr(state), n, _lib.i64_array(cl),
         len(cl), int(qubit0), int(qubit1))
    return state


def apply_block(state, block, qubits):
    """Dense 2^k x 2^k block on `qubits` (row bit j <-> qubits[j]).

    k = 5 with n >= 10, k = 4 with n >= 11 and k = 6 with n >= 13 run on FP32
    MFMA (qsim_apply_block_mfma); other blocks of up to 5 qubits go through a
    one-op LDS tile program.
    """
    n = n_qubits_of(state)
    qs = [int(q) for q in qubits]
    k = len(qs)
    b = _dev_matrix(block, (1 << k, 1 << k), "block", state.device)
    if (k == 5 and n >= 10) or (k == 4 and n >= 11) or (k == 6 and n >= 13):
        call("qsim_apply_block_mfma", stream_handle(state.device), ptr(state), n, ptr(b),
             _lib.int_array(qs), k)
        return state
    if k > 5:
        raise RuntimeError("blocks of 6 qubits need n >= 13; wider blocks are not supported")
    mask = 0
    for q in qs:
        mask |= 1 << q
    tile = _tile_mask(n, mask)
    ops = np.zeros(1, dtype=_lib.OP_DTYPE)
    ops["kind"] = _lib.OP_MK
    ops["nq"] = k
    for j, q in enumerate(qs):
        ops["q"][0, j] = bin(tile & ((1 << q) - 1)).count("1")
    run_tile_program(state, tile, ops, b.reshape(-1))
    return state


def block_step_records(state_ptrs, n_qubits, mat_ptrs, qubits):
    """Records of batched block steps (numpy BLOCK_REC_DTYPE) and the
    offsets of each step in them.  state_ptrs[i] / n_qubits[i]: instance i's
    state; mat_ptrs[i]: device address of its complex64 [B_i, 32, 32]
    blocks; qubits[i]: host int array [B_i, >= 5].  Step s holds block s of
    every instance with s < B_i, in instance order."""
    counts = np.array([len(q) for q in qubits], dtype=np.int64)
    steps = int(counts.max()) if counts.size else 0
    inst = np.repeat(np.arange(counts.size), counts)
    step = np.concatenate([np.arange(c) for c in counts]) if counts.size else np.zeros(0, np.int64)
    order = np.lexsort((inst, step))             # by step, then instance
    inst, step = inst[order], step[order]
    qcat = np.concatenate([np.asarray(q, dtype=np.int64)[:, :5] for q in qubits]) \
        if counts.size else np.zeros((0, 5), np.int64)
    qoff = np.concatenate([[0], np.cumsum(counts)[:-1]]) if counts.size else np.zeros(0, np.int64)
    nq = np.asarray(n_qubits, dtype=np.int64)
    recs = np.zeros(inst.size, dtype=_lib.BLOCK_REC_DTYPE)
    recs["state"] = np.asarray(state_ptrs, dtype=np.uint64)[inst]
    recs["mat"] = np.asarray(mat_ptrs, dtype=np.uint64)[inst] + step.astype(np.uint64) * np.uint64(32 * 32 * 8)
    recs["n_qubits"] = nq[inst]
    recs["q"] = qcat[qoff[inst] + step]
    sets = np.left_shift(np.int64(1), nq[inst] - 10)
    off = np.zeros(steps + 1, dtype=np.int64)
    off[1:] = np.cumsum(np.bincount(step, minlength=steps))
    csum = np.cumsum(sets) - sets                 # exclusive prefix over all records
    recs["set_begin"] = (csum - csum[off[step]]).astype(np.uint64)
    return recs, off


def run_block_steps(stream, recs, recs_dev, off):
    """One qsim_apply_block_batch launch per step (recs: host records,
    recs_dev: a device tensor holding the same bytes)."""
    base_h = recs.ctypes.data
    base_d = recs_dev.data_ptr()
    size = _lib.BLOCK_REC_DTYPE.itemsize
    for s in range(len(off) - 1):
        a, b = int(off[s]), int(off[s + 1])
        call("qsim_apply_block_batch", stream, _lib.ctypes.c_void_p(base_h + a * size),
             _lib.ctypes.c_void_p(base_d + a * size), b - a)


def apply_blocks_batched(states, mats, qubits):
    """Apply each instance's 5-qubit block list (mats[i]: complex64 device
    [B_i, 32, 32], qubits[i]: host int [B_i, >= 5]) to states[i] (10..32
    qubits), one launch per block step for all instances together."""
    if not states:
        return states
    for st in states:
        _check_state(st)
    dev = states[0].device
    ns = [n_qubits_of(st) for st in states]
    for m, q in zip(mats, qubits):
        if m.dtype != torch.complex64 or not m.is_cuda or not m.is_contiguous() or \
                tuple(m.shape[1:]) != (32, 32) or m.shape[0] != len(q):
            raise RuntimeError("mats[i] must be a contiguous CUDA complex64 [B_i, 32, 32] tensor")
    recs, off = block_step_records([st.data_ptr() for st in states], ns,
                                   [m.data_ptr() for m in mats], qubits)
    if recs.size == 0:
        return states
    recs_dev = torch.from_numpy(recs.view(np.uint8).copy()).to(dev)
    run_block_steps(stream_handle(dev), recs, recs_dev, off)
    return states


def _tile_mask(n, must, bits=None):
    """Tile mask holding `must` plus the lowest other qubits (qubit 0 always)."""
    bits = min(n, bits or 10)
    mask = must | 1
    q = 0
    while bin(mask).count("1") < bits and q < n:
        mask |= 1 << q
        q += 1
    return mask


def run_tile_program(state, tile_mask, ops, coefs):
    """Apply an op program (numpy OP_DTYPE array, tile-local qubits) to every
    tile of `state`; coefs is a complex64 tensor/array of coefficients."""
    n = n_qubits_of(state)
    ops_d = torch.from_numpy(np.ascontiguousarray(ops).view(np.uint8).copy()).to(state.device)
    c = _dev_matrix(coefs, (int(np.prod(np.shape(coefs))),), "coefs", state.device)
    call("qsim_tile_program", stream_handle(state.device), ptr(state), n, int(tile_mask),
         ptr(ops_d), len(ops), ptr(c))
    return state


# ------------------------------------------------------------------ init
def state_init(n_qubits, device="cuda"):
    """|0...0> as a new complex64 tensor [2^n] (cuquantum.py:213-219)."""
    n = int(n_qubits)
    if n < 1 or n > 34:
        raise RuntimeError(f"n_qubits={n} outside 1..34")
    st = torch.empty(1 << n, dtype=torch.complex64, device=device)
    call("qsim_state_init", stream_handle(st.device), ptr(st), n)
    return st


def state_init_batched(n_states, n_qubits, device="cuda"):
    """N states |0...0> as [N, 2^n] (cuquantum.py:444-450)."""
    n = int(n_qubits)
    st = torch.empty(int(n_states), 1 << n, dtype=torch.complex64, device=device)
    call("qsim_state_init_batched", stream_handle(st.device), ptr(st), int(n_states), n)
    return st


def apply_gate_batched(states, gate, target_qubit):
    """Same 2x2 gate on each of N states (cuquantum.py:414-425)."""
    ns, n = _batched_dims(states)
    g = _dev_matrix(gate, (2, 2), "gate", states.device)
    call("qsim_apply_gate_batched", stream_handle(states.device), ptr(states), ns, n, ptr(g),
         int(target_qubit))
    return states


def apply_gates_batched_fused(states, gates, targets):
    """Same gate list (targets < 8) on each of N states (cuquantum.py:428-441)."""
    ns, n = _batched_dims(states)
    if isinstance(targets, torch.Tensor) and targets.is_cuda:
        t = targets.to(torch.int32).contiguous()
    else:
        tl = [int(x) for x in (targets.tolist() if isinstance(targets, torch.Tensor) else targets)]
        lim = min(n, 8)
        for x in tl:
            if x < 0 or x >= lim:
                raise RuntimeError(f"batched fused targets must be in 0..{lim - 1} (got {x})")
        t = torch.tensor(tl, dtype=torch.int32, device=states.device)
    g = _dev_matrix(gates, (t.numel(), 2, 2), "gates", states.device)
    call("qsim_apply_gates_batched_fused", stream_handle(states.device), ptr(states), ns, n,
         ptr(g), ptr(t), t.numel())
    return states


# ------------------------------------------------------------------ reductions
def norm2(state):
    """||psi||^2 in FP64 (deterministic)."""
    n = n_qubits_of(state)
    ws = _lib.workspace("reduce", _lib.load().qsim_reduce_workspace_bytes(n), state.device)
    out = torch.empty(1, dtype=torch.float64, device=state.device)
    call("qsim_norm2", stream_handle(state.device), ptr(state), n, ptr(ws), ptr(out))
    return float(out.item())


def measure_qubit(state, target_qubit, random_val):
    """Measure and collapse; returns 0 or 1 (cuquantum.py:257-268)."""
    n = n_qubits_of(state)
    ws = _lib.workspace("reduce", _lib.load().qsim_reduce_workspace_bytes(n), state.device)
    out = _lib.ctypes.c_int(-1)
    call("qsim_measure_qubit", stream_handle(state.device), ptr(state), n, int(target_qubit),
         float(random_val), ptr(ws), _lib.ctypes.byref(out))
    return int(out.value)


def measure_probs(state):
    """|state|^2 as float32 [2^n] (cuquantum.py:271-280)."""
    n = n_qubits_of(state)
    probs = torch.empty(1 << n, dtype=torch.float32, device=state.device)
    call("qsim_measure_probs", stream_handle(state.device), ptr(state), n, ptr(probs))
    return probs


def sample_with_uniforms(state, uniforms):
    """Smallest i with cdf(i) > u * ||psi||^2 for each uniform (FP64)."""
    n = n_qubits_of(state)
    u = uniforms
    if not isinstance(u, torch.Tensor):
        u = torch.as_tensor(np.asarray(u, dtype=np.float64))
    u = u.to(device=state.device, dtype=torch.float64).contiguous()
    shots = u.numel()
    out = torch.empty(shots, dtype=torch.int64, device=state.device)
    if shots == 0:
        return out
    ws = _lib.workspace("sample", _lib.load().qsim_sample_workspace_bytes(n), state.device)
    call("qsim_sample", stream_handle(state.device), ptr(state), n, ptr(u), shots, ptr(ws), ptr(out))
    return out


def uniforms_splitmix(seed, count, device="cuda"):
    """Device float64 [count] of the seeded SplitMix64 uniforms the LDS
    serving lane draws in-kernel (u_k from seed + (k+1) * golden gamma)."""
    out = torch.empty(int(count), dtype=torch.float64, device=device)
    if count:
        call("qsim_uniforms_splitmix", stream_handle(out.device), int(seed) & 0xFFFFFFFFFFFFFFFF,
             int(count), ptr(out))
    return out


def sample_records(state_ptrs, n_qubits, out_ptrs, seeds, shots):
    """Records (numpy SAMPLE_REC_DTYPE) of one batched sampling call and the
    number of leaves its workspace must hold."""
    nq = np.asarray(n_qubits, dtype=np.int64)
    sh = np.asarray(shots, dtype=np.int64)
    recs = np.zeros(nq.size, dtype=_lib.SAMPLE_REC_DTYPE)
    recs["state"] = np.asarray(state_ptrs, dtype=np.uint64)
    recs["out"] = np.asarray(out_ptrs, dtype=np.uint64)
    recs["seed"] = np.asarray([int(x) & 0xFFFFFFFFFFFFFFFF for x in seeds], dtype=np.uint64)
    leaves = np.left_shift(np.int64(1), nq - 8)
    recs["leaf_off"] = (np.cumsum(leaves) - leaves).astype(np.uint64)
    recs["shot_off"] = (np.cumsum(sh) - sh).astype(np.uint64)
    recs["n_qubits"] = nq
    recs["shots"] = sh.astype(np.uint32)
    return recs, int(leaves.sum())


def run_sample_batch(stream, recs, recs_dev, n_leaves, device):
    """qsim_sample_batch over host records `recs` (device copy recs_dev)."""
    ws = _lib.workspace("sample_batch", _lib.load().qsim_sample_batch_workspace_bytes(n_leaves, len(recs)),
                        device)
    call("qsim_sample_batch", stream, _lib.ctypes.c_void_p(recs.ctypes.data),
         _lib.ctypes.c_void_p(recs_dev.data_ptr()), len(recs), ptr(ws))


def sample_batch(states, seeds, shots):
    """int64 device tensors of samples for each state (8..20 qubits), drawn
    with the seeded SplitMix64 uniforms of uniforms_splitmix(seed, shots):
    bitwise sample_with_uniforms(state, uniforms_splitmix(seed, shots)) for
    every state, in three launches for all of them."""
    if not states:
        return []
    for st in states:
        _check_state(st)
    dev = states[0].device
    ns = [n_qubits_of(st) for st in states]
    outs = [torch.empty(int(k), dtype=torch.int64, device=dev) for k in shots]
    recs, leaves = sample_records([st.data_ptr() for st in states], ns,
                                  [o.data_ptr() for o in outs], seeds, shots)
    recs_dev = torch.from_numpy(recs.view(np.uint8).copy()).to(dev)
    run_sample_batch(stream_handle(dev), recs, recs_dev, leaves, dev)
    return outs


def sample(state, n_shots):
    """n_shots indices from |state|^2 (cuquantum.py:283-294); uniforms from
    torch's generator on the state's device, in FP64."""
    u = torch.rand(int(n_shots), dtype=torch.float64, device=state.device)
    return sample_with_uniforms(state, u)


def renormalize(state):
    """Scale to unit norm, FP64 norm and scale (cuquantum.py:501-513)."""
    n = n_qubits_of(state)
    ws = _lib.workspace("reduce", _lib.load().qsim_reduce_workspace_bytes(n), state.device)
    call("qsim_renormalize", stream_handle(state.device), ptr(state), n, ptr(ws))
    return state


# ------------------------------------------------------------------ Chebyshev
def _c64(t, name):
    if t.dtype != torch.complex64 or not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA complex64 tensor")
    return t


def spmv_csr(row_ptr, col_idx, values, x, y, alpha_re=1.0, alpha_im=0.0, beta_re=0.0, beta_im=0.0):
    """y = alpha*A*x + beta*y, CSR int32/complex64 (cuquantum.py:518-531)."""
    if not (row_ptr.is_cuda and col_idx.is_cuda and values.is_cuda):
        raise RuntimeError("CSR arrays must be CUDA tensors")
    if row_ptr.dtype != torch.int32 or col_idx.dtype != torch.int32:
        raise RuntimeError("row_ptr and col_idx must be int32")
    _c64(values, "values"); _c64(x, "x"); _c64(y, "y")
    call("qsim_spmv_csr", stream_handle(x.device), ptr(row_ptr), ptr(col_idx), ptr(values),
         row_ptr.numel() - 1, ptr(x), ptr(y), float(alpha_re), float(alpha_im),
         float(beta_re), float(beta_im))


def chebyshev_step(spmv_result, t_prev, t_next):
    """t_next = 2*spmv_result - t_prev (cuquantum.py:534-539)."""
    _c64(spmv_result, "spmv_result"); _c64(t_prev, "t_prev"); _c64(t_next, "t_next")
    call("qsim_chebyshev_step", stream_handle(t_next.device), ptr(spmv_result), ptr(t_prev),
         ptr(t_next), t_next.numel())


def chebyshev_accum(accum, t_k, c_re, c_im):
    """accum += c * t_k (cuquantum.py:542-547)."""
    _c64(accum, "accum"); _c64(t_k, "t_k")
    call("qsim_chebyshev_accum", stream_handle(accum.device), ptr(accum), ptr(t_k), accum.numel(),
         float(c_re), float(c_im))


def chebyshev_step_accum(spmv_result, t_prev, t_next, accum, c_re, c_im):
    """Fused step + accumulation (cuquantum.py:550-555)."""
    for t, nm in ((spmv_result, "spmv_result"), (t_prev, "t_prev"), (t_next, "t_next"),
                  (accum, "accum")):
        _c64(t, nm)
    call("qsim_chebyshev_step_accum", stream_handle(accum.device), ptr(spmv_result), ptr(t_prev),
         ptr(t_next), ptr(accum), accum.numel(), float(c_re), float(c_im))


# ------------------------------------------------------------------ graphs
class CircuitGraph:
    """HIP-graph capture of a 1Q gate list, replayed with one launch
    (cuquantum.py:101-130).  torch.cuda.CUDAGraph is a hipGraph on ROCm."""

    def __init__(self, state, gates, targets):
        if isinstance(targets, torch.Tensor):
            targets = targets.tolist()
        gates = _dev_matrix(gates, (len(targets), 2, 2), "gates", state.device)
        self._gates = gates
        for i, t in enumerate(targets):
            apply_gate(state, gates[i], int(t))
        torch.cuda.synchronize(state.device)
        self._graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self._graph):
            for i, t in enumerate(targets):
                apply_gate(state, gates[i], int(t))
        self._state = state

    def replay(self):
        self._graph.replay()


# ------------------------------------------------------------------ constants
def hadamard_gate(device="cuda"):
    return torch.tensor([[1, 1], [1, -1]], dtype=torch.complex64, device=device) / (2 ** 0.5)


def pauli_x_gate(device="cuda"):
    return torch.tensor([[0, 1], [1, 0]], dtype=torch.complex64, device=device)


def cnot_gate(device="cuda"):
    """CX with the first (control) qubit as the row MSB."""
    return torch.tensor([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]],
                        dtype=torch.complex64, device=device)


def cz_gate(device="cuda"):
    return torch.diag(torch.tensor([1, 1, 1, -1], dtype=torch.complex64)).to(device)


def swap_gate(device="cuda"):
    return torch.tensor([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]],
                        dtype=torch.complex64, device=device)


def cswap_gate(device="cuda"):
    """SWAP matrix for apply_mc2qgate(state, swap, [control], t0, t1)."""
    return swap_gate(device)


def _haar(dim, generator=None):
    z = torch.randn(dim, dim, dtype=torch.complex128, generator=generator)
    q, r = torch.linalg.qr(z)
    d = torch.diagonal(r)
    return (q * (d / d.abs()).unsqueeze(0)).to(torch.complex64)


def random_unitary(device="cuda"):
    """Haar-random 2x2 unitary (QR of a complex Gaussian, phase-fixed R)."""
    return _haar(2).to(device)


def random_unitary_4x4(device="cuda"):
    """Haar-random 4x4 unitary (QR of a complex Gaussian, phase-fixed R)."""
    return _haar(4).to(device)


def rz_gate_diag(theta, device="cuda"):
    return torch.tensor([complex(math.cos(theta / 2), -math.sin(theta / 2)),
                         complex(math.cos(theta / 2), math.sin(theta / 2))],
                        dtype=torch.complex64, device=device)


def phase_gate_diag(theta, device="cuda"):
    return torch.tensor([1.0, complex(math.cos(theta), math.sin(theta))],
                        dtype=torch.complex64, device=device)


def t_gate_diag(device="cuda"):
    return phase_gate_diag(math.pi / 4, device)


def s_gate_diag(device="cuda"):
    return torch.tensor([1.0, 1j], dtype=torch.complex64, device=device)


def z_gate_diag(device="cuda"):
    return torch.tensor([1.0, -1.0], dtype=torch.complex64, device=device)


def cz_gate_diag(device="cuda"):
    return torch.tensor([1.0, 1.0, 1.0, -1.0], dtype=torch.complex64, device=device)


# ------------------------------------------------------------------ Pauli strings
def _mask(v, name):
    v = int(v)
    if v < 0:
        raise RuntimeError(f"{name} must be non-negative")
    return v


def expectation_paulis(state, x_masks, z_masks):
    """<psi|P_k|psi> for Pauli strings P_k = i^popcount(x&z) X^x Z^z as a
    float64 device tensor [K]; terms sharing an x mask share one state pass.
    FP64, fixed summation order (deterministic).  The reference has one Z
    string or one X/Y qubit per call (cuquantum.py:458-487)."""
    _check_state(state)
    n = n_qubits_of(state)
    xs = [_mask(v, "x mask") for v in x_masks]
    zs = [_mask(v, "z mask") for v in z_masks]
    if len(xs) != len(zs):
        raise RuntimeError("x_masks and z_masks differ in length")
    out = torch.empty(len(xs), dtype=torch.float64, device=state.device)
    if not xs:
        return out
    xa = (_lib.ctypes.c_uint64 * len(xs))(*xs)
    za = (_lib.ctypes.c_uint64 * len(zs))(*zs)
    ws = _lib.workspace("expect", _lib.load().qsim_expectation_workspace_bytes(n), state.device)
    call("qsim_expectation_paulis", stream_handle(state.device), ptr(state), n, len(xs),
         _lib.ctypes.cast(xa, _lib.ctypes.c_void_p), _lib.ctypes.cast(za, _lib.ctypes.c_void_p),
         ptr(ws), ptr(out))
    return out


def expectation_pauli_string(state, x_mask, z_mask):
    """<psi|i^popcount(x&z) X^x Z^z|psi> as a float (FP64)."""
    return float(expectation_paulis(state, [x_mask], [z_mask]).item())


def expectation_z(state, z_mask):
    """<psi|Z on the z_mask qubits|psi>, FP64 (cuquantum.py:458-469)."""
    _check_state(state)
    n = n_qubits_of(state)
    ws = _lib.workspace("expect", _lib.load().qsim_expectation_workspace_bytes(n), state.device)
    out = torch.empty(1, dtype=torch.float64, device=state.device)
    call("qsim_expectation_z", stream_handle(state.device), ptr(state), n, _mask(z_mask, "z_mask"),
         ptr(ws), ptr(out))
    return float(out.item())


def expectation_pauli(state, z_mask, xy_qubit, op):
    """<psi|(Z on z_mask) (op on xy_qubit)|psi>, op 0 = I/Z, 1 = X, 2 = Y;
    a z_mask bit on xy_qubit is ignored (cuquantum.py:472-487)."""
    _check_state(state)
    n = n_qubits_of(state)
    ws = _lib.workspace("expect", _lib.load().qsim_expectation_workspace_bytes(n), state.device)
    out = torch.empty(1, dtype=torch.float64, device=state.device)
    call("qsim_expectation_pauli", stream_handle(state.device), ptr(state), n,
         _mask(z_mask, "z_mask"), int(xy_qubit), int(op), ptr(ws), ptr(out))
    return float(out.item())


def apply_pauli_string(state, x_mask, z_mask):
    """state <- i^popcount(x&z) X^x Z^z state in place, bit-exact."""
    _check_state(state)
    n = n_qubits_of(state)
    call("qsim_apply_pauli_string", stream_handle(state.device), ptr(state), n,
         _mask(x_mask, "x_mask"), _mask(z_mask, "z_mask"))
    return state


# ------------------------------------------------------------------ noise channels
def apply_depolarizing(state, target_qubit, p, random_val):
    """Depolarizing step: 0 = identity (random_val < 1 - p), else 1/2/3 for
    X/Y/Z on target_qubit, applied bit-exactly (cuquantum.py:352-369)."""
    _check_state(state)
    n = n_qubits_of(state)
    out = _lib.ctypes.c_int(-1)
    call("qsim_apply_depolarizing", stream_handle(state.device), ptr(state), n, int(target_qubit),
         float(p), float(random_val), _lib.ctypes.byref(out))
    return int(out.value)


def _damping(entry, state, target_qubit, prob, random_val):
    _check_state(state)
    n = n_qubits_of(state)
    ws = _lib.workspace("reduce", _lib.load().qsim_reduce_workspace_bytes(n), state.device)
    out = _lib.ctypes.c_int(-1)
    call(entry, stream_handle(state.device), ptr(state), n, int(target_qubit), float(prob),
         float(random_val), ptr(ws), _lib.ctypes.byref(out))
    return int(out.value)


def apply_amplitude_damping(state, target_qubit, gamma, random_val):
    """Amplitude damping step: 1 (decay to |0>) when random_val < gamma *
    P(target = 1), else 0; renormalised, FP64 scales (cuquantum.py:372-390)."""
    return _damping("qsim_apply_amplitude_damping", state, target_qubit, gamma, random_val)


def apply_dephasing(state, target_qubit, lam, random_val):
    """Dephasing step: 1 (E1 = diag(0, sqrt(lam))) when random_val < lam *
    P(target = 1), else 0 (E0 = diag(1, sqrt(1 - lam))); renormalised
    (cuquantum.py:393-411)."""
    return _damping("qsim_apply_dephasing", state, target_qubit, lam, random_val)
