"""QV-8 batched simulation surface (bwk/python/blackwell_kernels/qv8.py).

generate_qv8_circuits follows the reference's recipe (qv8.py:55-113): one
seeded numpy Generator draws the per-layer qubit permutations first, then
all Haar U(4) gates at once as QR of a complex Gaussian with the R-diagonal
phase folded into Q.  qv8_simulate runs every circuit in one launch on the
gfx950 library (one wave per circuit, state in LDS).
"""
import math

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream_handle

N_QUBITS = 8
STATE_SIZE = 1 << N_QUBITS


def haar_u4_batch(rng, count):
    """count Haar-random U(4) as complex64 [count, 4, 4] (qv8.py:55-73)."""
    z = (rng.standard_normal((count, 4, 4)) + 1j * rng.standard_normal((count, 4, 4))) / math.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r, axis1=-2, axis2=-1)
    q = q * (d / np.abs(d))[:, np.newaxis, :]
    return q.astype(np.complex64)


def generate_qv8_circuits(num_circuits, num_layers=N_QUBITS, seed=42):
    """gate_matrices f32 [C,G,4,4,2], gate_qubits i32 [C,G,2], G (qv8.py:76-113)."""
    rng = np.random.default_rng(seed)
    num_gates = num_layers * (N_QUBITS // 2)
    perms = [rng.permutation(N_QUBITS) for _ in range(num_layers)]
    u = haar_u4_batch(rng, num_circuits * num_gates).reshape(num_circuits, num_gates, 4, 4)
    mats = np.stack([u.real, u.imag], axis=-1).astype(np.float32)
    qubits = np.zeros((num_circuits, num_gates, 2), dtype=np.int32)
    g = 0
    for layer in range(num_layers):
        for pair in range(N_QUBITS // 2):
            qubits[:, g, 0] = int(perms[layer][2 * pair])
            qubits[:, g, 1] = int(perms[layer][2 * pair + 1])
            g += 1
    return (torch.from_numpy(np.ascontiguousarray(mats)),
            torch.from_numpy(np.ascontiguousarray(qubits)), num_gates)


def generate_qv8_circuits_gpu(num_circuits, num_layers=N_QUBITS, seed=42, device="cuda"):
    """Same layout as generate_qv8_circuits, tensors placed on `device`."""
    m, q, g = generate_qv8_circuits(num_circuits, num_layers, seed)
    return m.to(device), q.to(device), g


def qv8_simulate(gate_matrices, gate_qubits, num_circuits):
    """probs f32 [C, 256]; row index r = 2*bit(q1) + bit(q0), q0 the first
    listed qubit (qv8_sim_sm120.cu:88-90)."""
    if not (gate_matrices.is_cuda and gate_qubits.is_cuda):
        raise RuntimeError("gate_matrices and gate_qubits must be CUDA")
    if gate_matrices.dtype != torch.float32:
        raise RuntimeError("float32 required")
    if gate_qubits.dtype != torch.int32:
        raise RuntimeError("int32 required")
    gm = gate_matrices.contiguous()
    gq = gate_qubits.contiguous()
    C = int(num_circuits)
    G = gm.shape[1]
    probs = torch.empty(C, STATE_SIZE, dtype=torch.float32, device=gm.device)
    call("qsim_qv8_simulate", stream_handle(gm.device), ptr(gm), ptr(gq), C, G, ptr(probs))
    return probs


def qv8_simulate_ref(gate_matrices, gate_qubits, num_circuits):
    """The reference's own vectorised torch restatement (qv8.py:207-257),
    kept for API parity; it is not used by any path of this package."""
    device = gate_matrices.device
    gates = torch.complex(gate_matrices[..., 0].double(), gate_matrices[..., 1].double())
    C, G = int(num_circuits), gate_matrices.size(1)
    state = torch.zeros(C, STATE_SIZE, dtype=torch.complex128, device=device)
    state[:, 0] = 1.0
    idx = torch.arange(STATE_SIZE, device=device)
    for g in range(G):
        q0 = int(gate_qubits[0, g, 0]); q1 = int(gate_qubits[0, g, 1])
        m0, m1 = 1 << q0, 1 << q1
        b = idx[((idx & m0) == 0) & ((idx & m1) == 0)]
        cols = [b, b | m0, b | m1, b | m0 | m1]
        amps = torch.stack([state[:, c] for c in cols], dim=-1)
        new = torch.einsum("cij,cnj->cni", gates[:, g], amps)
        for k, c in enumerate(cols):
            state[:, c] = new[..., k]
    return (state.real ** 2 + state.imag ** 2).float()
