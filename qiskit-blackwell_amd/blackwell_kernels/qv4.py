"""QV-4 batched simulation surface (bwk/python/blackwell_kernels/qv4.py).

Pair ids follow the reference's table (qv4.py:14-21): 0:(0,1) 1:(0,2)
2:(0,3) 3:(1,2) 4:(1,3) 5:(2,3), with row index r = 2*bit(qb) + bit(qa),
qa < qb.  Gate data per gate is 16 real parts then 16 imaginary parts of the
row-major 4x4 matrix.
"""
import numpy as np
import torch

from ._lib import call, ptr, stream_handle
from .qv8 import haar_u4_batch

PAIR_MAP = {(0, 1): 0, (0, 2): 1, (0, 3): 2, (1, 2): 3, (1, 3): 4, (2, 3): 5}
PAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]


def generate_qv4_circuits(n_circuits, rng=None):
    """gate_data f32 [C, 8, 32] and pair_ids i32 [C, 8]: 4 layers of 2 Haar
    gates on the pairs of a random permutation (qv4.py:22-50)."""
    if rng is None:
        rng = np.random.default_rng()
    gate_data = np.empty((n_circuits, 8, 32), dtype=np.float32)
    pair_ids = np.empty((n_circuits, 8), dtype=np.int32)
    for c in range(n_circuits):
        for layer in range(4):
            perm = rng.permutation(4)
            for sub, pr in enumerate((sorted(perm[:2]), sorted(perm[2:]))):
                g = layer * 2 + sub
                u = haar_u4_batch(rng, 1)[0]
                gate_data[c, g, :16] = u.real.ravel()
                gate_data[c, g, 16:] = u.imag.ravel()
                pair_ids[c, g] = PAIR_MAP[(int(pr[0]), int(pr[1]))]
    return gate_data, pair_ids


def qv4_simulate_cuda(gate_data, pair_ids):
    """probs f32 [C, 16] from the gfx950 kernel (16 circuits per wave)."""
    if isinstance(gate_data, np.ndarray):
        gate_data = torch.from_numpy(gate_data).cuda()
    if isinstance(pair_ids, np.ndarray):
        pair_ids = torch.from_numpy(pair_ids).cuda()
    if gate_data.dtype != torch.float32:
        raise RuntimeError("gate_data must be float32")
    if pair_ids.dtype != torch.int32:
        raise RuntimeError("pair_ids must be int32")
    gd = gate_data.contiguous()
    pid = pair_ids.contiguous()
    C, G = gd.shape[0], gd.shape[1]
    probs = torch.empty(C, 16, dtype=torch.float32, device=gd.device)
    call("qsim_qv4_simulate", stream_handle(gd.device), ptr(gd), ptr(pid), C, G, ptr(probs))
    return probs


def qv4_simulate_numpy(gate_data_np, pair_ids_np):
    """The reference's numpy restatement (qv4.py:74-110), kept for API parity;
    not used by any path of this package."""
    C, G = gate_data_np.shape[0], gate_data_np.shape[1]
    probs = np.empty((C, 16), dtype=np.float32)
    for c in range(C):
        st = np.zeros(16, dtype=np.complex128)
        st[0] = 1.0
        for g in range(G):
            u = gate_data_np[c, g, :16].reshape(4, 4) + 1j * gate_data_np[c, g, 16:].reshape(4, 4)
            qa, qb = PAIRS[int(pair_ids_np[c, g])]
            for base in range(16):
                if base & ((1 << qa) | (1 << qb)):
                    continue
                ii = [base, base | 1 << qa, base | 1 << qb, base | 1 << qa | 1 << qb]
                st[ii] = u @ st[ii]
        probs[c] = np.abs(st) ** 2
    return probs


def heavy_output_probability(probs):
    """Probability mass above the median outcome probability (qv4.py:113-124)."""
    if isinstance(probs, torch.Tensor):
        probs = probs.cpu().numpy()
    med = np.median(probs, axis=1, keepdims=True)
    return (probs * (probs > med)).sum(axis=1)
