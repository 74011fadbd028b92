"""Shared test helpers (host side)."""
import numpy as np


def rand_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return (v / np.linalg.norm(v)).astype(np.complex64)


def rand_unitary(dim, seed):
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    d = np.diagonal(r)
    return q * (d / np.abs(d))[None, :]


def to_dev(v, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(v)).to(dev)


def close(ref, got, rtol=2e-6, atol=2e-7):
    """Max abs error relative to the FP64 reference: compensated FP32
    kernels stay within a few ulp of |a| ~ 2^-n/2 amplitudes."""
    ref = np.asarray(ref, dtype=np.complex128)
    got = np.asarray(got, dtype=np.complex128)
    err = np.max(np.abs(ref - got))
    scale = max(np.max(np.abs(ref)), 1e-30)
    return err <= atol + rtol * scale, err
