"""qsim_sample_noalloc, the sampling entry point that allocates nothing (its
scratch is a static device array), against qsim_sample with a caller
workspace: the same indices, bit for bit, on two calls in a row, the second
on a smaller state, and a state above its 26-qubit limit rejected."""
import ctypes

import numpy as np
import pytest
import torch

from helpers import rand_state

pytestmark = pytest.mark.gpu


def _sample_ws(st, n, u):
    from blackwell_kernels import _lib
    dev = st.device
    ws = torch.empty(int(_lib.load().qsim_sample_workspace_bytes(n)), dtype=torch.uint8, device=dev)
    out = torch.full((u.numel(),), -1, dtype=torch.int64, device=dev)
    _lib.call("qsim_sample", _lib.stream_handle(dev), _lib.ptr(st), n, _lib.ptr(u), u.numel(),
              _lib.ptr(ws), _lib.ptr(out))
    return out


def _sample_noalloc(st, n, u):
    from blackwell_kernels import _lib
    dev = st.device
    out = torch.full((u.numel(),), -1, dtype=torch.int64, device=dev)
    _lib.call("qsim_sample_noalloc", _lib.stream_handle(dev), _lib.ptr(st), n, _lib.ptr(u), u.numel(),
              _lib.ptr(out))
    return out


def test_sample_noalloc_equals_workspace_entry(bk, dev):
    import blackwell_kernels as B
    for n, seed, shots in ((22, 5, 50000), (13, 6, 7000)):
        print(f"[noalloc] n={n} shots={shots}", flush=True)
        st = torch.from_numpy(rand_state(n, seed)).to(dev)
        u = B.uniforms_splitmix(100 + n, shots, dev)
        want = _sample_ws(st, n, u).cpu().numpy()
        got = _sample_noalloc(st, n, u).cpu().numpy()
        assert got.min() >= 0 and got.max() < (1 << n)
        assert np.array_equal(got, want), n


def test_sample_noalloc_rejects_above_limit(bk, dev):
    from blackwell_kernels import _lib
    lib = _lib.load()
    v8 = ctypes.c_void_p(8)
    assert lib.qsim_sample_noalloc(None, v8, 27, v8, 16, v8) == 2        # QSIM_ERANGE
    assert b"at most 26" in lib.qsim_last_error()
