"""ctypes binding of libqsim_gfx950.so (C ABI: include/qsim_gfx950.h).

torch is imported first on purpose: it loads its own HIP runtime
(libamdhip64.so.7), and the dynamic linker then binds the library's
libamdhip64.so.7 dependency to that same runtime, so device pointers from the
torch allocator and streams from torch.cuda are valid on both sides.

There is no fallback: if the library is missing or fails to load, every
entry point raises.
"""
import ctypes
import os
import threading

import numpy as np
import torch

LIB_NAME = "libqsim_gfx950.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

QSIM_OK = 0

_vp, _i, _i64, _u64, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pint = ctypes.POINTER(ctypes.c_int)
_pd = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); one entry per declaration of qsim_gfx950.h
PROTOTYPES = {
    "qsim_version": (ctypes.c_char_p, []),
    "qsim_abi_version": (_i, []),
    "qsim_last_error": (ctypes.c_char_p, []),
    "qsim_device_count": (_i, [_pint]),
    "qsim_state_init": (_i, [_vp, _vp, _i]),
    "qsim_state_init_batched": (_i, [_vp, _vp, _i64, _i]),
    "qsim_apply_gate": (_i, [_vp, _vp, _i, _vp, _i]),
    "qsim_apply_gate_h": (_i, [_vp, _vp, _i, _pd, _i]),
    "qsim_apply_gate_2q": (_i, [_vp, _vp, _i, _vp, _i, _i]),
    "qsim_apply_gate_2q_h": (_i, [_vp, _vp, _i, _pd, _i, _i]),
    "qsim_apply_diagonal": (_i, [_vp, _vp, _i, _vp, _i]),
    "qsim_apply_diagonal_h": (_i, [_vp, _vp, _i, _pd, _i]),
    "qsim_apply_diagonal_phase": (_i, [_vp, _vp, _i, _d, _d, _i]),
    "qsim_apply_diagonal_2q": (_i, [_vp, _vp, _i, _vp, _i, _i]),
    "qsim_apply_mcgate": (_i, [_vp, _vp, _i, _vp, _pi64, _i, _i]),
    "qsim_apply_mc2qgate": (_i, [_vp, _vp, _i, _vp, _pi64, _i, _i, _i]),
    "qsim_apply_mcx": (_i, [_vp, _vp, _i, _pi64, _i, _i]),
    "qsim_apply_mcswap": (_i, [_vp, _vp, _i, _pi64, _i, _i, _i]),
    "qsim_apply_gates_fused": (_i, [_vp, _vp, _i, _vp, _pi32, _i]),
    "qsim_apply_gates_reg_tiled": (_i, [_vp, _vp, _i, _vp, _vp, _i]),
    "qsim_tile_program": (_i, [_vp, _vp, _i, _u64, _vp, _i, _vp]),
    "qsim_apply_block_mfma": (_i, [_vp, _vp, _i, _vp, _pint, _i]),
    "qsim_plan_blocks": (_i, [_i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _pint]),
    "qsim_plan_blocks_search": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _pint]),
    "qsim_apply_blocks": (_i, [_vp, _vp, _i, _vp, _vp, _i]),
    "qsim_apply_block_batch": (_i, [_vp, _vp, _vp, _i64]),
    "qsim_reduce_workspace_bytes": (_i64, [_i]),
    "qsim_norm2": (_i, [_vp, _vp, _i, _vp, _vp]),
    "qsim_measure_probs": (_i, [_vp, _vp, _i, _vp]),
    "qsim_measure_qubit": (_i, [_vp, _vp, _i, _i, _d, _vp, _pint]),
    "qsim_renormalize": (_i, [_vp, _vp, _i, _vp]),
    "qsim_sample_workspace_bytes": (_i64, [_i]),
    "qsim_sample": (_i, [_vp, _vp, _i, _vp, _i64, _vp, _vp]),
    "qsim_sample_noalloc": (_i, [_vp, _vp, _i, _vp, _i64, _vp]),
    "qsim_uniforms_splitmix": (_i, [_vp, _u64, _i64, _vp]),
    "qsim_sample_batch_workspace_bytes": (_i64, [_i64, _i64]),
    "qsim_sample_batch": (_i, [_vp, _vp, _vp, _i64, _vp]),
    "qsim_expectation_workspace_bytes": (_i64, [_i]),
    "qsim_expectation_paulis": (_i, [_vp, _vp, _i, _i64, _vp, _vp, _vp, _vp]),
    "qsim_expectation_z": (_i, [_vp, _vp, _i, _u64, _vp, _vp]),
    "qsim_expectation_pauli": (_i, [_vp, _vp, _i, _u64, _i, _i, _vp, _vp]),
    "qsim_apply_pauli_string": (_i, [_vp, _vp, _i, _u64, _u64]),
    "qsim_apply_depolarizing": (_i, [_vp, _vp, _i, _i, _d, _d, _pint]),
    "qsim_apply_amplitude_damping": (_i, [_vp, _vp, _i, _i, _d, _d, _vp, _pint]),
    "qsim_apply_dephasing": (_i, [_vp, _vp, _i, _i, _d, _d, _vp, _pint]),
    "qsim_apply_gate_batched": (_i, [_vp, _vp, _i64, _i, _vp, _i]),
    "qsim_apply_gates_batched_fused": (_i, [_vp, _vp, _i64, _i, _vp, _vp, _i]),
    "qsim_run_instances": (_i, [_vp, _vp, _u64, _vp, _vp, _i, _vp, _vp, _u64, _vp, _u64, _vp, _u64, _i]),
    "qsim_qv8_simulate": (_i, [_vp, _vp, _vp, _i64, _i, _vp]),
    "qsim_qv4_simulate": (_i, [_vp, _vp, _vp, _i64, _i, _vp]),
    "qsim_pauli_csr_workspace_bytes": (_i64, [_i]),
    "qsim_pauli_csr_count": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _pi64]),
    "qsim_pauli_csr_fill": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "qsim_spmv_csr": (_i, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _d, _d, _d, _d]),
    "qsim_chebyshev_step": (_i, [_vp, _vp, _vp, _vp, _i64]),
    "qsim_chebyshev_accum": (_i, [_vp, _vp, _vp, _i64, _d, _d]),
    "qsim_chebyshev_step_accum": (_i, [_vp, _vp, _vp, _vp, _vp, _i64, _d, _d]),
    "qsim_cheb_spmm_step": (_i, [_vp, _vp, _vp, _vp, _i64, _i, _d, _vp, _vp, _vp, _vp, _d, _d, _i, _i]),
    "qsim_csr_split_workspace_bytes": (_i64, [_i64]),
    "qsim_csr_split_count": (_i, [_vp, _vp, _vp, _i64, _i, _vp, _pi64, _pi64]),
    "qsim_csr_split_fill": (_i, [_vp, _vp, _vp, _vp, _i64, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "qsim_cheb_spmm_step_split": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _i, _i, _d, _vp, _vp, _vp,
                                       _vp, _vp, _d, _d, _i]),
    "qsim_cheb_pauli_step": (_i, [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp, _vp, _d, _d, _i]),
    "qsim_cheb_init_acc": (_i, [_vp, _vp, _vp, _vp, _i64, _d, _d]),
    "qsim_cheb_finish": (_i, [_vp, _vp, _vp, _vp, _i64]),
}

# numpy mirrors of the structs of the C ABI
OP_DTYPE = np.dtype([("kind", "u1"), ("nq", "u1"), ("q", "u1", (6,)),
                     ("cmask", "<u4"), ("coef", "<u4")])
INSTANCE_DTYPE = np.dtype([("state_off", "<u8"), ("seed", "<u8"), ("shot_off", "<u8"),
                           ("op_off", "<u4"), ("n_ops", "<u4"), ("shots", "<u4"),
                           ("n_qubits", "u1"), ("init", "u1"), ("want", "u1"), ("pad", "u1")])
BLOCK_REC_DTYPE = np.dtype([("state", "<u8"), ("mat", "<u8"), ("set_begin", "<u8"),
                            ("n_qubits", "<i4"), ("q", "i1", (5,)), ("pad", "i1", (7,))])
SAMPLE_REC_DTYPE = np.dtype([("state", "<u8"), ("out", "<u8"), ("seed", "<u8"), ("leaf_off", "<u8"),
                             ("shot_off", "<u8"), ("n_qubits", "<i4"), ("shots", "<u4")])
assert SAMPLE_REC_DTYPE.itemsize == 48
assert OP_DTYPE.itemsize == 16 and INSTANCE_DTYPE.itemsize == 40 and BLOCK_REC_DTYPE.itemsize == 40

OP_M1, OP_M2, OP_D1, OP_D2, OP_PH, OP_X, OP_SWAP, OP_MK, OP_DK = range(1, 10)

_lib = None
_lock = threading.Lock()


class QsimError(RuntimeError):
    """A libqsim_gfx950 entry point returned a non-zero status."""


def load():
    """Load the library once and attach the prototypes."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_NAME} is not built at {LIB_PATH}; run "
                    "`python qiskit-blackwell_amd/build.py` (the gfx950 library has no fallback)")
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in PROTOTYPES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def call(name, *args):
    """Call a status-returning entry point; raise QsimError on failure."""
    rc = getattr(load(), name)(*args)
    if rc != QSIM_OK:
        msg = load().qsim_last_error().decode(errors="replace")
        raise QsimError(f"{name}: {msg} (status {rc})")
    return rc


def stream_handle(device=None):
    """HIP stream handle of torch's current stream on `device`."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def i64_array(values):
    arr = (ctypes.c_int64 * max(1, len(values)))(*[int(v) for v in values])
    return arr


def i32_array(values):
    return (ctypes.c_int32 * max(1, len(values)))(*[int(v) for v in values])


def int_array(values):
    return (ctypes.c_int * max(1, len(values)))(*[int(v) for v in values])


def f64_array(values):
    return (ctypes.c_double * len(values))(*[float(v) for v in values])


class _Workspace(threading.local):
    def __init__(self):
        self.bufs = {}


_ws = _Workspace()


def workspace(kind, nbytes, device):
    """Scratch tensor of at least nbytes for launches on torch's current
    stream of `device`, one per (thread, kind, device, stream), grown on
    demand.  A buffer is only ever used on the stream it was allocated on:
    torch's caching allocator hands a freed block out again in that stream's
    order, so replacing a buffer that queued kernels still read is safe, and
    two streams never share scratch."""
    stream = torch.cuda.current_stream(device)
    key = (kind, str(device), stream.cuda_stream)
    buf = _ws.bufs.get(key)
    if buf is None or buf.numel() < nbytes:
        buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
        _ws.bufs[key] = buf
    return buf
