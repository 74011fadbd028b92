"""Builds the C FP64 oracle (oracle/sv_fp64.c) into oracle/_build/libsv_fp64.so.

The reference itself cannot be compiled here: its kernels are CUDA for
sm_120a (nvcc is absent) and its Python layer imports qiskit/qiskit-aer
(absent), so there is no oracle/_ref build; the oracle is pinned by closed
forms and library routines instead (tests/test_oracle.py)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_build", "libsv_fp64.so")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fopenmp", "-fPIC", "-shared",
           os.path.join(HERE, "sv_fp64.c"), "-o", OUT]
    subprocess.run(cmd, check=True)
    if verbose:
        print(f"built {OUT}")
    return OUT


def load():
    import ctypes
    if not os.path.exists(OUT):
        build()
    lib = ctypes.CDLL(OUT)
    vp, i, pd, pi = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    lib.sv_set_threads.argtypes = [i]
    lib.sv_set_threads.restype = i
    lib.sv_init.argtypes = [vp, i]
    lib.sv_apply_1q.argtypes = [vp, i, pd, i]
    lib.sv_apply_2q.argtypes = [vp, i, pd, i, i]
    lib.sv_apply_mcx.argtypes = [vp, i, pi, i, i]
    lib.sv_norm2.argtypes = [vp, i]
    lib.sv_norm2.restype = ctypes.c_double
    return lib


if __name__ == "__main__":
    build(verbose=True)
