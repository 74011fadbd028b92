"""Builds the C FP64 oracle (oracle/sv_fp64.c, oracle/cheb_fp64.c) into
oracle/_build/libsv_fp64_<isa>.so, once per x86-64 ISA level, and loads the
highest level the host CPU supports (the box that runs the checks need not be
the machine that built them).

The reference itself cannot be compiled here: its kernels are CUDA for
sm_120a (nvcc is absent) and its Python layer imports qiskit/qiskit-aer
(absent), so there is no oracle/_ref build; the oracle is pinned by closed
forms and library routines instead (tests/test_oracle.py)."""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
SOURCES = ("sv_fp64.c", "cheb_fp64.c")
# x86-64 micro-architecture levels: v2 (SSE4.2), v3 (AVX2 + FMA), v4 (AVX-512)
ISA_LEVELS = ("x86-64-v2", "x86-64-v3", "x86-64-v4")
_ISA_FLAGS = {
    # /proc/cpuinfo names LZCNT "abm"
    "x86-64-v3": {"avx", "avx2", "fma", "bmi1", "bmi2", "movbe", "f16c", "abm", "xsave"},
    "x86-64-v4": {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"},
}


def lib_path(isa):
    return os.path.join(BUILD, f"libsv_fp64_{isa.replace('x86-64-', '')}.so")


OUT = lib_path("x86-64-v2")     # the baseline build every x86-64 host can load


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [os.path.join(HERE, s) for s in SOURCES]
    for isa in ISA_LEVELS:
        cmd = ["gcc", "-O3", f"-march={isa}", "-fcx-limited-range", "-fopenmp", "-fPIC", "-shared", *srcs,
               "-o", lib_path(isa)]
        subprocess.run(cmd, check=True)
        if verbose:
            print(f"built {lib_path(isa)}")
    return OUT


def _cpu_flags():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_isa():
    """Highest ISA level of ISA_LEVELS that the host CPU supports."""
    flags = _cpu_flags()
    best = "x86-64-v2"
    if _ISA_FLAGS["x86-64-v3"] <= flags:
        best = "x86-64-v3"
        if _ISA_FLAGS["x86-64-v4"] <= flags:
            best = "x86-64-v4"
    return best


_loaded = None


def load():
    """The oracle library built for host_isa() (built here if missing)."""
    global _loaded
    if _loaded is not None:
        return _loaded
    isa = host_isa()
    path = lib_path(isa)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    vp, i, pd, pi = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    lib.sv_set_threads.argtypes = [i]
    lib.sv_set_threads.restype = i
    lib.sv_init.argtypes = [vp, i]
    lib.sv_apply_1q.argtypes = [vp, i, pd, i]
    lib.sv_apply_2q.argtypes = [vp, i, pd, i, i]
    lib.sv_apply_mcx.argtypes = [vp, i, pi, i, i]
    lib.sv_norm2.argtypes = [vp, i]
    lib.sv_norm2.restype = ctypes.c_double
    lib.ph_set_threads.argtypes = [i]
    lib.ph_set_threads.restype = i
    lib.ph_matvec.argtypes = [i, i, vp, vp, vp, vp, i, vp, vp, ctypes.c_double]
    lib.ph_matvec.restype = None
    lib.ph_cheb_evolve.argtypes = [i, i, vp, vp, vp, vp, i, ctypes.c_double, i, vp, vp, vp]
    lib.ph_cheb_evolve.restype = i
    lib.isa = isa
    _loaded = lib
    return lib


if __name__ == "__main__":
    build(verbose=True)
