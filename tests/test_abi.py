"""The C ABI on the CPU: the library loads, exports every function declared
in include/qsim_gfx950.h, the ctypes prototypes agree with the header, the
struct layouts agree with a C compiler's, and argument validation answers
before any device work (no GPU needed)."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qsim_gfx950.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?\w+\**\s+\**(qsim_\w+)\s*\(([^;]*?)\)\s*;", src, re.M | re.S):
        args = m.group(2).strip()
        n = 0 if args in ("", "void") else args.count(",") + 1
        out[m.group(1)] = n
    return out


def test_header_parses_and_prototypes_match():
    from blackwell_kernels import _lib
    decl = declared_functions()
    assert len(decl) >= 40
    assert set(decl) == set(_lib.PROTOTYPES), set(decl) ^ set(_lib.PROTOTYPES)
    for name, nargs in decl.items():
        assert len(_lib.PROTOTYPES[name][1]) == nargs, name


def test_library_exports_every_symbol():
    from blackwell_kernels import _lib
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(declared_functions()) <= exported


def test_struct_layout_matches_c_compiler():
    from blackwell_kernels import _lib
    prog = r'''
#include <stdio.h>
#include <stddef.h>
#include "qsim_gfx950.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(qsim_op), offsetof(qsim_op, cmask), offsetof(qsim_op, coef),
         offsetof(qsim_op, q));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(qsim_instance), offsetof(qsim_instance, op_off),
         offsetof(qsim_instance, shots), offsetof(qsim_instance, n_qubits),
         offsetof(qsim_instance, want), offsetof(qsim_instance, shot_off));
  printf("%zu %zu %zu %zu %zu\n", sizeof(qsim_block_rec), offsetof(qsim_block_rec, mat),
         offsetof(qsim_block_rec, set_begin), offsetof(qsim_block_rec, n_qubits),
         offsetof(qsim_block_rec, q));
  printf("%zu %zu %zu %zu %zu\n", sizeof(qsim_sample_rec), offsetof(qsim_sample_rec, seed),
         offsetof(qsim_sample_rec, shot_off), offsetof(qsim_sample_rec, n_qubits),
         offsetof(qsim_sample_rec, shots));
  return 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "probe.c")
        exe = os.path.join(d, "probe")
        open(c, "w").write(prog)
        subprocess.run(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe], check=True)
        l1, l2, l3, l4 = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split("\n")[:4]
    op = [int(x) for x in l1.split()]
    ins = [int(x) for x in l2.split()]
    rec = [int(x) for x in l3.split()]
    sd = _lib.SAMPLE_REC_DTYPE
    assert [int(x) for x in l4.split()] == [sd.itemsize, sd.fields["seed"][1], sd.fields["shot_off"][1],
                                            sd.fields["n_qubits"][1], sd.fields["shots"][1]]
    bd = _lib.BLOCK_REC_DTYPE
    assert rec == [bd.itemsize, bd.fields["mat"][1], bd.fields["set_begin"][1], bd.fields["n_qubits"][1],
                   bd.fields["q"][1]]
    od, idt = _lib.OP_DTYPE, _lib.INSTANCE_DTYPE
    assert op == [od.itemsize, od.fields["cmask"][1], od.fields["coef"][1], od.fields["q"][1]]
    assert ins == [idt.itemsize, idt.fields["op_off"][1], idt.fields["shots"][1],
                   idt.fields["n_qubits"][1], idt.fields["want"][1], idt.fields["shot_off"][1]]


def test_meta_and_validation_without_device():
    from blackwell_kernels import _lib
    lib = _lib.load()
    assert lib.qsim_version().startswith(b"qsim-gfx950")
    assert lib.qsim_abi_version() == 2
    c = ctypes.c_int(-1)
    assert lib.qsim_device_count(ctypes.byref(c)) == 0 and c.value >= 0
    # argument errors are reported before any HIP call
    assert lib.qsim_state_init(None, None, 0) == 2                      # QSIM_ERANGE
    assert b"n_qubits" in lib.qsim_last_error()
    assert lib.qsim_apply_gate(None, ctypes.c_void_p(8), 4, ctypes.c_void_p(8), 4) == 1
    assert b"target_qubit" in lib.qsim_last_error()
    assert lib.qsim_apply_gate_2q(None, ctypes.c_void_p(8), 4, ctypes.c_void_p(8), 1, 1) == 1
    ctl = _lib.i64_array([2])
    assert lib.qsim_apply_mcx(None, ctypes.c_void_p(8), 4, ctl, 1, 2) == 1
    assert b"controls" in lib.qsim_last_error()
    assert lib.qsim_apply_block_mfma(None, ctypes.c_void_p(8), 12, ctypes.c_void_p(8),
                                     _lib.int_array([0, 1, 2, 3, 4, 5, 6]), 7) == 2
    assert lib.qsim_apply_block_mfma(None, ctypes.c_void_p(8), 12, ctypes.c_void_p(8),
                                     _lib.int_array([0, 1, 2, 3, 4, 5]), 6) == 2      # k = 6 needs n >= 13
    assert lib.qsim_apply_block_mfma(None, ctypes.c_void_p(8), 13, ctypes.c_void_p(8),
                                     _lib.int_array([0, 1, 2, 3, 4, 4]), 6) == 1      # repeated qubit
    assert lib.qsim_sample_workspace_bytes(20) > 0
    assert lib.qsim_sample_workspace_bytes(0) == -1
    assert lib.qsim_tile_program(None, ctypes.c_void_p(8), 10, 0b110, ctypes.c_void_p(8), 1,
                                 ctypes.c_void_p(8)) == 1                 # mask without qubit 0
    v8_ = ctypes.c_void_p(8)
    assert lib.qsim_run_instances(None, None, 0, v8_, v8_, 1, v8_, v8_, 1, v8_, 1, None, 0, 15) == 2
    # Pauli strings and noise channels
    v8 = ctypes.c_void_p(8)
    assert lib.qsim_expectation_pauli(None, v8, 4, 0, 1, 3, v8, v8) == 1
    assert b"op must be" in lib.qsim_last_error()
    assert lib.qsim_expectation_pauli(None, v8, 4, 0, 4, 1, v8, v8) == 1
    xm = (ctypes.c_uint64 * 2)(1, 16)
    zm = (ctypes.c_uint64 * 2)(0, 0)
    assert lib.qsim_expectation_paulis(None, v8, 4, 2, xm, zm, v8, v8) == 1
    assert b"term 1" in lib.qsim_last_error()
    assert lib.qsim_apply_pauli_string(None, v8, 4, 0, 32) == 1
    b = ctypes.c_int(-1)
    assert lib.qsim_apply_depolarizing(None, v8, 4, 0, 1.5, 0.5, ctypes.byref(b)) == 1
    assert lib.qsim_apply_amplitude_damping(None, v8, 4, 0, -0.1, 0.5, v8, ctypes.byref(b)) == 1
    assert lib.qsim_apply_dephasing(None, v8, 4, 4, 0.1, 0.5, v8, ctypes.byref(b)) == 1
    assert b"target_qubit" in lib.qsim_last_error()
    assert lib.qsim_expectation_workspace_bytes(20) >= 8 * 1024 * 8
    import numpy as np
    recs = np.zeros(2, dtype=_lib.BLOCK_REC_DTYPE)
    recs["state"], recs["mat"], recs["n_qubits"] = 8, 8, 12
    recs["q"] = [0, 1, 2, 3, 4]
    recs[1]["set_begin"] = 3                                           # prefix must be 4
    rp = recs.ctypes.data_as(ctypes.c_void_p)
    assert lib.qsim_apply_block_batch(None, rp, v8, 2) == 1
    assert b"set_begin" in lib.qsim_last_error()
    recs[1]["set_begin"] = 4
    recs[1]["q"] = [0, 1, 2, 3, 3]
    assert lib.qsim_apply_block_batch(None, rp, v8, 2) == 1
    recs[1]["q"] = [0, 1, 2, 3, 4]
    recs[1]["n_qubits"] = 9
    assert lib.qsim_apply_block_batch(None, rp, v8, 2) == 2
    from blackwell_kernels.cuquantum import sample_records
    srecs, leaves = sample_records([8, 8], [10, 12], [8, 8], [1, 2], [16, 0])
    assert leaves == 4 + 16 and list(srecs["leaf_off"]) == [0, 4] and list(srecs["shot_off"]) == [0, 16]
    sp_ = srecs.ctypes.data_as(ctypes.c_void_p)
    srecs[1]["leaf_off"] = 3
    assert lib.qsim_sample_batch(None, sp_, v8, 2, v8) == 1
    srecs[1]["leaf_off"] = 4
    srecs[1]["n_qubits"] = 21
    assert lib.qsim_sample_batch(None, sp_, v8, 2, v8) == 2


def test_python_layer_raises_runtime_error():
    import torch
    import blackwell_kernels as bk
    with pytest.raises(RuntimeError):
        bk.apply_gate(torch.zeros(4, dtype=torch.complex64), bk.pauli_x_gate("cpu"), 0)
    with pytest.raises(RuntimeError):
        bk.expectation_z(torch.zeros(4, dtype=torch.complex64), 1)
    with pytest.raises(RuntimeError):
        bk.apply_depolarizing(torch.zeros(4, dtype=torch.complex64), 0, 0.1, 0.5)


def test_kernel_path_loading(tmp_path, monkeypatch):
    """The simulator finds its kernel module by kernel_path, then
    $BLACKWELL_KERNEL_PATH, then its own package (simulator.py:46-63 of the
    reference); a second build at another path replaces the loaded one, as
    the reference's kernel A/B harness does (tests/kernel_ab.py:183-230)."""
    import shutil
    import sys
    from blackwell_backend.simulator import _load_kernels
    pkg = os.path.join(ROOT, "qiskit-blackwell_amd")
    bk = _load_kernels(pkg)
    assert os.path.abspath(bk.__file__).startswith(os.path.abspath(pkg))
    other = tmp_path / "build_b"
    shutil.copytree(os.path.join(pkg, "blackwell_kernels"), other / "blackwell_kernels",
                    ignore=shutil.ignore_patterns("__pycache__"))
    try:
        bk_b = _load_kernels(str(other))
        assert os.path.abspath(bk_b.__file__).startswith(str(other))
        assert bk_b._lib.load().qsim_abi_version() == 2
        monkeypatch.setenv("BLACKWELL_KERNEL_PATH", pkg)
        bk_env = _load_kernels(None)
        assert os.path.abspath(bk_env.__file__).startswith(os.path.abspath(pkg))
    finally:
        if str(other) in sys.path:
            sys.path.remove(str(other))
        _load_kernels(pkg)


def test_run_instances_rejects_bad_records_without_device():
    """qsim_run_instances checks the host copy of every instance record and
    op before it launches: an instance larger than max_qubits, ops, states,
    shots or coefficients beyond the buffers passed, an op qubit >= n, a
    repeated target, a control on a target and an unknown kind are all
    QSIM_EINVAL with the instance named (no device is touched)."""
    from blackwell_kernels import _lib
    lib = _lib.load()
    v8 = ctypes.c_void_p(8)

    def rec(n=4, op_off=0, n_ops=1, shots=16, state_off=0, init=1, want=1):
        r = np.zeros(1, dtype=_lib.INSTANCE_DTYPE)
        r["n_qubits"], r["op_off"], r["n_ops"], r["shots"] = n, op_off, n_ops, shots
        r["state_off"], r["init"], r["want"] = state_off, init, want
        return r

    def op(kind=_lib.OP_M1, q=(1,), cmask=0, coef=0, nq=None):
        o = np.zeros(1, dtype=_lib.OP_DTYPE)
        o["kind"], o["nq"], o["cmask"], o["coef"] = kind, len(q) if nq is None else nq, cmask, coef
        o["q"][0, :len(q)] = q
        return o

    def run(r, ops, n_state=16, n_coefs=64, n_shots=16, max_q=4):
        return (lib.qsim_run_instances(None, v8, n_state, r.ctypes.data, v8, len(r), ops.ctypes.data, v8,
                                       len(ops), v8, n_coefs, v8, n_shots, max_q),
                lib.qsim_last_error())

    good = run(rec(), op())
    assert good[0] != 1, good          # valid records get past validation (launch fails without a GPU)
    cases = [
        (rec(n=5), op(), {}, b"n_qubits"),
        (rec(n_ops=2), op(), {}, b"ops"),
        (rec(state_off=8), op(), {}, b"state"),
        (rec(init=0, want=0, state_off=8), op(), {}, b"state"),
        (rec(shots=17), op(), {}, b"shots"),
        (rec(), op(q=(4,)), {}, b"qubit 4"),
        (rec(), op(kind=_lib.OP_M2, q=(2, 2)), {}, b"repeated"),
        (rec(), op(q=(1,), cmask=0b10), {}, b"control"),
        (rec(), op(q=(1,), cmask=0b10000), {}, b"control"),
        (rec(), op(kind=_lib.OP_M2, q=(0, 1), coef=60), {}, b"coefficients"),
        (rec(), op(kind=_lib.OP_MK, q=(0, 1, 2), coef=8), {"n_coefs": 64}, b"coefficients"),
        (rec(), op(kind=_lib.OP_MK, q=(0, 1, 2, 3, 0, 0), nq=6), {}, b"dense block"),
        (rec(), op(kind=42), {}, b"unknown kind"),
    ]
    for r, o, kw, msg in cases:
        rc, err = run(r, o, **kw)
        assert rc == 1 and msg in err and b"instance 0" in err, (rc, err, msg)
