"""Benchmark: amp-updates/s and %HBM peak of the statevector gate kernels on
one MI355X.

Workload (one step): the 30-qubit random-SU(4) volume circuit, depth 30
(450 Haar 4x4 gates on seeded random pairs, complex64 state of 8 GiB), run
from |0...0> through the plain per-gate kernels.  value = 450 * 2^30 /
seconds per step.  The same circuit also runs as fused blocks on FP32 MFMA
("fused"), at the fusion width (4, 5 or 6 qubits) the tuner measures fastest
for it; the plain and 5-qubit-block forms are checked against an FP64 CPU
statevector of the same circuit family at 22 qubits ("parity").

Extra figures in the same JSON line:
  roofline        dominant kernel (dense 2Q gate), 16 B per amplitude update,
                  average launch duration from HIP events on the launch stream
  per_gate_30q    each gate class on a 30-qubit state (8 GiB)
  c2_28q          the 28-qubit SU(4) circuit (depth 28), plain and fused
  c3/c5/serving   the 32-qubit ansatz, the 22-qubit Chebyshev evolution and
                  the mixed 10..24-qubit request stream
  cpu_baseline    FP64 C/OpenMP statevector (oracle/sv_fp64.c) on the host
                  cores, bounded sample of the same 30-qubit circuit
  cpu_legs        the CPU path on the 12-qubit ladder circuit (whole circuit)
                  and on the Chebyshev workload (scipy expm_multiply, FP64)
Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "qiskit-blackwell_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

HBM_PEAK_GBS = 8000.0
# HBM bytes per amplitude update that k_gate2q moves, measured with two
# rocprofv3 --pmc runs of this script (FETCH_SIZE, then WRITE_SIZE, both scaled
# by the counts of a device-to-device copy of the state run first with
# --pmc-calibration): one read and one write of the state, nothing more.
GATE2Q_TRAFFIC_B_PER_AMP = 16.0
GATE2Q_TRAFFIC_SOURCE = ("rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, per k_gate2q dispatch of this "
                         "script at 30 qubits, each scaled by the same counter on a 8 GiB device copy "
                         "(--pmc-calibration)")
MFMA_FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class EventTimer:
    """HIP events on the stream the kernels are launched on."""

    def __init__(self, stream):
        self.stream = stream

    def __enter__(self):
        self.a = torch.cuda.Event(enable_timing=True)
        self.b = torch.cuda.Event(enable_timing=True)
        self.a.record(self.stream)
        return self

    def __exit__(self, *exc):
        self.b.record(self.stream)

    def ms(self):
        self.b.synchronize()
        return self.a.elapsed_time(self.b)


def circuit_plan(sim, qc, mode):
    cc = sim.compile(qc)
    return cc, cc.plan(cc.resolve_mode(mode), sim.device)


def run_steps(sim, qc, mode, steps, warmup, stream):
    """Median/mean seconds per step and the kernel-only event time per step."""
    import blackwell_kernels as bk
    cc, plan = circuit_plan(sim, qc, mode)
    state = torch.empty(1 << cc.n_qubits, dtype=torch.complex64, device=sim.device)
    for _ in range(warmup):
        bk.cuquantum._lib.call("qsim_state_init", bk.cuquantum.stream_handle(sim.device),
                               bk.cuquantum.ptr(state), cc.n_qubits)
        plan.run(state)
    torch.cuda.synchronize()
    wall, kern = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        bk.cuquantum._lib.call("qsim_state_init", bk.cuquantum.stream_handle(sim.device),
                               bk.cuquantum.ptr(state), cc.n_qubits)
        with EventTimer(stream) as ev:
            plan.run(state)
        torch.cuda.synchronize()
        wall.append(time.perf_counter() - s0)
        kern.append(ev.ms() / 1e3)
    total = time.perf_counter() - t0
    return state, total, wall, kern, plan.n_launches


def fidelity_check(sim, n, depth, seed, modes):
    """Fidelity / max|d amp| of the GPU result vs the FP64 C statevector."""
    from blackwell_backend import workloads as W
    from blackwell_backend.circuit import instructions
    from oracle import build_ref
    from oracle import statevector as O
    lib = build_ref.load()
    qc = W.su4_volume(n, depth, seed)
    psi = np.zeros(1 << n, dtype=np.complex128)
    lib.sv_init(psi.ctypes.data, n)
    for name, qs, _, params in instructions(qc):
        u = np.asarray(params[0], dtype=np.complex128)
        v = np.empty(32)
        v[0::2] = u.reshape(-1).real
        v[1::2] = u.reshape(-1).imag
        lib.sv_apply_2q(psi.ctypes.data, n, v.ctypes.data_as(_pd()), qs[0], qs[1])
    out = {}
    for mode in modes:
        sv, _ = sim.run_statevector(qc, mode=mode)
        out[mode] = {"fidelity": O.fidelity(psi, sv),
                     "max_abs_damp": O.max_abs_diff_phase_aligned(psi, sv)}
    return out


def _pd():
    import ctypes
    return ctypes.POINTER(ctypes.c_double)


def per_gate_sweep(n=30, reps=4):
    """amp-updates/s and GB/s of each gate class on an n-qubit state."""
    import blackwell_kernels as bk
    from blackwell_kernels import _lib
    dev = torch.device("cuda", 0)
    st = bk.state_init(n, device=dev)
    stream = torch.cuda.current_stream(dev)
    N = 1 << n
    u2 = np.array([[0.6, 0.8j], [0.8j, 0.6]], dtype=np.complex64)
    u4 = np.linalg.qr(np.random.default_rng(1).standard_normal((4, 4)) + 0j)[0].astype(np.complex64)
    g2 = torch.from_numpy(u2).to(dev)
    g4 = torch.from_numpy(u4).to(dev)
    d2 = torch.tensor([np.exp(-0.3j), np.exp(0.3j)], dtype=torch.complex64, device=dev)
    cases = [
        ("1q_dense_t0", lambda: bk.apply_gate(st, g2, 0), N, 16 * N),
        ("1q_dense_t12", lambda: bk.apply_gate(st, g2, 12), N, 16 * N),
        ("1q_dense_t29", lambda: bk.apply_gate(st, g2, n - 1), N, 16 * N),
        ("2q_dense_0_1", lambda: bk.apply_gate_2q(st, g4, 0, 1), N, 16 * N),
        ("2q_dense_7_21", lambda: bk.apply_gate_2q(st, g4, 7, 21), N, 16 * N),
        ("2q_dense_28_29", lambda: bk.apply_gate_2q(st, g4, n - 2, n - 1), N, 16 * N),
        ("diag_t5", lambda: bk.apply_diagonal(st, d2, 5), N, 16 * N),
        ("phase_t9", lambda: bk.apply_diagonal_phase(st, 1j, 9), N // 2, 8 * N),
        ("cnot_17_3", lambda: bk.apply_mcx(st, [17], 3), N // 2, 8 * N),
        ("cnot_3_17", lambda: bk.apply_mcx(st, [3], 17), N // 2, 8 * N),
        ("mcx3_11_19_25_to_2", lambda: bk.apply_mcx(st, [11, 19, 25], 2), N // 8, 2 * N),
        ("mcx3_2_11_25_to_19", lambda: bk.apply_mcx(st, [2, 11, 25], 19), N // 8, 2 * N),
        ("swap_4_23", lambda: bk.apply_mcswap(st, [], 4, 23), N // 2, 8 * N),
        ("norm2", lambda: bk.norm2(st), N, 8 * N),
    ]
    out = {}
    for name, fn, upd, nbytes in cases:
        fn()
        torch.cuda.synchronize()
        with EventTimer(stream) as ev:
            for _ in range(reps):
                fn()
        ms = ev.ms() / reps
        gbs = nbytes / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "amp_updates_per_s": upd / (ms * 1e-3),
                     "GB_per_s": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    # sampling 2^20 shots
    u = torch.rand(1 << 20, dtype=torch.float64, device=dev)
    bk.sample_with_uniforms(st, u)
    torch.cuda.synchronize()
    with EventTimer(stream) as ev:
        bk.sample_with_uniforms(st, u)
    out["sample_2^20_shots"] = {"ms": round(ev.ms(), 3)}
    del st
    torch.cuda.empty_cache()
    return out


def stride_sweep(n=30, reps=6):
    """GB/s of the dense 1Q kernel for every target and of the 2Q kernel for
    (7, q) and (12, q) on an n-qubit state: the partner stream of target t
    sits 2^(t+3) bytes away, and the rate depends on that distance."""
    import blackwell_kernels as bk
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    st = bk.state_init(n, device=dev)
    u2 = torch.tensor([[0.6, 0.8j], [0.8j, 0.6]], dtype=torch.complex64, device=dev)
    u4 = torch.from_numpy(np.linalg.qr(np.random.default_rng(1).standard_normal((4, 4)) + 0j)[0]
                          .astype(np.complex64)).to(dev)
    N = 1 << n

    def gbs(fn):
        fn()
        torch.cuda.synchronize()
        with EventTimer(stream) as ev:
            for _ in range(reps):
                fn()
        return round(16.0 * N / (ev.ms() / reps * 1e-3) / 1e9, 1)

    out = {"1q": {t: gbs(lambda: bk.apply_gate(st, u2, t)) for t in range(n)},
           "2q_7_q": {q: gbs(lambda: bk.apply_gate_2q(st, u4, 7, q)) for q in range(8, n)},
           "2q_12_q": {q: gbs(lambda: bk.apply_gate_2q(st, u4, 12, q)) for q in range(13, n)}}
    del st
    torch.cuda.empty_cache()
    return out


def observables_sweep(n=30, reps=4):
    """GB/s of the Pauli-string expectation, Pauli application and noise
    channel kernels on an n-qubit state, plus the grouped energy of the
    22-qubit chain Hamiltonian (85 Pauli terms)."""
    import blackwell_kernels as bk
    from blackwell_backend import workloads as W
    from blackwell_backend.hamiltonian import expectation
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream(dev)
    N = 1 << n
    g = torch.Generator(device=dev)
    g.manual_seed(3)
    st = torch.randn(N, dtype=torch.complex64, device=dev, generator=g)
    bk.renormalize(st)
    z8 = [(0b11 << (3 * k)) for k in range(8)]
    cases = [
        ("expect_z_string", lambda: bk.expectation_paulis(st, [0], [0b1011 << 20]), 8 * N, 1),
        ("expect_x_low_y", lambda: bk.expectation_paulis(st, [0b101], [0b110]), 8 * N, 1),
        ("expect_xy_high", lambda: bk.expectation_paulis(st, [(1 << 29) | 1], [1 << 29]), 8 * N, 1),
        ("expect_8_z_terms_one_pass", lambda: bk.expectation_paulis(st, [0] * 8, z8), 8 * N, 8),
        ("apply_pauli_x_low", lambda: bk.apply_pauli_string(st, 0b1010, 0b11), 16 * N, 1),
        ("apply_pauli_xy_high", lambda: bk.apply_pauli_string(st, 1 << 28, 1 << 28), 16 * N, 1),
        ("apply_pauli_z_high", lambda: bk.apply_pauli_string(st, 0, 1 << 25), 8 * N, 1),
        ("depolarizing_y_t17", lambda: bk.apply_depolarizing(st, 17, 0.3, 0.85), 16 * N, 1),
        ("amplitude_damping_t9", lambda: bk.apply_amplitude_damping(st, 9, 0.05, 0.99), 24 * N, 1),
    ]
    out = {}
    for name, fn, nbytes, terms in cases:
        fn()
        torch.cuda.synchronize()
        with EventTimer(stream) as ev:
            for _ in range(reps):
                fn()
        ms = ev.ms() / reps
        gbs = nbytes / (ms * 1e-3) / 1e9
        row = {"ms": round(ms, 4), "GB_per_s": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
        if terms > 1:
            row["terms_per_s"] = terms / (ms * 1e-3)
        out[name] = row
    del st
    torch.cuda.empty_cache()
    h = W.heisenberg_chain(22, seed=5)
    psi = torch.randn(1 << 22, dtype=torch.complex64, device=dev, generator=g)
    bk.renormalize(psi)
    expectation(psi, h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e = expectation(psi, h)
    dt = time.perf_counter() - t0
    out["energy_chain_22q"] = {"terms": len(h.labels), "ms": round(dt * 1e3, 3), "energy": e}
    return out


class StepEvents:
    """step_timer for _Plan.run: a pair of HIP events around each step on
    the launch stream, kept for aggregation after one synchronize."""

    def __init__(self, stream):
        self.stream = stream
        self.rows = []

    def __call__(self, kind, args):
        timer = self

        class _Ctx:
            def __enter__(self_):
                self_.a = torch.cuda.Event(enable_timing=True)
                self_.a.record(timer.stream)

            def __exit__(self_, *exc):
                b = torch.cuda.Event(enable_timing=True)
                b.record(timer.stream)
                timer.rows.append((kind, args, self_.a, b))
        return _Ctx()


def _c3_class(kind, a, n):
    """(class name, algorithmic bytes) of one plain-plan step."""
    N = 1 << n
    if kind == "qsim_apply_gate_h":
        return "1q_dense", 16 * N
    if kind == "qsim_apply_diagonal_h":
        return "1q_diagonal", 16 * N
    if kind == "qsim_apply_diagonal_phase":
        return "phase", 8 * N
    if kind == "qsim_apply_mcx":
        nc = a[2]
        name = {0: "x", 1: "cnot", 2: "toffoli"}.get(nc, f"mcx{nc}")
        return name, 16 * (N >> max(nc, 1))
    return kind, 16 * N


def c3_bench(n=32, reps=8, seed=77, shots=1 << 20):
    """Layered rotation ansatz on 32 qubits (32 GiB complex64): 8 reps of
    Ry, Rz on every qubit, CNOT q -> q+1, 4 MCX with 3 random controls, run
    through the plain per-gate kernels (Ry.Rz fused per qubit), then 2^20
    shots from seeded uniforms.  Per-class GB/s from events per launch."""
    import blackwell_kernels as bk
    from blackwell_backend import BlackwellSimulator, workloads as W
    from blackwell_backend.serving import splitmix_uniforms
    sim = BlackwellSimulator("cuda:0")
    qc = W.c3_rotation_mcx(n, reps, seed)
    cc = sim.compile(qc)
    plan = cc.plan("plain", sim.device)
    stream = torch.cuda.current_stream(sim.device)
    st = bk.state_init(n, device=sim.device)
    torch.cuda.synchronize()
    timer = StepEvents(stream)
    t0 = time.perf_counter()
    plan.run(st, step_timer=timer)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    classes = {}
    for kind, a, ea, eb in timer.rows:
        name, nbytes = _c3_class(kind, a, n)
        c = classes.setdefault(name, {"launches": 0, "ms": 0.0, "bytes": 0})
        c["launches"] += 1
        c["ms"] += ea.elapsed_time(eb)
        c["bytes"] += nbytes
    for c in classes.values():
        gbs = c["bytes"] / (c["ms"] * 1e-3) / 1e9
        c["ms"] = round(c["ms"], 3)
        c["GB_per_s"] = round(gbs, 1)
        c["frac_hbm_peak"] = round(gbs / HBM_PEAK_GBS, 4)
        del c["bytes"]
    norm = float(bk.norm2(st))
    u = torch.from_numpy(splitmix_uniforms(seed, shots)).to(sim.device)
    bk.sample_with_uniforms(st, u[:16])
    torch.cuda.synchronize()
    with EventTimer(stream) as ev:
        bk.sample_with_uniforms(st, u)
    samp_ms = ev.ms()
    del st
    torch.cuda.empty_cache()
    return {"workload": f"{n}-qubit rotation + CNOT ladder + 4 MCX(3 controls) ansatz, {reps} reps, "
                        f"seed {seed}, plain per-gate kernels, then {shots} shots",
            "launches": len(timer.rows), "circuit_s": round(wall, 3), "norm2_after": norm,
            "classes": classes,
            "sample": {"shots": shots, "ms": round(samp_ms, 3),
                       "state_read_GB_per_s": round((8 << n) / (samp_ms * 1e-3) / 1e9, 1)}}


def c5_bench(n=22, n_rhs=32, times=(1.0, 5.0, 10.0), seed=5, precision=1e-7):
    """Chebyshev evolution exp(-iHt) of 32 normalised complex-Gaussian
    vectors under the 22-qubit open Heisenberg-type chain, one fused SpMM
    launch per Chebyshev term.  Algorithmic bytes per term:
    12 nnz + 4 (N+1) + 5 * 8 N R (CSR once; T_k, T_{k-1} read, T_{k+1}
    written, accumulator read and written).  Checks that need no FP64
    solve at this size: norm and energy conservation, and
    exp(-iH 5) exp(-iH 5) = exp(-iH 10)."""
    from blackwell_backend import workloads as W
    from blackwell_backend.hamiltonian import HamiltonianCSR
    from blackwell_backend.chebyshev import chebyshev_evolve_batch
    import blackwell_kernels as bk
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    h = HamiltonianCSR.from_sparse_pauli_op(W.heisenberg_chain(n, seed), device=dev)
    build_s = time.perf_counter() - t0
    N = 1 << n
    rng = np.random.default_rng(seed + 1)
    X = (rng.standard_normal((n_rhs, N)) + 1j * rng.standard_normal((n_rhs, N))).astype(np.complex64)
    X /= np.linalg.norm(X, axis=1, keepdims=True).astype(np.float32)
    X = torch.from_numpy(X).to(dev)

    def energy(Y):
        """<y|H|y> per row (FP64 host reduction of the device SpMV)."""
        out = []
        for r in range(min(4, Y.shape[0])):
            hy = torch.empty_like(Y[r])
            bk.spmv_csr(h.row_ptr, h.col_idx, h.values, Y[r].contiguous(), hy)
            out.append(complex(torch.vdot(Y[r].to(torch.complex128), hy.to(torch.complex128))).real)
        return np.array(out)

    e0 = energy(X)
    chebyshev_evolve_batch(X[:2], h, 0.1, precision)          # warm
    res = {}
    outs = {}
    for t in times:
        torch.cuda.synchronize()
        Y, info = chebyshev_evolve_batch(X, h, t, precision)
        K = info["chebyshev_degree"]
        el = info["elapsed_s"]
        per = 12.0 * h.nnz + 4.0 * (N + 1) + 5.0 * 8.0 * N * n_rhs
        gbs = per * K / el / 1e9
        nrm = torch.linalg.vector_norm(Y.to(torch.complex128), dim=1).cpu().numpy()
        res[f"t={t:g}"] = {"K": K, "s": round(el, 4), "ms_per_term": round(el / K * 1e3, 4),
                           "GB_per_s": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
                           "max_norm_dev": float(np.max(np.abs(nrm - 1.0))),
                           "max_energy_drift": float(np.max(np.abs(energy(Y) - e0)))}
        outs[t] = Y
    if 5.0 in outs and 10.0 in outs:
        Y55, _ = chebyshev_evolve_batch(outs[5.0][:4].contiguous(), h, 5.0, precision)
        d = (Y55.to(torch.complex128) - outs[10.0][:4].to(torch.complex128)).abs().max().item()
        res["composition_5+5_vs_10_max_abs"] = d
    # the reference's compensated FP32 row sum (TwoProduct + Neumaier) at t = 1, for comparison
    torch.cuda.synchronize()
    Yc, infoc = chebyshev_evolve_batch(X, h, times[0], precision, accumulate="fp32c")
    Kc = infoc["chebyshev_degree"]
    res["fp32c_row_sums"] = {
        "t": times[0], "K": Kc, "ms_per_term": round(infoc["elapsed_s"] / Kc * 1e3, 4),
        "max_abs_diff_vs_fp64_row_sums": (Yc.to(torch.complex128) -
                                          outs[times[0]].to(torch.complex128)).abs().max().item()}
    del Yc
    # one right-hand side (the single-state chebyshev_evolve / run_hamiltonian
    # path) at t = 1: the first vector alone
    torch.cuda.synchronize()
    Y1, info1 = chebyshev_evolve_batch(X[0].contiguous(), h, times[0], precision, matrix_free=False)
    K1 = info1["chebyshev_degree"]
    per1 = 12.0 * h.nnz + 4.0 * (N + 1) + 5.0 * 8.0 * N
    res["single_rhs"] = {
        "t": times[0], "K": K1, "ms_per_term": round(info1["elapsed_s"] / K1 * 1e3, 4),
        "GB_per_s": round(per1 * K1 / info1["elapsed_s"] / 1e9, 1),
        "max_abs_diff_vs_batched": (Y1.to(torch.complex128) -
                                    outs[times[0]][0].to(torch.complex128)).abs().max().item()}
    # the same vector with matrix-free terms (the default for one state):
    # the algorithmic bytes are the five state vectors alone
    torch.cuda.synchronize()
    Ym, infom = chebyshev_evolve_batch(X[0].contiguous(), h, times[0], precision, matrix_free=True)
    Km = infom["chebyshev_degree"]
    perm = 5.0 * 8.0 * N
    res["single_rhs_matrix_free"] = {
        "t": times[0], "K": Km, "ms_per_term": round(infom["elapsed_s"] / Km * 1e3, 4),
        "GB_per_s": round(perm * Km / infom["elapsed_s"] / 1e9, 1),
        "max_abs_diff_vs_csr_terms": (Ym.to(torch.complex128) - Y1.to(torch.complex128)).abs().max().item()}
    del Y1, Ym
    # the opt-in two-pass term over the CSR split at bit n/2, at t = 1
    if n_rhs == 32:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        h.split()
        torch.cuda.synchronize()
        split_build_s = time.perf_counter() - t0
        Ys, infos = chebyshev_evolve_batch(X, h, times[0], precision, split=True)
        Ks = infos["chebyshev_degree"]
        res["split_term"] = {
            "t": times[0], "K": Ks, "ms_per_term": round(infos["elapsed_s"] / Ks * 1e3, 4),
            "csr_split_s": round(split_build_s, 4),
            "max_abs_diff_vs_one_pass": (Ys.to(torch.complex128) -
                                         outs[times[0]].to(torch.complex128)).abs().max().item()}
        del Ys
    return {"workload": f"{n}-qubit open chain ZZ + (XX+YY) + X, seed {seed}, CSR complex64 "
                        f"nnz {h.nnz}, {n_rhs} RHS, Bessel tail < {precision:g}",
            "row_sums": "FP64 sums of exact FP32 products (accumulate='fp64')",
            "csr_build_s": round(build_s, 2), "results": res}


# The reference's end-to-end table: BlackwellSimulator.run(circuit + measure_all,
# shots=8192), median of 5 after 3 warmups, milliseconds, RTX 5090
# (README.md:121-152).
REFERENCE_README_MS = {
    "GHZ": (3.8, 4.0, 4.2, 4.3, 4.9),
    "QFT": (4.0, 4.5, 5.8, 6.2, 7.8),
    "EfficientSU2": (4.1, 4.3, 4.8, 6.2, 27.0),
    "DeepRz": (6.6, 9.9, 12.8, 15.9, 25.3),
    "Toffoli": (3.8, 4.1, 10.7, 11.0, 5.4),
    "QV": (11.2, 8.9, 15.5, 29.0, 126.0),
}
AB_QUBITS = (4, 8, 12, 16, 20)


def ab_suite_bench(shots=8192, warmup=3, repeats=5):
    """The reference's own A/B performance suite, timed the way its harness
    times it (tests/ab_test.py:293-335): run(circuit + measure_all, 8192
    shots), median wall clock of 5 after 3 warmups."""
    from blackwell_backend import BlackwellSimulator, workloads as W
    sim = BlackwellSimulator("cuda:0")
    rows = {}
    ratios = []
    for name, build in W.AB_SUITE.items():
        for i, n in enumerate(AB_QUBITS):
            qc = build(n)
            qc.measure_all()
            for _ in range(warmup):
                sim.run(qc, shots=shots)
            ts = []
            for _ in range(repeats):
                t0 = time.perf_counter()
                sim.run(qc, shots=shots)
                ts.append(time.perf_counter() - t0)
            ms = float(np.median(ts)) * 1e3
            ref = REFERENCE_README_MS[name][i]
            ratios.append(ref / ms)
            rows[f"{name}-{n}"] = {"ms": round(ms, 3), "reference_ms": ref, "speedup": round(ref / ms, 2)}
    return {"workload": f"run(circuit + measure_all, shots={shots}), median of {repeats} after "
                        f"{warmup} warmups; reference_ms from its README (RTX 5090)",
            "geomean_speedup": float(np.exp(np.mean(np.log(ratios)))),
            "min_speedup": float(min(ratios)), "circuits": rows}


def host_cores():
    """(threads to use, description): the CPU share this process may use (its
    affinity mask, capped by OMP_NUM_THREADS when set), with the CPU model and
    the ISA level of the oracle build that runs on it."""
    from oracle import build_ref
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, omp) if omp > 0 else avail
    return threads, {"cpu_model": build_ref.cpu_model(), "isa": build_ref.host_isa(),
                     "affinity_cpus": avail, "os_cpu_count": os.cpu_count(),
                     "omp_num_threads_env": omp or None}


def cpu_leg(sim, n, depth, seed, time_it=True):
    """The CPU leg: the FP64 C/OpenMP statevector (oracle/sv_fp64.c) checks
    the GPU result of the same circuit family at 22 qubits ("parity") and,
    with time_it, is timed on the host cores as the CPU baseline, beside the
    CPU path on the 12-qubit ladder and the Chebyshev workload."""
    out = {}
    try:
        out["parity"] = fidelity_check(sim, 22, 22, seed, ("plain", "blocks"))
    except Exception as e:  # reported, not hidden
        out["parity"] = {"error": repr(e)}
    if time_it:
        try:
            out["cpu_baseline"] = cpu_baseline(n, depth, seed)
        except Exception as e:
            out["cpu_baseline"] = {"error": repr(e)}
        out["cpu_legs"] = {}
        for key, fn in (("c1_ladder_12q", lambda: cpu_c1(sim)), ("c5_chebyshev_22q", cpu_c5)):
            try:
                log(f"[bench] CPU leg {key}")
                out["cpu_legs"][key] = fn()
            except Exception as e:
                out["cpu_legs"][key] = {"error": repr(e)}
    return out


def cpu_baseline(n, depth, seed, budget_s=15.0):
    """FP64 C/OpenMP statevector on the host cores: amp-updates/s on a
    bounded prefix of the same circuit (gates applied until budget_s
    elapses).  Aer is not installed, so the FP64 C oracle stands in for the
    reference's Aer CPU path."""
    from blackwell_backend import workloads as W
    from oracle import build_ref
    lib = build_ref.load()
    threads, info = host_cores()
    threads = lib.sv_set_threads(threads)
    psi = np.empty(1 << n, dtype=np.complex128)
    lib.sv_init(psi.ctypes.data, n)
    layers = W.su4_layers(n, depth, seed)
    done = 0
    t0 = time.perf_counter()
    for pairs, us in layers:
        for (a, b), u in zip(pairs, us):
            v = np.empty(32)
            v[0::2] = u.reshape(-1).real
            v[1::2] = u.reshape(-1).imag
            lib.sv_apply_2q(psi.ctypes.data, n, v.ctypes.data_as(_pd()), a, b)
            done += 1
            if time.perf_counter() - t0 > budget_s:
                break
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    del psi
    return {"value": done * (1 << n) / el, "unit": "amp-updates/s", "cores": threads,
            "kind": "port", **info,
            "sample": f"first {done} of {n * depth // 2} gates of the {n}-qubit depth-{depth} "
                      f"SU(4) circuit (FP64 state {(16 << n) >> 30} GiB), FP64 C/OpenMP statevector "
                      f"(oracle/sv_fp64.c, built -march={lib.isa}), {el:.1f} s; Aer CPU is not installed"}


def cpu_c1(sim, warmup=3, repeats=5, seed=1234):
    """The 12-qubit rotation ladder (564 gates) on the CPU path, whole
    circuit, median of 5 after 3 warmups (tests/ab_test.py:154-165), beside
    the GPU's run_statevector of the same circuit."""
    from blackwell_backend import workloads as W
    from blackwell_backend.circuit import instructions
    from oracle import build_ref
    from oracle import statevector as O
    lib = build_ref.load()
    threads, info = host_cores()
    n = 12
    qc = W.c1_rotation_ladder(n, 12, seed)
    prog = []
    for name, qs, _, params in instructions(qc):
        if name == "cx":
            prog.append(("cx", (ctypes_int_array([qs[0]]), qs[1])))
        else:
            u = O.gate_1q(name, params)
            v = np.empty(8)
            v[0::2] = u.reshape(-1).real
            v[1::2] = u.reshape(-1).imag
            prog.append(("1q", (v, qs[0])))
    psi = np.empty(1 << n, dtype=np.complex128)

    def run_cpu():
        lib.sv_init(psi.ctypes.data, n)
        for kind, a in prog:
            if kind == "cx":
                lib.sv_apply_mcx(psi.ctypes.data, n, a[0], 1, a[1])
            else:
                lib.sv_apply_1q(psi.ctypes.data, n, a[0].ctypes.data_as(_pd()), a[1])

    # a 12-qubit state is small for a thread team: time one thread and all
    # of them, and keep the faster as the baseline
    per_threads = {}
    for nt in sorted({1, threads}):
        lib.sv_set_threads(nt)
        for _ in range(warmup):
            run_cpu()
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            run_cpu()
            ts.append(time.perf_counter() - t0)
        per_threads[nt] = float(np.median(ts))
    lib.sv_set_threads(threads)
    best = min(per_threads, key=per_threads.get)
    ref = psi.copy()
    for _ in range(warmup):
        sim.run_statevector(qc)
    gts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        sv, _ = sim.run_statevector(qc)
        gts.append(time.perf_counter() - t0)
    return {"workload": f"{n}-qubit rotation ladder, 12 layers, {len(prog)} gates, from |0..0>",
            "cpu_ms": round(per_threads[best] * 1e3, 3), "cpu_threads": best,
            "cpu_ms_by_threads": {str(k): round(v * 1e3, 3) for k, v in per_threads.items()},
            "cpu_kind": "FP64 C statevector (oracle/sv_fp64.c); Aer CPU is not installed",
            "gpu_run_statevector_ms": round(float(np.median(gts)) * 1e3, 3),
            "gpu_fidelity": O.fidelity(ref, sv), "gpu_max_abs_damp": O.max_abs_diff_phase_aligned(ref, sv),
            **info}


def ctypes_int_array(vals):
    import ctypes
    return (ctypes.c_int * len(vals))(*vals)


def cpu_c5(n=22, seed=5, t=1.0, n_rhs=1):
    """The Chebyshev workload's CPU path: scipy expm_multiply (FP64) of
    exp(-iHt) on the first checked right-hand side, on the CSR the GPU
    uses (copied to the host and widened to complex128), beside the GPU's
    time per right-hand side at 32 RHS."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import expm_multiply
    from blackwell_backend import workloads as W
    from blackwell_backend.hamiltonian import HamiltonianCSR
    from blackwell_backend.chebyshev import chebyshev_evolve_batch
    from oracle import statevector as O
    dev = torch.device("cuda", 0)
    h = HamiltonianCSR.from_sparse_pauli_op(W.heisenberg_chain(n, seed), device=dev)
    N = 1 << n
    rng = np.random.default_rng(seed + 1)
    X = (rng.standard_normal((32, N)) + 1j * rng.standard_normal((32, N))).astype(np.complex64)
    X /= np.linalg.norm(X, axis=1, keepdims=True).astype(np.float32)
    Xd = torch.from_numpy(X).to(dev)
    chebyshev_evolve_batch(Xd[:2], h, 0.1, 1e-7)
    Y, info = chebyshev_evolve_batch(Xd, h, t, 1e-7)
    gpu = Y[:n_rhs].cpu().numpy()
    A = sp.csr_matrix((h.values.cpu().numpy().astype(np.complex128), h.col_idx.cpu().numpy(),
                       h.row_ptr.cpu().numpy()), shape=(N, N))
    del Y, Xd
    torch.cuda.empty_cache()
    t0 = time.perf_counter()
    ref = expm_multiply(-1j * t * A, X[:n_rhs].astype(np.complex128).T).T
    cpu_s = time.perf_counter() - t0
    fid = [O.fidelity(ref[k], gpu[k]) for k in range(n_rhs)]
    return {"workload": f"exp(-iHt), {n}-qubit chain, t = {t:g}, nnz {h.nnz}",
            "cpu_kind": "scipy.sparse.linalg.expm_multiply, FP64, one thread (scipy's sparse products)",
            "cpu_s_per_rhs": round(cpu_s / n_rhs, 3), "cpu_rhs": n_rhs,
            "gpu_s_per_rhs_at_32": round(info["elapsed_s"] / 32, 5),
            "gpu_K": info["chebyshev_degree"],
            "gpu_fidelity_vs_expm": fid,
            "gpu_max_abs_damp_vs_expm": [float(np.max(np.abs(ref[k] - gpu[k]))) for k in range(n_rhs)]}


def c2_bench(n=28, depth=28, seed=2024, steps=5, warmup=2):
    """The 28-qubit random-SU(4) circuit (depth 28, 392 gates, 2 GiB):
    plain per-gate kernels and fused 5-qubit MFMA blocks."""
    from blackwell_backend import BlackwellSimulator, workloads as W
    sim = BlackwellSimulator("cuda:0")
    stream = torch.cuda.current_stream(sim.device)
    qc = W.su4_volume(n, depth, seed)
    gates = (n // 2) * depth
    out = {"workload": f"random-SU(4) volume circuit, {n} qubits, depth {depth}, {gates} gates"}
    for mode in ("plain", "blocks"):
        st, total, wall, kern, launches = run_steps(sim, qc, mode, steps, warmup, stream)
        del st
        per = float(np.median(kern)) / launches
        row = {"ms_per_step": round(total / steps * 1e3, 3),
               "amp_updates_per_s": gates * float(1 << n) / (total / steps), "launches": launches,
               "ms_per_launch": round(per * 1e3, 4),
               "GB_per_s_per_launch": round(16.0 * (1 << n) / per / 1e9, 1)}
        row["frac_hbm_peak_per_launch"] = round(row["GB_per_s_per_launch"] / HBM_PEAK_GBS, 4)
        out[mode] = row
    out["fused_speedup"] = out["plain"]["ms_per_step"] / out["blocks"]["ms_per_step"]
    torch.cuda.empty_cache()
    return out


def serving_bench(n_requests=4096, seed=99, shots=1024):
    """Mixed 10..24-qubit random-SU(4) request stream (depth n, 1024 shots):
    the continuous-batching server vs the plain one-circuit-at-a-time path.
    Capacity with all requests queued at t = 0, then Poisson arrivals at
    1.2x the plain capacity (latency p50/p99)."""
    from blackwell_backend import BlackwellSimulator, serving as S, workloads as W
    sim = BlackwellSimulator("cuda:0")
    ns = W.serving_qubit_counts(n_requests, seed)

    def reqs(rate):
        rng = np.random.default_rng(seed + 1)
        arr = np.cumsum(rng.exponential(1.0 / rate, n_requests)) if rate else np.zeros(n_requests)
        return [S.su4_request(i, float(arr[i]), int(ns[i]), shots, seed * 100003 + i)
                for i in range(n_requests)]

    # warm both paths (compiles kernels, allocator)
    S.serve_plain(reqs(None)[:8], sim, realtime=False)
    S.Server(sim).serve(reqs(None)[:8], realtime=False)
    plain_cap = S.serve_plain(reqs(None), sim, realtime=False)
    served_cap = S.Server(sim).serve(reqs(None), realtime=False)
    rate = 1.2 * plain_cap["circuits_per_s"]
    plain_rt = S.serve_plain(reqs(rate), sim, realtime=True)
    served_rt = S.Server(sim).serve(reqs(rate), realtime=True)
    fid = served_fidelity(sim, ns, seed, shots)
    return {
        "workload": f"{n_requests} requests, n = 10 + k, P(k) ~ 0.25*0.75^k (k <= 14), random-SU(4) "
                    f"depth n, {shots} shots each",
        "qubit_histogram": {int(k): int(v) for k, v in zip(*np.unique(ns, return_counts=True))},
        "capacity": {"served_circuits_per_s": served_cap["circuits_per_s"],
                     "plain_circuits_per_s": plain_cap["circuits_per_s"],
                     "speedup": served_cap["circuits_per_s"] / plain_cap["circuits_per_s"]},
        "poisson_rate_per_s": rate,
        "poisson_served": {k: served_rt[k] for k in ("circuits_per_s", "p50_latency_s", "p99_latency_s")},
        "poisson_plain": {k: plain_rt[k] for k in ("circuits_per_s", "p50_latency_s", "p99_latency_s")},
        "fidelity": fid,
    }


def served_fidelity(sim, ns, seed, shots, per_lane=16):
    """Re-serve a seeded sample of the stream's requests, 16 from each lane,
    as one mixed run with their final states copied back, and compare every
    state with the FP64 C statevector of the same gates (outside the timed
    runs)."""
    from blackwell_backend import serving as S, workloads as W
    from oracle import build_ref
    from oracle import statevector as O
    rng = np.random.default_rng(seed + 7)
    lds = [i for i in range(len(ns)) if ns[i] <= S.LDS_MAX_QUBITS]
    hbm = [i for i in range(len(ns)) if ns[i] > S.LDS_MAX_QUBITS]
    pick = sorted(list(rng.choice(lds, min(per_lane, len(lds)), replace=False)) +
                  list(rng.choice(hbm, min(per_lane, len(hbm)), replace=False)))
    reqs = [S.su4_request(int(i), 0.0, int(ns[i]), shots, seed * 100003 + int(i), keep_state=True)
            for i in pick]
    S.Server(sim).serve(reqs, realtime=False)
    lib = build_ref.load()
    lib.sv_set_threads(host_cores()[0])
    rows = []
    for r in reqs:
        n = r.n_qubits
        pairs, us = W.su4_arrays(n, n, r.seed)
        psi = np.empty(1 << n, dtype=np.complex128)
        lib.sv_init(psi.ctypes.data, n)
        for (a, b), u in zip(pairs, us):
            v = np.empty(32)
            v[0::2] = np.asarray(u).reshape(-1).real
            v[1::2] = np.asarray(u).reshape(-1).imag
            lib.sv_apply_2q(psi.ctypes.data, n, v.ctypes.data_as(_pd()), int(a), int(b))
        rows.append((n, r.lane, O.fidelity(psi, r.state), float(np.max(np.abs(psi - r.state)))))
    return {"requests": len(rows), "qubits": sorted({x[0] for x in rows}),
            "lanes": {ln: sum(1 for x in rows if x[1] == ln) for ln in ("lds", "hbm")},
            "min_fidelity": min(x[2] for x in rows), "max_abs_damp": max(x[3] for x in rows),
            "reference": "FP64 C statevector (oracle/sv_fp64.c) of the same gates"}


# extra sections: name -> (JSON key, function)
SECTIONS = {
    "per_gate": ("per_gate_30q", lambda: per_gate_sweep(30)),
    "c2": ("c2_28q", lambda: c2_bench()),
    "ab": ("ab_suite", lambda: ab_suite_bench()),
    "c3": ("c3_rotation_mcx_32q", lambda: c3_bench()),
    "c5": ("c5_chebyshev_22q", lambda: c5_bench()),
    "serving": ("serving", lambda: serving_bench()),
    "observables": ("observables_30q", lambda: observables_sweep(30)),
    "stride": ("stride_sweep_30q", lambda: stride_sweep(30)),
}
DEFAULT_SECTIONS = ("per_gate", "observables", "c2", "ab", "c3", "c5", "serving")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the extra sections and the timed CPU legs")
    ap.add_argument("--no-fused", action="store_true", help="skip the fused-block run of the headline circuit")
    ap.add_argument("--only", choices=sorted(SECTIONS),
                    help="run one extra section alone and print its JSON (for profiling it)")
    ap.add_argument("--pmc-calibration", action="store_true",
                    help="first run one plain device copy of a 2^qubits complex64 vector (a kernel "
                         "of known traffic, for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the statevector path here runs on a single MI355X)", file=sys.stderr)
        sys.exit(2)
    if args.only:
        key, fn = SECTIONS[args.only]
        log(f"[bench] section {args.only} alone")
        print(json.dumps({key: fn()}), flush=True)
        return
    from blackwell_backend import BlackwellSimulator, workloads as W
    n = args.qubits
    depth = n
    seed = 2024
    sim = BlackwellSimulator("cuda:0")
    stream = torch.cuda.current_stream(sim.device)
    qc = W.su4_volume(n, depth, seed)
    n_gates = (n // 2) * depth
    if args.pmc_calibration:
        src = torch.ones(1 << n, dtype=torch.complex64, device=sim.device)
        dst = torch.empty_like(src)
        dst.copy_(src)                       # reads and writes 8 * 2^n bytes
        torch.cuda.synchronize()
        del src, dst
        log(f"[bench] calibration copy of {(8 << n) >> 20} MiB done")
    log(f"[bench] compile + warmup plain ({n} qubits, {n_gates} gates)")
    st, total, wall, kern, launches = run_steps(sim, qc, "plain", args.steps, args.warmup, stream)
    del st
    ms_step = total / args.steps * 1e3
    updates = n_gates * float(1 << n)
    value = updates / (total / args.steps)
    per_launch_s = float(np.median(kern)) / launches
    achieved = 16.0 * (1 << n) / per_launch_s / 1e9
    log(f"[bench] plain: {ms_step:.1f} ms/step, {value:.3e} amp-updates/s, "
        f"{achieved:.0f} GB/s per 2Q launch")
    line = {
        "metric": f"amp-updates/s at {n} qubits (plain per-gate kernels, random-SU(4) depth {depth})",
        "value": value,
        "unit": "amp-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (fp32 compensated)",
        "data": f"synthetic (seeded Haar SU(4) gates, seed {seed})",
        "config": {"workload": f"random-SU(4) volume circuit, {n} qubits, depth {depth}, "
                               f"{n_gates} gates, complex64 state {(8 << n) >> 30} GiB, plain per-gate",
                   "qubits": n, "depth": depth, "gates": n_gates, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": GATE2Q_TRAFFIC_B_PER_AMP * (1 << n),
                     "traffic_source": GATE2Q_TRAFFIC_SOURCE,
                     "kernel": "k_gate2q (dense 4x4, 16 B/amp-update)",
                     "note": "achieved = 16 B x 2^n algorithmic bytes / median kernel-event time per "
                             "launch over the timed steps; the FETCH_SIZE / WRITE_SIZE counters see one "
                             "read and one write of the state per launch"},
    }
    if not args.no_fused:
        # the fusion width the tuner picks for this circuit (blocks of 4, 5 or 6
        # qubits, each timed from |0...0> with HIP events), then the timed steps
        from blackwell_backend.executor import block_width
        log("[bench] fused MFMA blocks: width tuner")
        fmode, width_ms = sim.compile(qc).tuned_block_width(sim.device)
        k = block_width(fmode)
        log(f"[bench] fused {k}-qubit MFMA blocks (tuner ms per run: "
            + ", ".join(f"k={w}: {t:.1f}" for w, t in sorted(width_ms.items())) + ")")
        st, ftotal, fwall, fkern, flaunch = run_steps(sim, qc, fmode, args.steps, args.warmup, stream)
        del st
        fused_value = updates / (ftotal / args.steps)
        blk_s = float(np.median(fkern)) / flaunch
        blk_bytes = 16.0 * (1 << n)
        blk_flops = 8.0 * (1 << k) * (1 << n)
        log(f"[bench] fused: {ftotal / args.steps * 1e3:.1f} ms/step, {fused_value:.3e} gate-amp-updates/s,"
            f" {flaunch} blocks")
        line["fused"] = {"value": fused_value, "unit": "gate-equivalent amp-updates/s",
                         "ms_per_step": ftotal / args.steps * 1e3, "blocks": flaunch,
                         "block_qubits": k,
                         "tuner_ms_per_run_by_width": {str(w): round(t, 3) for w, t in sorted(width_ms.items())},
                         "speedup_vs_plain": fused_value / value,
                         "block_GB_per_s": blk_bytes / blk_s / 1e9,
                         "block_mfma_TFLOP_per_s": blk_flops / blk_s / 1e12,
                         "frac_hbm": blk_bytes / blk_s / 1e9 / HBM_PEAK_GBS,
                         "frac_mfma_fp32": blk_flops / blk_s / 1e12 / MFMA_FP32_PEAK_TFLOPS}
    torch.cuda.empty_cache()
    if not args.no_extras:
        for name in DEFAULT_SECTIONS:
            key, fn = SECTIONS[name]
            try:
                log(f"[bench] {key}")
                line[key] = fn()
            except Exception as e:  # reported, not hidden
                line[key] = {"error": repr(e)}
            torch.cuda.empty_cache()
    log("[bench] CPU leg: FP64 C/OpenMP statevector (parity at 22 qubits, then the timed baselines)")
    line.update(cpu_leg(sim, n, depth, seed, time_it=not args.no_extras))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
