"""Parity at the benchmark workloads' own sizes (28-qubit SU(4), 32-qubit
rotation + MCX ansatz), through properties that need no FP64 statevector of
that size:
  * a circuit followed by its inverse returns |0...0>: |<0|psi>|^2 >= 1 - 1e-6
    and the FP64 norm stays within 1e-6 of 1;
  * a layer of permutation gates (CNOT ladder, MCX with controls on window,
    lane and float4 bits) followed by its reverse restores a random state
    bit for bit;
  * the FP64 leaf sampler at 32 qubits returns only indices whose amplitude
    is nonzero and reproduces them for the same uniforms;
  * at 26 qubits (just above the MALL size) a gate sequence through the
    streaming forms of the kernels matches the FP64 oracle directly.
The tests keep every state on the device (no 2^32-element host copies)."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim():
    from blackwell_backend import BlackwellSimulator
    return BlackwellSimulator("cuda:0")


def _inverse_appended(qc):
    """qc followed by its inverse (reverse order, each gate inverted)."""
    from blackwell_backend.circuit import QuantumCircuit, instructions
    out = QuantumCircuit(qc.num_qubits, name=qc.name + "_and_inverse")
    ins = instructions(qc)
    for stage in (ins, list(reversed(ins))):
        inverse = stage is not ins
        for name, qs, _, params in stage:
            if name == "unitary":
                u = np.asarray(params[0])
                out.unitary(u.conj().T if inverse else u, qs)
            elif name in ("rx", "ry", "rz"):
                getattr(out, name)(-params[0] if inverse else params[0], qs[0])
            elif name == "cx":
                out.cx(qs[0], qs[1])
            elif name == "mcx":
                out.mcx(qs[:-1], qs[-1])
            else:
                raise AssertionError(f"no inverse rule for {name}")
    return out


def _run_on_device(sim, qc, mode):
    import blackwell_kernels as bk
    cc = sim.compile(qc)
    plan = cc.plan(mode, sim.device)
    st = bk.state_init(qc.num_qubits, device=sim.device)
    plan.run(st)
    return st


@pytest.mark.parametrize("mode", ["plain", "blocks"])
def test_su4_28q_circuit_then_inverse(sim, mode):
    import blackwell_kernels as bk
    from blackwell_backend import workloads as W
    print(f"[fullsize] su4 28q forward+inverse mode={mode}", flush=True)
    st = _run_on_device(sim, _inverse_appended(W.su4_volume(28, 28, 2024)), mode)
    p0 = abs(complex(st[0].item())) ** 2
    nrm = float(bk.norm2(st))
    del st
    torch.cuda.empty_cache()
    assert p0 >= 1 - 1e-6 and abs(nrm - 1) <= 1e-6, (p0, nrm)


def test_c3_32q_circuit_then_inverse(sim):
    import blackwell_kernels as bk
    from blackwell_backend import workloads as W
    print("[fullsize] rotation+MCX 32q forward+inverse (plain)", flush=True)
    st = _run_on_device(sim, _inverse_appended(W.c3_rotation_mcx(32, 8, 77)), "plain")
    p0 = abs(complex(st[0].item())) ** 2
    nrm = float(bk.norm2(st))
    del st
    torch.cuda.empty_cache()
    assert p0 >= 1 - 1e-6 and abs(nrm - 1) <= 1e-6, (p0, nrm)


def test_permutation_layer_32q_round_trip_bit_exact():
    import blackwell_kernels as bk
    n = 32
    print("[fullsize] permutation layer 32q round trip", flush=True)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(7)
    st = torch.randn(1 << n, dtype=torch.complex64, device="cuda:0", generator=g)
    ref = st.clone()
    rng = np.random.default_rng(3)
    gates = [([q], q + 1) for q in range(n - 1)]                       # CNOT ladder
    gates += [([0], 17), ([3], 29), ([2, 11, 25], 19), ([11, 19, 25], 2), ([31, 30, 29], 0),
              ([5, 6], 31), ([1, 2, 3, 4, 20, 21, 22], 12)]
    for _ in range(6):
        qs = [int(q) for q in rng.choice(n, 4, replace=False)]
        gates.append((qs[:-1], qs[-1]))
    for ctl, t in gates:
        bk.apply_mcx(st, ctl, t)
    for ctl, t in reversed(gates):
        bk.apply_mcx(st, ctl, t)
    same = bool(torch.equal(st.view(torch.int32), ref.view(torch.int32)))
    del st, ref
    torch.cuda.empty_cache()
    assert same


def test_sampler_32q_support_and_determinism():
    import blackwell_kernels as bk
    n = 32
    print("[fullsize] sampler 32q", flush=True)
    st = bk.state_init(n, device="cuda:0")
    # GHZ-like support {0, 2^n - 1, 2^31} with known weights 1/2, 1/4, 1/4
    st[0] = math.sqrt(0.5)
    st[(1 << n) - 1] = math.sqrt(0.25)
    st[1 << 31] = -math.sqrt(0.25)
    u = bk.uniforms_splitmix(2024, 1 << 20, "cuda:0")
    a = bk.sample_with_uniforms(st, u)
    b = bk.sample_with_uniforms(st, u)
    vals, cnt = torch.unique(a, return_counts=True)
    del st
    torch.cuda.empty_cache()
    assert torch.equal(a, b)
    assert set(vals.tolist()) <= {0, (1 << n) - 1, 1 << 31}
    frac = {int(v): int(c) / (1 << 20) for v, c in zip(vals.tolist(), cnt.tolist())}
    assert abs(frac.get(0, 0) - 0.5) < 0.005 and abs(frac.get(1 << 31, 0) - 0.25) < 0.005


def test_pauli_conjugation_30q():
    """<Q psi|P|Q psi> = (-1)^[P,Q] <psi|P|psi> on an 8 GiB state, and Q twice
    restores psi bit for bit."""
    import blackwell_kernels as bk
    n = 30
    print("[fullsize] Pauli conjugation 30q", flush=True)
    g = torch.Generator(device="cuda:0")
    g.manual_seed(11)
    st = torch.randn(1 << n, dtype=torch.complex64, device="cuda:0", generator=g)
    ref = st.clone()
    rng = np.random.default_rng(12)
    xs = [0, 0, 1 << 29, 0b101 << 12, (1 << 30) - 1] + [int(v) for v in rng.integers(0, 1 << n, 5)]
    zs = [1 << 5, (1 << 30) - 1, 1 << 29, 0b11 << 12, 0] + [int(v) for v in rng.integers(0, 1 << n, 5)]
    e0 = bk.expectation_paulis(st, xs, zs).cpu().numpy()
    qx, qz = (1 << 29) | (1 << 3) | 1, (1 << 17) | (1 << 3)
    bk.apply_pauli_string(st, qx, qz)
    e1 = bk.expectation_paulis(st, xs, zs).cpu().numpy()
    bk.apply_pauli_string(st, qx, qz)
    same = bool(torch.equal(st.view(torch.int32), ref.view(torch.int32)))
    del st, ref
    torch.cuda.empty_cache()
    sign = np.array([(-1) ** ((bin(x & qz).count("1") + bin(z & qx).count("1")) & 1)
                     for x, z in zip(xs, zs)])
    scale = float(1 << n)   # |<P>| <= ||psi||^2 ~ 2^n for this unnormalised state
    assert same
    assert np.max(np.abs(e1 - sign * e0)) <= 1e-12 * scale, (e0, e1, sign)


def test_33_qubit_state_64bit_paths():
    """A 33-qubit state (64 GiB, beyond 32-bit amplitude indices) through the
    64-bit index variants: init, 1Q on the top qubit, controlled gates across
    bit 32, a dense 2Q gate and its inverse, diagonal and phase gates, the
    FP64 norm, a Z expectation, the sampler and a 5-qubit MFMA block."""
    import blackwell_kernels as bk
    n = 33
    print("[fullsize] 33-qubit state, 64-bit index paths", flush=True)
    dev = "cuda:0"
    st = bk.state_init(n, device=dev)
    assert abs(bk.norm2(st) - 1) < 1e-12 and complex(st[0].item()) == 1
    h = bk.hadamard_gate(dev)
    u4 = torch.from_numpy(np.linalg.qr(np.random.default_rng(5).standard_normal((4, 4))
                                       + 1j * np.random.default_rng(6).standard_normal((4, 4)))[0]
                          .astype(np.complex64)).to(dev)
    bk.apply_gate(st, h, 32)                                   # (|0> + |2^32>) / sqrt 2
    assert abs(bk.expectation_z(st, 1 << 32)) < 1e-6
    u = bk.uniforms_splitmix(9, 1 << 16, dev)
    smp = bk.sample_with_uniforms(st, u)
    vals, cnt = torch.unique(smp, return_counts=True)
    assert set(vals.tolist()) == {0, 1 << 32} and abs(cnt[0].item() / (1 << 16) - 0.5) < 0.01
    bk.apply_mcx(st, [32], 0)                                  # CNOT across bit 32
    bk.apply_gate_2q(st, u4, 5, 32)
    bk.apply_gate_2q(st, u4.conj().T.contiguous(), 5, 32)      # and its inverse
    bk.apply_diagonal(st, bk.rz_gate_diag(0.7, dev), 32)
    bk.apply_diagonal(st, bk.rz_gate_diag(-0.7, dev), 32)
    bk.apply_diagonal_phase(st, 1j, 32)
    bk.apply_diagonal_phase(st, -1j, 32)
    from helpers import rand_unitary
    ub = rand_unitary(32, 33).astype(np.complex64)
    qs = [0, 7, 15, 24, 32]                                    # MFMA block (LDS-staged above 32 qubits)
    bk.apply_block(st, torch.from_numpy(ub).to(dev), qs)
    bk.apply_block(st, torch.from_numpy(ub.conj().T.copy()).to(dev), qs)
    bk.apply_mcx(st, [32], 0)
    bk.apply_gate(st, h, 32)
    p0 = abs(complex(st[0].item())) ** 2
    nrm = bk.norm2(st)
    del st
    torch.cuda.empty_cache()
    assert p0 >= 1 - 1e-6 and abs(nrm - 1) <= 1e-6, (p0, nrm)


def test_hbm_streaming_paths_26q_match_oracle():
    """At 26 qubits (512 MiB) the state is larger than the 256 MB MALL, so the
    kernels take their nontemporal forms and the pass order across
    launches (alternating direction, cached stores on each pass's tail).  A sequence
    covering those paths is compared with the FP64 oracle on the same
    complex64 input (compensated FP32 bar, relative to max |a|), and the
    permutation gates in it are checked bit for bit on their own."""
    import blackwell_kernels as bk
    from oracle import statevector as SV
    from helpers import rand_state, rand_unitary, close
    n = 26
    print(f"[fullsize] streaming paths at {n} qubits vs the FP64 oracle", flush=True)
    dev = torch.device("cuda", 0)
    psi = rand_state(n, 26)
    st = torch.from_numpy(psi).to(dev)
    ref = psi.astype(np.complex128)
    u2 = [rand_unitary(2, 100 + k).astype(np.complex64) for k in range(5)]
    u4 = [rand_unitary(4, 200 + k).astype(np.complex64) for k in range(3)]
    d2 = np.array([np.exp(-0.4j), np.exp(0.7j)], dtype=np.complex64)
    ops = [("1q", u2[0], 7), ("1q", u2[1], 12), ("1q", u2[2], 25), ("1q", u2[3], 3),
           ("2q", u4[0], (7, 21)), ("2q", u4[1], (3, 20)), ("2q", u4[2], (0, 1)),
           ("diag", d2, 5), ("phase", np.complex64(np.exp(0.3j)), 18),
           ("mcx", None, ([17], 9)), ("mcx", None, ([2, 11, 25], 19)),
           ("swap", None, (4, 23)), ("mcg", u2[4], ([20], 8))]
    for kind, m, q in ops:
        if kind == "1q":
            bk.apply_gate(st, torch.from_numpy(m).to(dev), q)
            ref = SV.apply_1q(ref, m, q)
        elif kind == "2q":
            bk.apply_gate_2q(st, torch.from_numpy(m).to(dev), *q)
            ref = SV.apply_2q_kernel(ref, m, *q)
        elif kind == "diag":
            bk.apply_diagonal(st, torch.from_numpy(m).to(dev), q)
            ref = SV.apply_diag(ref, m, q)
        elif kind == "phase":
            bk.apply_diagonal_phase(st, complex(m), q)
            ref = SV.apply_diag_phase(ref, complex(m), q)
        elif kind == "mcx":
            bk.apply_mcx(st, *q)
            ref = ref[SV.mcx_perm(n, *q)]
        elif kind == "swap":
            bk.apply_mcswap(st, [], *q)
            ref = ref[SV.mcswap_perm(n, [], *q)]
        else:
            bk.apply_mcgate(st, torch.from_numpy(m).to(dev), q[0], q[1])
            ref = SV.apply_mc(ref, m, q[0], [q[1]])
    got = st.cpu().numpy()
    ok, err = close(ref, got, rtol=2e-5, atol=0)
    assert ok, err
    assert abs(float(bk.norm2(st)) - float(np.vdot(ref, ref).real)) < 1e-5
    # permutations alone are bit-exact on this path
    perm_in = torch.from_numpy(psi).to(dev)
    bk.apply_mcx(perm_in, [17], 9)
    bk.apply_mcx(perm_in, [2, 11, 25], 19)
    bk.apply_mcswap(perm_in, [], 4, 23)
    want = psi[SV.mcx_perm(n, [17], 9)]
    want = want[SV.mcx_perm(n, [2, 11, 25], 19)]
    want = want[SV.mcswap_perm(n, [], 4, 23)]
    assert np.array_equal(perm_in.cpu().numpy().view(np.uint32), want.view(np.uint32))
    del st, perm_in
    torch.cuda.empty_cache()


def test_chebyshev_22q_32rhs_matches_fp64():
    """The Chebyshev workload at its own size: the 22-qubit open chain
    (ZZ + XX + YY + X, seeded couplings), 32 normalised complex-Gaussian
    right-hand sides evolved together, Bessel tail < 1e-7.
      * t = 1: the first 4 against the FP64 Chebyshev oracle on host cores
        (oracle/cheb_fp64.c, matrix elements from the Pauli strings, tail
        < 1e-13; pinned against expm_multiply in tests/test_oracle.py):
        fidelity >= 1 - 1e-7 and max |d amp| <= 1e-5 max |a| per vector.
      * t = 10 (K ~ 490, the longest recurrence): exp(-iH 5) applied twice
        equals exp(-iH 10) (fidelity >= 1 - 1e-7), and norm and energy
        <H> (FP64 host products on the GPU result) are conserved."""
    from blackwell_backend import workloads as W
    from blackwell_backend.chebyshev import chebyshev_evolve_batch
    from blackwell_backend.hamiltonian import HamiltonianCSR
    from oracle import hamiltonian as OH
    from oracle import statevector as SV
    n, R, seed = 22, 32, 5
    dev = torch.device("cuda", 0)
    op = W.heisenberg_chain(n, seed)
    h = HamiltonianCSR.from_sparse_pauli_op(op, device=dev)
    N = 1 << n
    rng = np.random.default_rng(seed + 1)
    X = (rng.standard_normal((R, N)) + 1j * rng.standard_normal((R, N))).astype(np.complex64)
    X /= np.linalg.norm(X, axis=1, keepdims=True).astype(np.float32)
    Xd = torch.from_numpy(X).to(dev)
    print(f"[fullsize] chebyshev {n}q, {R} RHS, t = 1 on the GPU", flush=True)
    Y1, info = chebyshev_evolve_batch(Xd, h, 1.0, 1e-7)
    got = Y1[:4].cpu().numpy().astype(np.complex128)
    print(f"[fullsize] K = {info['chebyshev_degree']}; FP64 oracle on 4 RHS", flush=True)
    ref, K64 = OH.chebyshev_evolve_c(op.labels, op.coeffs, n, X[:4].astype(np.complex128), 1.0,
                                     h.spectral_bound, precision=1e-13)
    for k in range(4):
        f = SV.fidelity(ref[k], got[k])
        d = float(np.max(np.abs(ref[k] - got[k])))
        print(f"[fullsize]   rhs {k}: fidelity 1 - {1 - f:.3e}, max|d amp| {d:.3e} "
              f"(max|a| {np.max(np.abs(ref[k])):.3e})", flush=True)
        assert f >= 1 - 1e-7, (k, f)
        assert d <= 1e-5 * np.max(np.abs(ref[k])), (k, d)
    del Y1
    print("[fullsize] t = 5, 5 + 5 and 10 on the GPU", flush=True)
    Y10, info10 = chebyshev_evolve_batch(Xd, h, 10.0, 1e-7)
    X4 = Xd[:4].contiguous()
    Y5, _ = chebyshev_evolve_batch(X4, h, 5.0, 1e-7)
    Y55, _ = chebyshev_evolve_batch(Y5, h, 5.0, 1e-7)
    a10 = Y10[:4].cpu().numpy().astype(np.complex128)
    a55 = Y55.cpu().numpy().astype(np.complex128)
    nrm = np.linalg.norm(Y10.cpu().numpy().astype(np.complex128), axis=1)
    e0 = np.real(np.sum(np.conj(X[:4].astype(np.complex128)) *
                        OH.matvec_c(op.labels, op.coeffs, n, X[:4].astype(np.complex128)), axis=1))
    e10 = np.real(np.sum(np.conj(a10) * OH.matvec_c(op.labels, op.coeffs, n, a10), axis=1))
    print(f"[fullsize] K(10) = {info10['chebyshev_degree']}; max norm dev {np.max(np.abs(nrm - 1)):.3e}; "
          f"max energy drift {np.max(np.abs(e10 - e0)):.3e}", flush=True)
    for k in range(4):
        f = SV.fidelity(a55[k], a10[k])
        print(f"[fullsize]   rhs {k}: exp(-iH5)^2 vs exp(-iH10) fidelity 1 - {1 - f:.3e}, "
              f"max|d| {np.max(np.abs(a55[k] - a10[k])):.3e}", flush=True)
        assert f >= 1 - 1e-7, (k, f)
    assert np.max(np.abs(nrm - 1)) <= 1e-5
    assert np.max(np.abs(e10 - e0)) <= 1e-5 * np.max(np.abs(e0) + 1)
    del Y10, Y5, Y55, Xd
    torch.cuda.empty_cache()


def test_mfma_block_variants_26q_match_oracle():
    """Above the MALL size (26 qubits, 512 MiB) each 5-qubit block kernel
    takes its streaming form.  One block per dispatch case, applied in
    sequence and compared with the FP64 oracle (row bit j <-> qubits[j])
    after every block, so a transposed, conjugated or mis-ordered block is
    seen directly rather than through a round trip:
      (5, 9, 14, 20, 25) and the same set unsorted: 16-byte-lane kernel,
          nontemporal, no block qubit in 0..3;
      (2, 7, 11, 18, 24): 16-byte-lane kernel, lowest block qubit in 1..3;
      (0, 6, 13, 19, 23): 16-byte-lane kernel with qubit 0 as row bit 0;
      (1, 3, 4, 17, 22) and (22, 4, 1, 17, 3): LDS-staged kernel
          (three block qubits among 0..4), nontemporal;
      (20, 7, 25, 11, 16): 16-byte-lane kernel with every block qubit >= 7,
          unsorted, nontemporal."""
    import blackwell_kernels as bk
    from oracle import statevector as SV
    from helpers import rand_state, rand_unitary, close
    n = 26
    dev = torch.device("cuda", 0)
    psi = rand_state(n, 2626)
    st = torch.from_numpy(psi).to(dev)
    ref = psi.astype(np.complex128)
    blocks = [(5, 9, 14, 20, 25), (25, 5, 20, 9, 14), (2, 7, 11, 18, 24), (0, 6, 13, 19, 23),
              (1, 3, 4, 17, 22), (22, 4, 1, 17, 3), (20, 7, 25, 11, 16)]
    for k, qs in enumerate(blocks):
        print(f"[fullsize] MFMA block {qs} at {n} qubits vs the FP64 oracle", flush=True)
        u = rand_unitary(32, 500 + k).astype(np.complex64)
        bk.apply_block(st, torch.from_numpy(u).to(dev), qs)
        ref = SV.apply_matrix(ref, u.astype(np.complex128), list(qs))
        ok, err = close(ref, st.cpu().numpy(), rtol=8e-6 * (k + 1), atol=0)
        assert ok, (qs, err)
    del st
    torch.cuda.empty_cache()
