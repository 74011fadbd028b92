"""Kernel-level parity of libqsim_gfx950.so against the FP64 oracle.

Every test feeds the same complex64 input to the GPU kernel and to the
oracle (up-cast to complex128) and bounds the difference:
  * permutations (X, CNOT, Toffoli, MCX, SWAP, Fredkin, state init): bit-exact;
  * compensated dense/diagonal gates: max |delta| <= 1e-6 * max |a|
    (coefficients are rounded to FP32 on the way in, ~6e-8 each);
  * MFMA blocks (FP32 accumulation over 64 terms): <= 8e-6 * max |a|;
  * sampling: identical indices for identical uniforms except at CDF ties
    (>= 99.9 %), plus a chi-square test.
"""
import itertools

import numpy as np
import pytest
import torch

from oracle import statevector as O
from helpers import rand_state, rand_unitary, to_dev

pytestmark = pytest.mark.gpu


def _rel_err(ref, got):
    ref = np.asarray(ref, dtype=np.complex128)
    got = np.asarray(got, dtype=np.complex128)
    return float(np.max(np.abs(ref - got)) / max(np.max(np.abs(ref)), 1e-30))


def _bits(x):
    return np.ascontiguousarray(x).view(np.uint32)


def test_state_init(bk, dev):
    for n in (1, 2, 3, 7, 12, 16):
        st = bk.state_init(n, device=dev).cpu().numpy()
        ref = np.zeros(1 << n, dtype=np.complex64)
        ref[0] = 1
        assert np.array_equal(_bits(st), _bits(ref)), n
    sb = bk.state_init_batched(5, 4, device=dev).cpu().numpy()
    ref = np.zeros((5, 16), dtype=np.complex64)
    ref[:, 0] = 1
    assert np.array_equal(_bits(sb), _bits(ref))


@pytest.mark.parametrize("n", [1, 2, 3, 5, 10])
def test_gate_1q_all_targets(bk, dev, n):
    for t in range(n):
        psi = rand_state(n, 100 * n + t)
        u = rand_unitary(2, t + 7)
        st = to_dev(psi, dev)
        bk.apply_gate(st, torch.from_numpy(u.astype(np.complex64)).to(dev), t)
        ref = O.apply_1q(psi.astype(np.complex128), u.astype(np.complex64).astype(np.complex128), t)
        assert _rel_err(ref, st.cpu().numpy()) < 1e-6, (n, t)


def test_gate_1q_by_value_matches_device_matrix(bk, dev):
    from blackwell_kernels import _lib
    n = 11
    for t in (0, 1, 4, 10):
        psi = rand_state(n, t)
        u = rand_unitary(2, 3 + t).astype(np.complex64)
        a = to_dev(psi, dev)
        b = to_dev(psi, dev)
        bk.apply_gate(a, torch.from_numpy(u).to(dev), t)
        v = np.empty(8)
        v[0::2] = u.reshape(-1).real
        v[1::2] = u.reshape(-1).imag
        _lib.call("qsim_apply_gate_h", _lib.stream_handle(dev), _lib.ptr(b), n, _lib.f64_array(v), t)
        assert np.array_equal(_bits(a.cpu().numpy()), _bits(b.cpu().numpy()))


@pytest.mark.parametrize("n", [2, 3, 6])
def test_gate_2q_all_pairs(bk, dev, n):
    for q0, q1 in itertools.permutations(range(n), 2):
        psi = rand_state(n, 10 * q0 + q1)
        u = rand_unitary(4, q0 * 31 + q1).astype(np.complex64)
        st = to_dev(psi, dev)
        bk.apply_gate_2q(st, torch.from_numpy(u).to(dev), q0, q1)
        ref = O.apply_2q_kernel(psi.astype(np.complex128), u.astype(np.complex128), q0, q1)
        assert _rel_err(ref, st.cpu().numpy()) < 1e-6, (n, q0, q1)


def test_gate_2q_wide(bk, dev):
    n = 14
    for q0, q1 in [(0, 13), (13, 0), (1, 2), (7, 3), (12, 13), (0, 1), (7, 8), (13, 9), (8, 11)]:
        psi = rand_state(n, q0 + 17 * q1)
        u = rand_unitary(4, q0 + q1).astype(np.complex64)
        st = to_dev(psi, dev)
        bk.apply_gate_2q(st, torch.from_numpy(u).to(dev), q0, q1)
        ref = O.apply_2q_kernel(psi.astype(np.complex128), u.astype(np.complex128), q0, q1)
        assert _rel_err(ref, st.cpu().numpy()) < 1e-6


def test_diagonals(bk, dev):
    n = 9
    rng = np.random.default_rng(3)
    for t in range(n):
        psi = rand_state(n, t)
        d = np.exp(1j * rng.uniform(0, 6.28, 2)).astype(np.complex64)
        st = to_dev(psi, dev)
        bk.apply_diagonal(st, torch.from_numpy(d).to(dev), t)
        assert _rel_err(O.apply_diag(psi.astype(np.complex128), d.astype(np.complex128), t),
                        st.cpu().numpy()) < 1e-6
        ph = complex(np.exp(1j * rng.uniform(0, 6.28)))
        st = to_dev(psi, dev)
        bk.apply_diagonal_phase(st, ph, t)
        ref = O.apply_diag_phase(psi.astype(np.complex128), complex(np.complex64(ph)), t)
        got = st.cpu().numpy()
        assert _rel_err(ref, got) < 1e-6
        # the bit=0 half is untouched, bit for bit
        idx = np.arange(1 << n)
        zero = ((idx >> t) & 1) == 0
        assert np.array_equal(_bits(got[zero]), _bits(psi[zero]))
    for q0, q1 in [(0, 1), (1, 0), (3, 8), (8, 2)]:
        psi = rand_state(n, q0 * 9 + q1)
        d = np.exp(1j * rng.uniform(0, 6.28, 4)).astype(np.complex64)
        st = to_dev(psi, dev)
        bk.apply_diagonal_2q(st, torch.from_numpy(d).to(dev), q0, q1)
        ref = O.apply_diag_2q(psi.astype(np.complex128), d.astype(np.complex128), q0, q1)
        assert _rel_err(ref, st.cpu().numpy()) < 1e-6


def test_mcx_mcswap_bit_exact(bk, dev):
    n = 6
    psi = rand_state(n, 42)
    for k in range(0, 4):
        for combo in itertools.permutations(range(n), k + 1):
            ctl, t = list(combo[:-1]), combo[-1]
            st = to_dev(psi, dev)
            bk.apply_mcx(st, ctl, t)
            ref = psi[O.mcx_perm(n, ctl, t)]
            assert np.array_equal(_bits(st.cpu().numpy()), _bits(ref)), (ctl, t)
    for k in range(0, 3):
        for combo in itertools.permutations(range(n), k + 2):
            ctl, a, b = list(combo[:-2]), combo[-2], combo[-1]
            st = to_dev(psi, dev)
            bk.apply_mcswap(st, ctl, a, b)
            ref = psi[O.mcswap_perm(n, ctl, a, b)]
            assert np.array_equal(_bits(st.cpu().numpy()), _bits(ref)), (ctl, a, b)


def test_mcx_seven_controls_wide(bk, dev):
    n = 16
    psi = rand_state(n, 5)
    for ctl, t in [([1, 3, 5, 7, 9, 11, 13], 0), ([0, 2, 4, 6, 8, 10, 12], 15), ([15, 14], 13)]:
        st = to_dev(psi, dev)
        bk.apply_mcx(st, ctl, t)
        assert np.array_equal(_bits(st.cpu().numpy()), _bits(psi[O.mcx_perm(n, ctl, t)]))


@pytest.mark.parametrize("n", [8, 9, 11])
def test_controlled_1q_all_layouts(bk, dev, n):
    """Controlled X (bit-exact) and controlled dense 2x2 for targets in the
    float4, cross-lane and stream layouts, with controls on qubit 0, on
    lane bits (1..6) and on window bits (>= 7), against the oracle."""
    rng = np.random.default_rng(100 + n)
    psi = rand_state(n, 9 + n)
    u = rand_unitary(2, 3 + n).astype(np.complex64)
    ud = torch.from_numpy(u).to(dev)
    cases = [([], 0), ([], 5), ([], n - 1), ([0], 7), ([0], 3), ([3], 0), ([2], n - 1),
             ([n - 1], 2), ([n - 1], 0), ([0, n - 1], 4), ([1, 2, 3], 7), ([4, 5], 6),
             ([7], 6), ([0, 6, 7], 1)]
    for _ in range(12):
        k = int(rng.integers(0, 4))
        qs = [int(q) for q in rng.choice(n, k + 1, replace=False)]
        cases.append((qs[:-1], qs[-1]))
    for ctl, t in cases:
        print(f"[mc1] n={n} controls={ctl} target={t}", flush=True)
        st = to_dev(psi, dev)
        bk.apply_mcx(st, ctl, t)
        assert np.array_equal(_bits(st.cpu().numpy()), _bits(psi[O.mcx_perm(n, ctl, t)])), (ctl, t)
        st = to_dev(psi, dev)
        bk.apply_mcgate(st, ud, ctl, t)
        ref = O.apply_mc(psi.astype(np.complex128), u.astype(np.complex128), ctl, [t])
        assert _rel_err(ref, st.cpu().numpy()) < 1e-6, (ctl, t)


def test_mcgate_and_mc2qgate(bk, dev):
    n = 7
    for ctl, t in [([1], 0), ([0], 3), ([2, 5], 6), ([6, 0, 3], 1), ([], 4)]:
        psi = rand_state(n, t)
        u = rand_unitary(2, 11 + t).astype(np.complex64)
        st = to_dev(psi, dev)
        bk.apply_mcgate(st, torch.from_numpy(u).to(dev), ctl, t)
        ref = O.apply_mc(psi.astype(np.complex128), u.astype(np.complex128), ctl, [t])
        assert _rel_err(ref, st.cpu().numpy()) < 1e-6, (ctl, t)
    for ctl, a, b in [([0], 1, 2), ([3], 6, 0), ([1, 2], 0, 5), ([], 3, 4)]:
        psi = rand_state(n, a + b)
        u = rand_unitary(4, 5 + a).astype(np.complex64)
        st = to_dev(psi, dev)
        bk.apply_mc2qgate(st, torch.from_numpy(u).to(dev), ctl, a, b)
        ref = O.apply_mc2q_kernel(psi.astype(np.complex128), u.astype(np.complex128), ctl, a, b)
        assert _rel_err(ref, st.cpu().numpy()) < 1e-6, (ctl, a, b)


def test_argument_errors(bk, dev):
    st = bk.state_init(4, device=dev)
    x = bk.pauli_x_gate(dev)
    with pytest.raises(RuntimeError):
        bk.apply_gate(st, x, 4)
    with pytest.raises(RuntimeError):
        bk.apply_gate_2q(st, bk.swap_gate(dev), 2, 2)
    with pytest.raises(RuntimeError, match="controls"):
        bk.apply_mcgate(st, x, [0], 0)
    with pytest.raises(RuntimeError):
        bk.apply_mcx(st, list(range(8)), 9)
    with pytest.raises(RuntimeError):
        bk.apply_gate(st.to(torch.complex128), x, 0)


def test_gates_fused_and_reg_tiled(bk, dev):
    n = 13
    rng = np.random.default_rng(8)
    targets = [int(t) for t in rng.integers(0, n, 40)]
    gates = np.stack([rand_unitary(2, 50 + i) for i in range(len(targets))]).astype(np.complex64)
    psi = rand_state(n, 1)
    ref = psi.astype(np.complex128)
    for g, t in zip(gates, targets):
        ref = O.apply_1q(ref, g.astype(np.complex128), t)
    st = to_dev(psi, dev)
    bk.apply_gates_fused(st, torch.from_numpy(gates).to(dev), torch.tensor(targets))
    assert _rel_err(ref, st.cpu().numpy()) < 1e-5
    # reg-tiled: targets < 8, device targets like the reference
    t8 = [int(t) for t in rng.integers(0, 8, 30)]
    g8 = np.stack([rand_unitary(2, 90 + i) for i in range(len(t8))]).astype(np.complex64)
    ref = psi.astype(np.complex128)
    for g, t in zip(g8, t8):
        ref = O.apply_1q(ref, g.astype(np.complex128), t)
    st = to_dev(psi, dev)
    bk.apply_gates_reg_tiled(st, torch.from_numpy(g8).to(dev),
                             torch.tensor(t8, dtype=torch.int32, device=dev))
    assert _rel_err(ref, st.cpu().numpy()) < 1e-5
    # batched: same list on 3 states of 9 qubits
    ps = np.stack([rand_state(9, s) for s in range(3)])
    sts = to_dev(ps, dev)
    t7 = [t % 8 for t in t8]
    bk.apply_gates_batched_fused(sts, torch.from_numpy(g8).to(dev),
                                 torch.tensor(t7, dtype=torch.int32, device=dev))
    for i in range(3):
        ref = ps[i].astype(np.complex128)
        for g, t in zip(g8, t7):
            ref = O.apply_1q(ref, g.astype(np.complex128), t)
        assert _rel_err(ref, sts[i].cpu().numpy()) < 1e-5
    u = rand_unitary(2, 1).astype(np.complex64)
    sts = to_dev(ps, dev)
    bk.apply_gate_batched(sts, torch.from_numpy(u).to(dev), 5)
    for i in range(3):
        assert _rel_err(O.apply_1q(ps[i].astype(np.complex128), u.astype(np.complex128), 5),
                        sts[i].cpu().numpy()) < 1e-6


def test_reductions_and_measure(bk, dev):
    for n in (1, 3, 10, 15):
        psi = rand_state(n, n) * np.float32(1.7)
        st = to_dev(psi, dev)
        ref = np.sum(np.abs(psi.astype(np.complex128)) ** 2)
        assert abs(bk.norm2(st) - ref) <= 1e-12 * ref
        pr = bk.measure_probs(st).cpu().numpy()
        assert np.allclose(pr, np.abs(psi.astype(np.complex128)) ** 2, rtol=1e-6, atol=0)
        bk.renormalize(st)
        got = st.cpu().numpy().astype(np.complex128)
        assert abs(np.sum(np.abs(got) ** 2) - 1) < 1e-6
        # the output is psi / ||psi|| itself, not just some unit vector:
        # element-wise against the FP64 scaling, relative to the largest amplitude
        want = psi.astype(np.complex128) / np.sqrt(ref)
        assert np.max(np.abs(got - want)) <= 1e-6 * np.max(np.abs(want)), n
        for t in range(n):
            for u in (0.1, 0.77):
                s2 = to_dev(psi, dev)
                o = bk.measure_qubit(s2, t, u)
                ro, rs = O.measure(psi.astype(np.complex128), t, u)
                assert o == ro
                assert _rel_err(rs, s2.cpu().numpy()) < 1e-6
    # determinism: same state, same bits
    psi = rand_state(18, 3)
    a = bk.norm2(to_dev(psi, dev))
    b = bk.norm2(to_dev(psi, dev))
    assert a == b


@pytest.mark.parametrize("n", [1, 2, 5, 8, 9, 12, 16])
def test_sample_matches_oracle(bk, dev, n):
    rng = np.random.default_rng(n)
    psi = rand_state(n, 77 + n)
    if n >= 5:
        psi[rng.choice(1 << n, (1 << n) // 3, replace=False)] = 0  # zero-probability outcomes
    u = rng.random(20000)
    got = bk.sample_with_uniforms(to_dev(psi, dev), torch.from_numpy(u)).cpu().numpy()
    ref = O.sample(psi.astype(np.complex128), u)
    assert np.mean(got == ref) >= 0.999
    p = np.abs(psi.astype(np.complex128)) ** 2
    assert np.all(p[got] > 0)              # never a zero-probability index
    pval, _, _ = O.chi_square_p(got, p, len(u))
    assert pval > 1e-3


def test_uniforms_splitmix_bit_exact(bk, dev):
    """Device SplitMix64 uniforms equal the oracle's, bit for bit (including
    seeds that wrap 2^64 and the empty draw)."""
    for seed, count in ((0, 1), (12345, 4097), (2**64 - 3, 1000), (99 * 100003 + 7, 1024)):
        got = bk.uniforms_splitmix(seed, count, dev).cpu().numpy()
        ref = O.splitmix_uniforms(seed, count)
        assert got.dtype == np.float64 and np.array_equal(got.view(np.uint64), ref.view(np.uint64))
    assert bk.uniforms_splitmix(5, 0, dev).numel() == 0


def test_tile_program_random(bk, dev):
    from blackwell_backend import compiler as C
    from blackwell_backend.executor import tile_ops
    rng = np.random.default_rng(12)
    n = 14
    for trial in range(4):
        m = [10, 12, 13, 14][trial]
        extra = rng.choice(np.arange(3, n), m - 3, replace=False)
        mask = 0b111
        for q in extra:
            mask |= 1 << int(q)
        tq = [q for q in range(n) if mask >> q & 1]
        ops = []
        for i in range(25):
            kind = rng.choice(["m1", "d1", "ph", "m2", "x", "swap", "mk"])
            qs = [int(q) for q in rng.choice(tq, 5, replace=False)]
            if kind == "m1":
                ops.append(C.Op("m1", qs[:1], rand_unitary(2, i), controls=qs[1:1 + i % 3]))
            elif kind == "d1":
                ops.append(C.Op("d1", qs[:1], np.exp(1j * rng.uniform(0, 6, 2))))
            elif kind == "ph":
                ops.append(C.Op("ph", qs[:1], np.exp(1j * rng.uniform(0, 6, 1)), controls=qs[1:2]))
            elif kind == "m2":
                ops.append(C.Op("m2", qs[:2], rand_unitary(4, i)))
            elif kind == "x":
                ops.append(C.Op("x", qs[:1], controls=qs[1:1 + i % 4]))
            elif kind == "swap":
                ops.append(C.Op("swap", qs[:2], controls=qs[2:2 + i % 2]))
            else:
                k = 3 + i % 3
                ops.append(C.Op("mk", qs[:k], rand_unitary(1 << k, i)))
        psi = rand_state(n, trial)
        ref = psi.astype(np.complex128)
        for op in ops:
            m64 = op.full_matrix().astype(np.complex64).astype(np.complex128)
            ref = O.apply_matrix(ref, m64, list(op.all_qubits))
        arr, cf = tile_ops(ops, mask)
        st = to_dev(psi, dev)
        bk.run_tile_program(st, mask, arr, cf.astype(np.complex64))
        assert _rel_err(ref, st.cpu().numpy()) < 2e-5, trial


@pytest.mark.parametrize("n,qubits", [
    (10, [0, 1, 2, 3, 4]), (10, [9, 5, 0, 7, 2]), (12, [11, 10, 9, 8, 7]), (16, [3, 15, 8, 0, 11]),
    # 16-byte-lane kernel: qubit 0 outside the block (lowest block qubit >= 4 or in 1..3) and
    # qubit 0 as block row bit 0; n = 11 is the smallest state it takes (one set of 64 groups)
    (11, [4, 5, 6, 9, 10]), (14, [1, 6, 8, 12, 13]), (14, [2, 3, 7, 9, 13]), (14, [0, 5, 6, 9, 12]),
    (13, [0, 1, 7, 10, 12]), (17, [16, 12, 8, 5, 4]), (15, [0, 14, 3, 9, 6]), (11, [0, 6, 7, 9, 10]),
    # every block qubit >= 7 (rows of 512 contiguous bytes per set), unsorted qubits
    (13, [12, 7, 9, 8, 11]), (16, [15, 8, 12, 10, 9]), (12, [7, 8, 9, 10, 11])])
def test_block_mfma(bk, dev, n, qubits):
    psi = rand_state(n, n + qubits[0])
    u = rand_unitary(32, sum(qubits)).astype(np.complex64)
    st = to_dev(psi, dev)
    bk.apply_block(st, torch.from_numpy(u).to(dev), qubits)
    ref = O.apply_matrix(psi.astype(np.complex128), u.astype(np.complex128), qubits)
    assert _rel_err(ref, st.cpu().numpy()) < 8e-6


@pytest.mark.parametrize("n,qubits", [
    (11, [0, 1, 2, 3]), (11, [3, 1, 10, 6]), (14, [7, 12, 2, 13]), (12, [11, 8, 9, 10]),
    (13, [0, 1, 2, 3, 4, 5]), (13, [12, 0, 7, 3, 9, 5]), (15, [14, 13, 12, 8, 9, 10]),
    (16, [2, 15, 6, 11, 1, 8]),
])
def test_block_mfma_widths_4_and_6(bk, dev, n, qubits):
    """k = 4 (v_mfma_f32_16x16x4f32) and k = 6 (two 32-row tiles of
    v_mfma_f32_32x32x2f32) blocks from the LDS tile kernel against the FP64
    oracle (row bit j <-> qubits[j]), for sorted, unsorted, low and high
    qubit sets; a block of 64 terms accumulates in FP32 like k = 5."""
    k = len(qubits)
    psi = rand_state(n, 7 * n + sum(qubits))
    u = rand_unitary(1 << k, 3 * n + qubits[-1]).astype(np.complex64)
    st = to_dev(psi, dev)
    print(f"[mfma] n={n} k={k} qubits={qubits}", flush=True)
    bk.apply_block(st, torch.from_numpy(u).to(dev), qubits)
    ref = O.apply_matrix(psi.astype(np.complex128), u.astype(np.complex128), qubits)
    assert _rel_err(ref, st.cpu().numpy()) < 8e-6


def test_qv8_qv4(bk, dev):
    m, q, g = bk.generate_qv8_circuits(37, seed=3)
    got = bk.qv8_simulate(m.to(dev), q.to(dev), 37).cpu().numpy()
    for c in range(0, 37, 6):
        psi = O.zero_state(8)
        for k in range(g):
            u = (m[c, k, ..., 0].numpy() + 1j * m[c, k, ..., 1].numpy()).astype(np.complex128)
            psi = O.apply_matrix(psi, u, [int(q[c, k, 0]), int(q[c, k, 1])])
        assert np.allclose(got[c], O.probs(psi), atol=2e-6)
    gd, pid = bk.generate_qv4_circuits(40, rng=np.random.default_rng(2))
    got = bk.qv4_simulate_cuda(gd, pid).cpu().numpy()
    ref = bk.qv4_simulate_numpy(gd, pid)
    assert np.allclose(got, ref, atol=2e-6)


def test_run_instances_with_sampling(bk, dev):
    from blackwell_kernels import _lib
    from blackwell_backend import compiler as C
    from blackwell_backend.executor import tile_ops
    rng = np.random.default_rng(4)
    ns = [1, 3, 8, 10, 12, 14, 9, 13]
    all_ops, all_cf, inst, refs = [], [], [], []
    op_off = cf_off = st_off = shot_off = 0
    shots = 3000
    for i, n in enumerate(ns):
        ops = []
        for j in range(12):
            if n >= 2 and j % 2:
                a, b = rng.choice(n, 2, replace=False)
                ops.append(C.Op("m2", [int(a), int(b)], rand_unitary(4, 10 * i + j)))
            else:
                ops.append(C.Op("m1", [int(rng.integers(n))], rand_unitary(2, 10 * i + j)))
        arr, cf = tile_ops(ops, (1 << n) - 1)
        arr["coef"] += cf_off
        rec = np.zeros(1, dtype=_lib.INSTANCE_DTYPE)
        rec["state_off"] = st_off
        rec["seed"] = 1000 + i
        rec["shot_off"] = shot_off
        rec["op_off"] = op_off
        rec["n_ops"] = len(arr)
        rec["shots"] = shots
        rec["n_qubits"] = n
        rec["init"] = 1
        rec["want"] = 1
        inst.append(rec)
        all_ops.append(arr)
        all_cf.append(cf)
        psi = O.zero_state(n)
        for op in ops:
            psi = O.apply_matrix(psi, op.full_matrix().astype(np.complex64).astype(np.complex128),
                                 list(op.all_qubits))
        refs.append(psi)
        op_off += len(arr)
        cf_off += cf.size
        st_off += 1 << n
        shot_off += shots
    states = torch.zeros(st_off, dtype=torch.complex64, device=dev)
    out = torch.full((shot_off,), -1, dtype=torch.int64, device=dev)
    inst_h = np.concatenate(inst)
    ops_h = np.concatenate(all_ops)
    inst_d = torch.from_numpy(inst_h.view(np.uint8).copy()).to(dev)
    ops_d = torch.from_numpy(ops_h.view(np.uint8).copy()).to(dev)
    cf_d = torch.from_numpy(np.concatenate(all_cf).astype(np.complex64)).to(dev)
    _lib.call("qsim_run_instances", _lib.stream_handle(dev), _lib.ptr(states), st_off,
              inst_h.ctypes.data, _lib.ptr(inst_d), len(ns), ops_h.ctypes.data, _lib.ptr(ops_d),
              len(ops_h), _lib.ptr(cf_d), cf_d.numel(), _lib.ptr(out), shot_off, max(ns))
    st = states.cpu().numpy()
    sh = out.cpu().numpy()
    off = 0
    for i, n in enumerate(ns):
        got = st[off:off + (1 << n)]
        assert _rel_err(refs[i], got) < 2e-5, n
        u = O.splitmix_uniforms(1000 + i, shots)
        ref_idx = O.sample(got.astype(np.complex128), u)
        g = sh[i * shots:(i + 1) * shots]
        assert np.mean(g == ref_idx) >= 0.999, n
        off += 1 << n


def test_spmv_and_chebyshev_kernels(bk, dev):
    import scipy.sparse as sp
    rng = np.random.default_rng(0)
    N = 700
    A = sp.random(N, N, density=0.02, random_state=1, format="csr")
    A = (A + 1j * sp.random(N, N, density=0.02, random_state=2, format="csr")).tocsr()
    A.sort_indices()
    x = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    y = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    rp = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    ci = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    va = torch.from_numpy(A.data.astype(np.complex64)).to(dev)
    yd = to_dev(y, dev)
    bk.spmv_csr(rp, ci, va, to_dev(x, dev), yd, 0.5, -0.25, 0.1, 0.3)
    Av = A.data.astype(np.complex64).astype(np.complex128)
    A64 = sp.csr_matrix((Av, A.indices, A.indptr), shape=A.shape)
    ref = (0.5 - 0.25j) * (A64 @ x.astype(np.complex128)) + (0.1 + 0.3j) * y
    assert np.max(np.abs(ref - yd.cpu().numpy())) < 1e-5 * np.max(np.abs(ref))
    s, tp = to_dev(x, dev), to_dev(y, dev)
    tn = torch.empty_like(s)
    bk.chebyshev_step(s, tp, tn)
    assert np.allclose(tn.cpu().numpy(), 2 * x - y, atol=1e-6)
    acc = to_dev(y, dev)
    bk.chebyshev_accum(acc, s, 0.3, -0.7)
    assert np.allclose(acc.cpu().numpy(), y + (0.3 - 0.7j) * x, atol=1e-6)
    acc = to_dev(y, dev)
    bk.chebyshev_step_accum(s, tp, tn, acc, 0.2, 0.1)
    t2 = 2 * x - y
    assert np.allclose(tn.cpu().numpy(), t2, atol=1e-6)
    assert np.allclose(acc.cpu().numpy(), y + (0.2 + 0.1j) * t2, atol=1e-5)


def _ragged_csr(n_rows, seed):
    """Random complex CSR with ragged rows: empty rows, short rows and rows of
    150-300 nonzeros (longer than a wave's 128-nonzero staging chunk), sorted
    column indices."""
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    A = sp.random(n_rows, n_rows, density=0.03, random_state=seed, format="lil")
    A = A.astype(np.complex128)
    for r in (3, 4, 11, n_rows - 1):
        cols = rng.choice(n_rows, size=int(rng.integers(150, 300)), replace=False)
        A[r, cols] = rng.standard_normal(len(cols)) + 1.0
    A = A.tocsr()
    A.data = A.data + 1j * rng.standard_normal(A.data.shape)
    for r in (0, 7, n_rows - 2):
        A.data[A.indptr[r]:A.indptr[r + 1]] = 0
    A.eliminate_zeros()
    A.sort_indices()
    assert A.indptr[1] == 0 and max(np.diff(A.indptr)) > 128
    return A


@pytest.mark.parametrize("n_rows", [701, 1536, 4096])
@pytest.mark.parametrize("accumulate", ["fp64", "fp32c"])
def test_cheb_spmm_step_matches_fp64(bk, dev, accumulate, n_rows):
    """One fused Chebyshev term, qsim_cheb_spmm_step, against the same term in
    FP64 on the host: w = scale A T_k, T_{k+1} = w (first) or 2 w - T_{k-1},
    acc + comp += c T_{k+1}.  Ragged row lengths; 32 RHS, 5 RHS (the
    generic kernel) and 1 RHS (16 lanes per row, merged partial sums).  At 32 RHS, 701 rows (odd: the last wave has one row)
    run the two-rows-per-wave kernel, and 1536 rows (3 windows, dispatch
    order) and 4096 rows (8 windows, one per XCD) the 512-row LDS window
    kernel, whose entries mix window reads and global gathers.  The 32 RHS
    run also equals, bit for bit, the same term computed as two 16-RHS halves
    through the generic kernel (same summation order per element)."""
    from blackwell_backend.chebyshev import ACCUMULATE
    from blackwell_kernels._lib import call, ptr, stream_handle
    mode = ACCUMULATE[accumulate]
    A = _ragged_csr(n_rows, 5)
    N = A.shape[0]
    A64 = A.astype(np.complex64).astype(np.complex128)
    rp = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    ci = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    va = torch.from_numpy(A.data.astype(np.complex64)).to(dev)
    rng = np.random.default_rng(9)
    scale, c = 0.37, 0.3 - 0.45j
    # the bound on each element's error scales with sum_k |a_ik| |t_kj|
    absA = abs(A64)
    for R in (32, 5, 1):
        def cplx(*shape):
            return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
        T, Tp, acc0 = cplx(N, R), cplx(N, R), cplx(N, R)
        for first in (1, 0):
            tpd, accd = to_dev(Tp, dev), to_dev(acc0, dev)
            compd = torch.zeros_like(accd)
            call("qsim_cheb_spmm_step", stream_handle(dev), ptr(rp), ptr(ci), ptr(va), N, R, scale,
                 ptr(to_dev(T, dev)), ptr(tpd), ptr(accd), ptr(compd), c.real, c.imag, first, mode)
            w = scale * (A64 @ T.astype(np.complex128))
            tn_ref = w if first else 2 * w - Tp.astype(np.complex128)
            bound = scale * (absA @ np.abs(T.astype(np.complex128))) * 2 + np.abs(Tp)
            tn = tpd.cpu().numpy().astype(np.complex128)
            err = np.abs(tn - tn_ref)
            # FP64 sums: one rounding to complex64 (<= 2^-24 per component);
            # compensated FP32: a few roundings of the row's term magnitudes
            tol = 2.0 ** -23 * np.abs(tn_ref) + (1e-15 if accumulate == "fp64" else 4e-7) * bound
            assert np.all(err <= tol + 1e-30), (R, first, float(np.max(err - tol)))
            got_acc = accd.cpu().numpy().astype(np.complex128) + compd.cpu().numpy()
            want_acc = acc0.astype(np.complex128) + c * tn
            assert np.max(np.abs(got_acc - want_acc)) <= 1e-6 * np.max(np.abs(want_acc)), (R, first)
            if R == 32:
                halves = []
                for h in range(2):
                    sl = slice(16 * h, 16 * h + 16)
                    tph = to_dev(np.ascontiguousarray(Tp[:, sl]), dev)
                    acch = to_dev(np.ascontiguousarray(acc0[:, sl]), dev)
                    comph = torch.zeros_like(acch)
                    call("qsim_cheb_spmm_step", stream_handle(dev), ptr(rp), ptr(ci), ptr(va), N, 16,
                         scale, ptr(to_dev(np.ascontiguousarray(T[:, sl]), dev)), ptr(tph), ptr(acch),
                         ptr(comph), c.real, c.imag, first, mode)
                    halves.append(tph.cpu().numpy())
                both = np.concatenate(halves, axis=1)
                assert np.array_equal(both.view(np.uint32), tpd.cpu().numpy().view(np.uint32)), first


def _rand_pauli_sum(n, n_terms, seed):
    """Random Pauli labels over I/X/Y/Z with real coefficients; every X mask
    appears with several Z/Y patterns (groups of more than one term), one
    term is the identity, one pair of terms cancels on half the rows, and
    some X-only groups have the same entry in every row (one of them zero)."""
    rng = np.random.default_rng(seed)
    labels, coeffs = [], []
    for _ in range(n_terms):
        lab = "".join(rng.choice(list("IXYZ"), n, p=[0.55, 0.15, 0.15, 0.15]))
        labels.append(lab)
        coeffs.append(float(rng.standard_normal()))
        # a partner with the same X mask and a different Z/Y pattern
        twin = "".join({"X": "Y", "Y": "X", "Z": "I", "I": "Z"}[ch] if rng.random() < 0.3 else ch for ch in lab)
        labels.append(twin)
        coeffs.append(float(rng.standard_normal()))
    labels.append("I" * n)
    coeffs.append(0.75)
    labels += ["X" + "I" * (n - 1), "X" + "Z" + "I" * (n - 2)]   # 0.5 X (1 + Z): zero where bit n-2 is set
    coeffs += [0.5, 0.5]
    # groups whose entry is the same in every row (X only, no Y or Z): one
    # term, two terms, and two that cancel exactly
    labels += ["I" * (n - 2) + "XX", "I" * (n - 3) + "X" + "I" + "X", "I" * (n - 3) + "X" + "I" + "X",
               "I" * (n - 4) + "XX" + "I" * 2, "I" * (n - 4) + "XX" + "I" * 2]
    coeffs += [0.3, 0.2, -0.45, 0.6, -0.6]
    return labels, coeffs


@pytest.mark.parametrize("n", [6, 11, 14])
def test_cheb_pauli_step_matches_csr_and_fp64(bk, dev, n):
    """qsim_cheb_pauli_step (matrix-free, one state) against the same term in
    FP64 on the host over the oracle's Kronecker-built matrix, and against the
    CSR term (qsim_cheb_spmm_step, 1 RHS, FP64 row sums) on the device CSR of
    the same Pauli sum: equal up to the order of the FP64 row sum (one
    rounding to complex64), for the first term and a later one."""
    from oracle import hamiltonian as OH
    from blackwell_backend.hamiltonian import HamiltonianCSR, SparsePauliOp
    from blackwell_backend.chebyshev import ACCUMULATE
    from blackwell_kernels._lib import call, ptr, stream_handle
    labels, coeffs = _rand_pauli_sum(n, 3 * n, 100 + n)
    h = HamiltonianCSR.from_sparse_pauli_op(SparsePauliOp.from_list(list(zip(labels, coeffs))), device=dev)
    assert h.pauli is not None and h.pauli.n_groups > 1
    gc = h.pauli.gconst.cpu().numpy().reshape(-1, 2)
    assert np.isfinite(gc[:, 0]).any() and np.isnan(gc[:, 0]).any()   # both kinds of group occur
    A64 = OH.pauli_sparse(labels, coeffs)
    A64 = A64.astype(np.complex64).astype(np.complex128)     # the entries rounded as the device stores them
    absA = abs(A64)
    N = 1 << n
    rng = np.random.default_rng(n)
    scale, c = 0.21, 0.3 - 0.45j
    cplx = lambda: (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    T, Tp, acc0 = cplx(), cplx(), cplx()
    s = stream_handle(dev)
    for first in (1, 0):
        outs = {}
        for how in ("pauli", "pauli_no_const", "csr"):
            tpd, accd = to_dev(Tp, dev), to_dev(acc0, dev)
            compd = torch.zeros_like(accd)
            if how.startswith("pauli"):
                pa = h.pauli.step_args()
                if how == "pauli_no_const":
                    pa = pa[:-1] + (None,)           # every group evaluated per row
                call("qsim_cheb_pauli_step", s, n, *pa, scale, ptr(to_dev(T, dev)), ptr(tpd),
                     ptr(accd), ptr(compd), c.real, c.imag, first)
            else:
                call("qsim_cheb_spmm_step", s, ptr(h.row_ptr), ptr(h.col_idx), ptr(h.values), N, 1, scale,
                     ptr(to_dev(T, dev)), ptr(tpd), ptr(accd), ptr(compd), c.real, c.imag, first,
                     ACCUMULATE["fp64"])
            outs[how] = (tpd.cpu().numpy().astype(np.complex128),
                         accd.cpu().numpy().astype(np.complex128) + compd.cpu().numpy())
        w = scale * (A64 @ T.astype(np.complex128))
        tn_ref = w if first else 2 * w - Tp.astype(np.complex128)
        bound = scale * (absA @ np.abs(T.astype(np.complex128))) * 2 + np.abs(Tp)
        tol = 2.0 ** -23 * np.abs(tn_ref) + 1e-15 * bound + 1e-30
        for how, (tn, acc) in outs.items():
            assert np.all(np.abs(tn - tn_ref) <= tol), (how, first, float(np.max(np.abs(tn - tn_ref) - tol)))
            want_acc = acc0.astype(np.complex128) + c * tn
            assert np.max(np.abs(acc - want_acc)) <= 1e-6 * np.max(np.abs(want_acc)), (how, first)
        # constant-group entries from the host table: the same bits as evaluating them per row
        assert np.array_equal(outs["pauli"][0], outs["pauli_no_const"][0]), first
        d = np.abs(outs["pauli"][0] - outs["csr"][0])
        assert np.all(d <= 2.0 ** -22 * np.abs(outs["csr"][0]) + 1e-15 * bound + 1e-30), (first, float(d.max()))


def test_chebyshev_matrix_free_equals_csr_evolution(bk, dev):
    """exp(-iHt)|psi> on a 12-qubit chain with the matrix-free terms (the
    default for one state) and with the CSR terms: the same result to 1e-6,
    and both against FP64 expm_multiply; the batch path and fp32c row sums
    stay on the CSR, and matrix_free=True is refused where it cannot run.
    The caller's state is not modified (for one state the RHS-contiguous
    transpose is a view of it, so the evolution works on a copy)."""
    from oracle import hamiltonian as OH
    from blackwell_backend import workloads as W
    from blackwell_backend.hamiltonian import HamiltonianCSR
    from blackwell_backend.chebyshev import chebyshev_evolve, chebyshev_evolve_batch
    spo = W.heisenberg_chain(12, seed=4)
    h = HamiltonianCSR.from_sparse_pauli_op(spo, device=dev)
    rng = np.random.default_rng(3)
    psi = rng.standard_normal(1 << 12) + 1j * rng.standard_normal(1 << 12)
    psi = (psi / np.linalg.norm(psi)).astype(np.complex64)
    x = to_dev(psi, dev)
    a, ia = chebyshev_evolve(x, h, 2.0, 1e-9)
    assert np.array_equal(x.cpu().numpy(), psi)          # the input state is left as it was
    b, ib = chebyshev_evolve(x, h, 2.0, 1e-9, matrix_free=False)
    assert ia["matrix_free"] and not ib["matrix_free"]
    a, b = a.cpu().numpy(), b.cpu().numpy()
    assert np.max(np.abs(a - b)) < 1e-6
    ref = OH.expm_reference(OH.pauli_sparse(spo.labels, spo.coeffs), psi.astype(np.complex128), 2.0)
    assert np.max(np.abs(a - ref)) < 1e-5
    _, ibatch = chebyshev_evolve_batch(torch.stack([x, x]), h, 0.5, 1e-8)
    assert not ibatch["matrix_free"]
    _, ic = chebyshev_evolve(x, h, 0.5, 1e-8, accumulate="fp32c")
    assert not ic["matrix_free"]
    with pytest.raises(ValueError):
        chebyshev_evolve_batch(torch.stack([x, x]), h, 0.5, 1e-8, matrix_free=True)


def test_csr_split_and_split_term(bk, dev):
    """qsim_csr_split_count / _fill split a ragged CSR of 2^10 rows at bit 4:
    the low part holds exactly the entries with (row ^ col) < 16 of each
    row, the high part (row p = rot(p)) the others, both in the original
    entry order.  The two-pass term on that split (32 RHS) equals the
    one-pass FP64 term up to the complex64 rounding of the low partial sum,
    for the first term and a later one."""
    import ctypes
    import scipy.sparse as sp
    from blackwell_kernels._lib import call, ptr, stream_handle
    n, sb, N = 10, 4, 1 << 10
    rng = np.random.default_rng(21)
    A = sp.random(N, N, density=0.01, random_state=3, format="lil").astype(np.complex128)
    for r in range(0, N, 37):             # add entries near the diagonal (low masks)
        for d in (1, 3, 6, 15):
            A[r, r ^ d] = rng.standard_normal() + 1j
    A = A.tocsr()
    A.data = A.data + 1j * rng.standard_normal(A.data.shape)
    A.sort_indices()
    rp = torch.from_numpy(A.indptr.astype(np.int32)).to(dev)
    ci = torch.from_numpy(A.indices.astype(np.int32)).to(dev)
    va = torch.from_numpy(A.data.astype(np.complex64)).to(dev)
    lib = bk._lib.load()
    ws = torch.empty(int(lib.qsim_csr_split_workspace_bytes(N)), dtype=torch.uint8, device=dev)
    nlo, nhi = ctypes.c_int64(0), ctypes.c_int64(0)
    st = stream_handle(dev)
    call("qsim_csr_split_count", st, ptr(rp), ptr(ci), N, sb, ptr(ws), ctypes.byref(nlo), ctypes.byref(nhi))
    assert nlo.value + nhi.value == A.nnz
    rp_lo = torch.empty(N + 1, dtype=torch.int32, device=dev)
    rp_hi = torch.empty(N + 1, dtype=torch.int32, device=dev)
    ci_lo = torch.empty(nlo.value, dtype=torch.int32, device=dev)
    ci_hi = torch.empty(nhi.value, dtype=torch.int32, device=dev)
    v_lo = torch.empty(nlo.value, dtype=torch.complex64, device=dev)
    v_hi = torch.empty(nhi.value, dtype=torch.complex64, device=dev)
    call("qsim_csr_split_fill", st, ptr(rp), ptr(ci), ptr(va), N, sb, ptr(ws), ptr(rp_lo), ptr(ci_lo),
         ptr(v_lo), ptr(rp_hi), ptr(ci_hi), ptr(v_hi))
    h_lo = (rp_lo.cpu().numpy(), ci_lo.cpu().numpy(), v_lo.cpu().numpy())
    h_hi = (rp_hi.cpu().numpy(), ci_hi.cpu().numpy(), v_hi.cpu().numpy())
    a32 = A.data.astype(np.complex64)

    def rot(p):
        return ((p & ((1 << (n - sb)) - 1)) << sb) | (p >> (n - sb))
    for r in range(N):
        cols = A.indices[A.indptr[r]:A.indptr[r + 1]]
        vals = a32[A.indptr[r]:A.indptr[r + 1]]
        low = ((cols ^ r) >> sb) == 0
        got_c = h_lo[1][h_lo[0][r]:h_lo[0][r + 1]]
        got_v = h_lo[2][h_lo[0][r]:h_lo[0][r + 1]]
        assert np.array_equal(got_c, cols[low]) and np.array_equal(got_v.view(np.uint32), vals[low].view(np.uint32)), r
    for p in range(N):
        r = rot(p)
        cols = A.indices[A.indptr[r]:A.indptr[r + 1]]
        vals = a32[A.indptr[r]:A.indptr[r + 1]]
        high = ((cols ^ r) >> sb) != 0
        got_c = h_hi[1][h_hi[0][p]:h_hi[0][p + 1]]
        got_v = h_hi[2][h_hi[0][p]:h_hi[0][p + 1]]
        assert np.array_equal(got_c, cols[high]) and np.array_equal(got_v.view(np.uint32), vals[high].view(np.uint32)), p
    R, scale, c = 32, 0.21, 0.4 + 0.25j
    A64 = A.astype(np.complex64).astype(np.complex128)
    bound = scale * (abs(A64) @ np.ones(N))
    for first in (1, 0):
        T = (rng.standard_normal((N, R)) + 1j * rng.standard_normal((N, R))).astype(np.complex64)
        Tp = (rng.standard_normal((N, R)) + 1j * rng.standard_normal((N, R))).astype(np.complex64)
        outs = []
        for split in (0, 1):
            tpd, acc = to_dev(Tp, dev), torch.zeros(N, R, dtype=torch.complex64, device=dev)
            comp = torch.zeros_like(acc)
            if split:
                w1 = torch.empty_like(acc)
                call("qsim_cheb_spmm_step_split", st, ptr(rp_lo), ptr(ci_lo), ptr(v_lo), ptr(rp_hi),
                     ptr(ci_hi), ptr(v_hi), N, sb, R, scale, ptr(to_dev(T, dev)), ptr(tpd), ptr(acc),
                     ptr(comp), ptr(w1), c.real, c.imag, first)
            else:
                call("qsim_cheb_spmm_step", st, ptr(rp), ptr(ci), ptr(va), N, R, scale, ptr(to_dev(T, dev)),
                     ptr(tpd), ptr(acc), ptr(comp), c.real, c.imag, first, 0)
            outs.append((tpd.cpu().numpy().astype(np.complex128), (acc + comp).cpu().numpy()))
        tmax = np.max(np.abs(T))
        tol = 2.0 ** -22 * (2 * bound[:, None] * tmax + np.abs(outs[0][0])) + 1e-30
        assert np.all(np.abs(outs[0][0] - outs[1][0]) <= tol), first
        assert np.max(np.abs(outs[0][1] - outs[1][1])) <= 1e-6 * np.max(np.abs(outs[0][1])), first


def test_block_batch_equals_per_instance_blocks(bk, dev):
    """Batched block steps (qsim_apply_block_batch) give bitwise the results
    of applying each instance's blocks alone (qsim_apply_block_mfma, whose
    variant depends on the qubits), for instances of different sizes and
    block counts, including qubit 0 and three or more low block qubits."""
    rng = np.random.default_rng(77)
    ns = [10, 11, 12, 15, 13]
    counts = [3, 1, 4, 2, 5]
    states, mats, qubits, refs = [], [], [], []
    for i, (n, b) in enumerate(zip(ns, counts)):
        psi = rand_state(n, 500 + i)
        qs = []
        for k in range(b):
            if k % 3 == 0:
                q = [0, 1, 2, int(n - 1), int(n - 2)]          # three low qubits, with qubit 0
            else:
                q = [int(x) for x in rng.choice(n, 5, replace=False)]
            qs.append(q + [-1, -1, -1])
        us = np.stack([rand_unitary(32, 900 + 10 * i + k) for k in range(b)]).astype(np.complex64)
        ref = to_dev(psi, dev)
        for k in range(b):
            bk.apply_block(ref, torch.from_numpy(us[k]).to(dev), qs[k][:5])
        refs.append(ref)
        states.append(to_dev(psi, dev))
        mats.append(torch.from_numpy(us).to(dev))
        qubits.append(np.array(qs, dtype=np.int32))
    bk.apply_blocks_batched(states, mats, qubits)
    for i, (a, b) in enumerate(zip(states, refs)):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32)), (i, ns[i])
    # against FP64 too, for one instance
    psi = rand_state(ns[2], 502).astype(np.complex128)
    for k in range(counts[2]):
        psi = O.apply_matrix(psi, mats[2][k].cpu().numpy().astype(np.complex128), list(qubits[2][k][:5]))
    assert _rel_err(psi, states[2].cpu().numpy()) < 8e-6 * counts[2]


def test_sample_batch_equals_per_state_sampling(bk, dev):
    """qsim_sample_batch gives, for every state, bitwise the samples of
    uniforms_splitmix + sample_with_uniforms on that state alone, including
    zero-shot records and states with zero-probability leaves."""
    ns = [8, 12, 10, 20, 15, 9]
    shots = [1000, 0, 4096, 777, 2048, 1]
    seeds = [3, 5, 7, 11, 13, 2 ** 63 + 17]
    states = []
    for i, n in enumerate(ns):
        psi = rand_state(n, 800 + i)
        if i == 2:
            psi[: 1 << (n - 1)] = 0            # first half of the leaves all zero
            psi /= np.linalg.norm(psi)
        states.append(to_dev(psi.astype(np.complex64), dev))
    got = bk.sample_batch(states, seeds, shots)
    for st, sd, k, g in zip(states, seeds, shots, got):
        ref = bk.sample_with_uniforms(st, bk.uniforms_splitmix(sd, k, dev))
        assert g.numel() == k and torch.equal(g, ref)
