"""Parity of the batched Chebyshev evolution at full size against FP64.

The 22-qubit open chain (ZZ + XX + YY + X, seeded couplings) evolves the 32
normalised complex-Gaussian right-hand sides of the benchmark at t = 1 on the
GPU; the first 4 are compared with scipy's expm_multiply on the oracle's FP64
CSR of the same operator.  This takes a few minutes of host work (the FP64
CSR build and about 60 sparse products), so it is a script to run by hand
rather than a collected test:

    python tests/check_c5_fp64.py [--t 1.0] [--check 4]

It prints one JSON line with max |delta amp| and the fidelity per checked
vector; the unit test of the same path at 10 qubits is
tests/test_gpu_backend.py::test_hamiltonian_evolution_matches_expm.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "qiskit-blackwell_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402
import torch  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=22)
    ap.add_argument("--rhs", type=int, default=32)
    ap.add_argument("--check", type=int, default=4)
    ap.add_argument("--t", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--precision", type=float, default=1e-7)
    a = ap.parse_args()
    import scipy.sparse as sp
    from scipy.sparse.linalg import expm_multiply
    from blackwell_backend import workloads as W
    from blackwell_backend.hamiltonian import HamiltonianCSR
    from blackwell_backend.chebyshev import chebyshev_evolve_batch
    from oracle import hamiltonian as OH
    dev = torch.device("cuda", 0)
    op = W.heisenberg_chain(a.n, a.seed)
    h = HamiltonianCSR.from_sparse_pauli_op(op, device=dev)
    N = 1 << a.n
    rng = np.random.default_rng(a.seed + 1)          # the benchmark's right-hand sides
    X = (rng.standard_normal((a.rhs, N)) + 1j * rng.standard_normal((a.rhs, N))).astype(np.complex64)
    X /= np.linalg.norm(X, axis=1, keepdims=True).astype(np.float32)
    Y, info = chebyshev_evolve_batch(torch.from_numpy(X).to(dev), h, a.t, a.precision)
    got = Y[:a.check].cpu().numpy().astype(np.complex128)
    print(f"[c5-fp64] GPU done: K = {info['chebyshev_degree']}; building the FP64 CSR", flush=True)
    t0 = time.perf_counter()
    rp, ci, v = OH.pauli_csr_grouped(op.labels, op.coeffs, a.n)
    A = sp.csr_matrix((v, ci, rp), shape=(N, N))
    print(f"[c5-fp64] CSR built in {time.perf_counter() - t0:.0f} s; expm_multiply", flush=True)
    t0 = time.perf_counter()
    ref = expm_multiply(-1j * a.t * A, X[:a.check].T.astype(np.complex128)).T
    print(f"[c5-fp64] expm_multiply done in {time.perf_counter() - t0:.0f} s", flush=True)
    out = []
    for k in range(a.check):
        f = abs(np.vdot(ref[k], got[k])) ** 2 / (np.vdot(ref[k], ref[k]).real * np.vdot(got[k], got[k]).real)
        out.append({"rhs": k, "max_abs_diff": float(np.max(np.abs(ref[k] - got[k]))),
                    "max_abs_ref": float(np.max(np.abs(ref[k]))), "fidelity": float(f)})
    print(json.dumps({"workload": f"{a.n}-qubit chain, t = {a.t:g}, {a.rhs} RHS, first {a.check} "
                                  f"checked, Bessel tail < {a.precision:g}",
                      "chebyshev_degree": info["chebyshev_degree"], "checks": out}), flush=True)


if __name__ == "__main__":
    main()
