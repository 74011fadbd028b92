"""Chebyshev evolution exp(-iHt)|psi> on the gfx950 library
(blackwell_backend/chebyshev.py:26-172).

Coefficients are FP64 on the host: c_k = (2 - delta_k0) (-i)^k J_k(alpha),
alpha = spectral_bound * |t|, J_k from scipy.special.jv, conjugated for
t < 0.  The degree is found by expanding until the coefficient tail is
negligible and then trimming while the cumulative tail stays below the
requested precision (the reference trims from a fixed estimate
K_est = alpha + 2 ln(1/eps) + 20, which can sit only a few terms above the
need at large alpha; here K_est only seeds the search).

Device work per term is ONE launch (qsim_cheb_spmm_step): SpMM of the
scaled Hamiltonian against the current term, the recurrence
T_{k+1} = 2 H~ T_k - T_{k-1} written over T_{k-1}, and the Kahan/Neumaier
accumulation acc += c_k T_{k+1} with its compensation vector.  Any number
of right-hand sides runs in the same launch, stored RHS-contiguous.

split=True runs each term as two launches over the CSR split at bit n/2
(HamiltonianCSR.split, qsim_cheb_spmm_step_split; 32 right-hand sides,
FP64 row sums): the entries that connect rows differing only in the low
half of the index bits, in natural row order, then the others in an order
where the high index bits vary fastest.  Each pass gathers T_k rows from a
smaller neighbourhood, at the price of a second copy of the CSR and one
more write and read of a partial-sum vector.  On a 22-qubit Heisenberg
chain at 32 RHS it saves a few per cent per term, so the one-pass term is
the default.

A Hamiltonian built from a Pauli sum keeps its grouped Pauli tables, and a
single state with FP64 row sums runs each term matrix-free by default
(qsim_cheb_pauli_step): every entry is recomputed from the tables as the
CSR builder computes it, so no index or value stream is read, and the CSR
stays for batches, fp32c row sums and the split term.

Row sums of the SpMM (accumulate=):
  "fp64"   every FP32 x FP32 product is exact in FP64 and is summed with
           FP64 FMAs (default: the more accurate form, and the cheaper one on
           CDNA4, where FP64 FMA issues at the non-packed FP32 rate);
  "fp32c"  TwoProduct + Neumaier compensated FP32, the reference's
           TwoProduct/Kahan row sum (bwk/csrc/spmv_sm120.cu:26-54).
"""
import math
import time

import numpy as np
import torch
from scipy.special import jv as bessel_j

from blackwell_kernels import _lib
from blackwell_kernels._lib import call, ptr, stream_handle


def chebyshev_coefficients(t, spectral_bound, precision=1e-10):
    """(coeffs complex128 [K+1], K, tail_bound)."""
    alpha = float(spectral_bound) * abs(float(t))
    k_est = max(int(alpha + 2.0 * np.log(1.0 / precision) + 20), 10)
    # grow until the last 8 coefficients are far below the precision
    kmax = k_est
    while True:
        ks = np.arange(kmax + 1)
        c = np.where(ks > 0, 2.0, 1.0) * ((-1j) ** (ks % 4)) * bessel_j(ks, alpha)
        if np.max(np.abs(c[-8:])) < precision * 1e-6 or kmax > 4 * alpha + 4000:
            break
        kmax = int(kmax * 1.5) + 8
    cumtail = np.cumsum(np.abs(c)[::-1])[::-1]
    K = kmax
    for k in range(kmax, 0, -1):
        if cumtail[k] < precision:
            K = k
        else:
            break
    K = max(K, 2)
    coeffs = c[:K + 1].astype(np.complex128)
    if t < 0:
        coeffs = np.conj(coeffs)
    tail = float(cumtail[K + 1]) if K + 1 < len(cumtail) else 0.0
    return coeffs, K, tail


ACCUMULATE = {"fp64": 0, "fp32c": 1}


def chebyshev_evolve_batch(states, hamiltonian_csr, t, precision=1e-10, accumulate="fp64", split=False,
                           matrix_free=None):
    """Evolve R states at once.  states: complex64 [R, 2^n] (or [2^n]) on the
    device.  Returns (evolved [R, 2^n] complex64, info).

    matrix_free: run each term from the Pauli tables the Hamiltonian was
    built from (qsim_cheb_pauli_step) instead of the CSR; None (default)
    does so for one state with accumulate="fp64" when the Hamiltonian
    carries its Pauli tables."""
    if accumulate not in ACCUMULATE:
        raise ValueError(f"accumulate must be one of {sorted(ACCUMULATE)}, got {accumulate!r}")
    mode = ACCUMULATE[accumulate]
    single = states.dim() == 1
    X = states.reshape(1, -1) if single else states
    R, N = X.shape
    h = hamiltonian_csr
    if N != h.n_rows:
        raise RuntimeError(f"state dimension {N} != Hamiltonian dimension {h.n_rows}")
    dev = X.device
    torch.cuda.synchronize(dev)
    t_start = time.perf_counter()
    coeffs, K, tail = chebyshev_coefficients(t, h.spectral_bound, precision)
    scale = 1.0 / h.spectral_bound
    # RHS-contiguous buffers [N, R]; A is always a copy (for one state the
    # transpose is already contiguous and would alias the caller's tensor,
    # which the recurrence overwrites)
    A = X.transpose(0, 1).contiguous()
    if A.data_ptr() == X.data_ptr():
        A = A.clone()
    B = torch.empty_like(A)
    acc = torch.empty_like(A)
    comp = torch.empty_like(A)
    s = stream_handle(dev)
    n_el = N * R
    c0 = complex(coeffs[0])
    call("qsim_cheb_init_acc", s, ptr(acc), ptr(comp), ptr(A), n_el, c0.real, c0.imag)
    use_split = bool(split)
    if use_split and not (mode == ACCUMULATE["fp64"] and R == 32 and 4 <= N <= (1 << 24)):
        raise ValueError("split=True needs accumulate='fp64', 32 states and 4..2^24 rows "
                         f"(got {accumulate!r}, {R} states, {N} rows)")
    pauli = getattr(h, "pauli", None)
    mf_ok = pauli is not None and R == 1 and mode == ACCUMULATE["fp64"] and not use_split
    use_mf = mf_ok if matrix_free is None else bool(matrix_free)
    if use_mf and not mf_ok:
        raise ValueError("matrix_free=True needs a Hamiltonian built from a Pauli sum, one state, "
                         f"accumulate='fp64' and split=False (got {R} states, {accumulate!r})")
    if use_mf:
        pargs = (h.n_qubits,) + pauli.step_args() + (scale,)

        def term(cur_, prv_, ck, first):
            call("qsim_cheb_pauli_step", s, *pargs, ptr(cur_), ptr(prv_), ptr(acc), ptr(comp),
                 ck.real, ck.imag, first)
    elif use_split:
        sb, rp_lo, ci_lo, v_lo, rp_hi, ci_hi, v_hi = h.split()
        w1 = torch.empty_like(A)
        sargs = (ptr(rp_lo), ptr(ci_lo), ptr(v_lo), ptr(rp_hi), ptr(ci_hi), ptr(v_hi), N, sb, R, scale)

        def term(cur_, prv_, ck, first):
            call("qsim_cheb_spmm_step_split", s, *sargs, ptr(cur_), ptr(prv_), ptr(acc), ptr(comp),
                 ptr(w1), ck.real, ck.imag, first)
    else:
        args = (ptr(h.row_ptr), ptr(h.col_idx), ptr(h.values), N, R, scale)

        def term(cur_, prv_, ck, first):
            call("qsim_cheb_spmm_step", s, *args, ptr(cur_), ptr(prv_), ptr(acc), ptr(comp),
                 ck.real, ck.imag, first, mode)
    term(A, B, complex(coeffs[1]), 1)
    cur, prv = B, A
    for k in range(2, K + 1):
        term(cur, prv, complex(coeffs[k]), 0)
        cur, prv = prv, cur
    call("qsim_cheb_finish", s, ptr(prv), ptr(acc), ptr(comp), n_el)
    out = prv.transpose(0, 1).contiguous()
    torch.cuda.synchronize(dev)
    elapsed = time.perf_counter() - t_start
    info = {
        "engine": "chebyshev-beta",
        "experimental": True,
        "chebyshev_degree": K,
        "spectral_bound": h.spectral_bound,
        "truncation_error_bound": tail,
        "spmv_count": K,
        "workspace_vectors": 5 if use_split else 4,
        "workspace_bytes": (5 if use_split else 4) * n_el * 8,
        "elapsed_s": elapsed,
        "n_rhs": R,
        "accumulate": accumulate,
        "split_term": bool(use_split),
        "matrix_free": bool(use_mf),
    }
    return (out[0] if single else out), info


def chebyshev_evolve(state, hamiltonian_csr, t, precision=1e-10, bk=None, accumulate="fp64",
                     matrix_free=None):
    """exp(-iHt)|state> for one state [2^n]; returns (result, info)
    (chebyshev.py:81-172).  `bk` is accepted for signature parity."""
    return chebyshev_evolve_batch(state, hamiltonian_csr, t, precision, accumulate, matrix_free=matrix_free)
