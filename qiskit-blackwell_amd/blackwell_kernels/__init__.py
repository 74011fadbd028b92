"""gfx950 statevector kernels with the interface of the reference's
blackwell_kernels plugin (bwk/python/blackwell_kernels/__init__.py:12-33).

Every function runs on libqsim_gfx950.so (HIP, MI355X) through its C ABI;
nothing here falls back to another implementation.
"""
__version__ = "0.1.0"

from .cuquantum import (
    apply_gate, apply_gate_2q, apply_gates_fused, apply_gates_reg_tiled,
    apply_diagonal, apply_diagonal_phase, apply_diagonal_2q, state_init,
    measure_qubit, measure_probs, sample, sample_with_uniforms, sample_batch, uniforms_splitmix, norm2,
    expectation_z, expectation_pauli, expectation_paulis, expectation_pauli_string,
    apply_pauli_string,
    apply_mcgate, apply_mc2qgate, apply_mcx, apply_mcswap, apply_block, apply_blocks_batched, run_tile_program,
    swap_gate, cswap_gate,
    apply_depolarizing, apply_amplitude_damping, apply_dephasing, pauli_x_gate,
    apply_gate_batched, apply_gates_batched_fused, state_init_batched,
    CircuitGraph, hadamard_gate, random_unitary,
    cnot_gate, cz_gate, random_unitary_4x4,
    rz_gate_diag, phase_gate_diag, t_gate_diag, s_gate_diag, z_gate_diag, cz_gate_diag,
    renormalize,
    spmv_csr, chebyshev_step, chebyshev_accum, chebyshev_step_accum,
)
from .qv8 import qv8_simulate, qv8_simulate_ref, generate_qv8_circuits, generate_qv8_circuits_gpu
from .qv4 import qv4_simulate_cuda, qv4_simulate_numpy, generate_qv4_circuits, heavy_output_probability
from ._lib import QsimError, load as load_library, LIB_PATH
