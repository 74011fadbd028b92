"""Qiskit backend on MI355X (gfx950) kernels: BlackwellSimulator and the
Chebyshev Hamiltonian entry (blackwell_backend/__init__.py:10-13 of the
reference exports BlackwellSimulator; its hybrid Aer router is not part of
this package)."""
from .simulator import BlackwellSimulator
from .circuit import QuantumCircuit
from .hamiltonian import HamiltonianCSR, SparsePauliOp, expectation
from .chebyshev import chebyshev_coefficients, chebyshev_evolve, chebyshev_evolve_batch

__all__ = ["BlackwellSimulator", "QuantumCircuit", "HamiltonianCSR", "SparsePauliOp",
           "chebyshev_coefficients", "chebyshev_evolve", "chebyshev_evolve_batch"]
