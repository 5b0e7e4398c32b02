"""nqhip -- MI355X backend for NeuralQuantum.jl's steady-state VMC loop.

Host-side mirror of the reference's Julia API for this path (operators,
Liouvillian, NDM parameters, Metropolis/LocalRule, SR, Descent and the
BatchedSampler driver) on top of libnqhip.so, the C ABI of include/nqhip.h.
"""
from . import hilbert, operators, problems  # noqa: F401
from ._lib import Context, NQHError, load  # noqa: F401

__all__ = ["hilbert", "operators", "problems", "Context", "NQHError", "load"]
