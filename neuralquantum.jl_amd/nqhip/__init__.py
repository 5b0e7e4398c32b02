"""nqhip -- MI355X backend for NeuralQuantum.jl's steady-state VMC loop.

Host-side mirror of the reference's Julia API for this path (operators,
Liouvillian, NDM parameters, Metropolis/LocalRule, SR, Descent and the
BatchedSampler driver) on top of libnqhip.so, the C ABI of include/nqhip.h.
"""
from . import hilbert, operators, problems  # noqa: F401
from ._lib import Batch, Context, NQHError, load  # noqa: F401
from .vmc import (NDM, RBM, SR, RBMSplit, BatchedSampler, Descent, Gradient, HIPBatchedSampler, LocalRule,  # noqa: F401
                  Measurement, MetropolisSampler, add_observable_, compute_observables, init_random_pars_,
                  precondition_, sample_, update_)

__all__ = ["hilbert", "operators", "problems", "Batch", "Context", "NQHError", "load", "NDM", "RBM", "RBMSplit", "SR",
           "BatchedSampler", "Descent", "Gradient", "HIPBatchedSampler", "LocalRule", "Measurement",
           "MetropolisSampler", "add_observable_", "compute_observables", "init_random_pars_", "precondition_",
           "sample_", "update_"]
