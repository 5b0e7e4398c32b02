"""The FP32 network's forces against the FP64 oracle at each tile class of the
matrix-core forces launch (k_force_mma): class 0 (2 nh <= 32, na <= 16, the
batched chains' shapes), class 1 (the 4x4 XYZ benchmark shapes, nh = na = 32)
and class 2 (nh = na = 64), with both activations.  F of 128 chains x 4
random samples against the oracle's sample! post-processing of its own
log-derivatives and local estimators (BatchedGradSampler.jl:99-118,
AccumulatorObsGrad.jl:39-127); bar: the FP32 form's 1e-4 of max |F|."""
import pytest

from _helpers import ising, params, random_keys, rel_err, xyz
from nqhip import Context, _lib
from oracle import nqref

pytestmark = pytest.mark.gpu

CASES = [("ising6", 6, 6, 6), ("xyz44", 16, 32, 32), ("ising8_wide", 8, 64, 64)]


@pytest.mark.parametrize("act", [_lib.ACT_SOFTPLUS, _lib.ACT_LOGCOSH])
@pytest.mark.parametrize("name,N,nh,na", CASES)
def test_fp32_forces_against_the_oracle(gpu_ok, name, N, nh, na, act):
    flat, tab, _ = xyz(4, 4) if name == "xyz44" else ising(N)
    ref = nqref.Ref(act=act)
    p = params(ref.ndm_np(N, nh, na), seed=60 + N, sigma=0.05)
    chains, Lc = 128, 4
    keys = random_keys(N, chains * Lc, seed=61 + N)
    ctx = Context(0, 3)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na, act)
    ctx.set_params(p)
    ctx.set_precision(_lib.PREC_FP32, _lib.PREC_FP64)
    _, _, F, _ = ctx.grad_from_samples(keys, chains, Lc)
    _, O = ref.logpsi_grad(N, nh, na, p, keys)
    L, dL = ref.local_estimator(tab, nh, na, p, keys)
    _, F_r, _, _ = ref.grad_step(chains, Lc, O, L, dL)
    assert rel_err(F, F_r) < 1e-4
    ctx.close()
