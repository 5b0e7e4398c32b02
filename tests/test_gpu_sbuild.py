"""The BF16 S build (split-K Gram on the matrix cores) against the same Gram
computed on the CPU from the same BF16-rounded centred O rows: only FP32
accumulation error remains, so the bound is tight, at shapes that use many
128 x 128 tiles (edges not multiples of 128) and many sample chunks."""
import numpy as np
import pytest

from _helpers import ising, params, random_keys, rel_err
from nqhip import Context, _lib

pytestmark = pytest.mark.gpu


def to_bf16(x):
    """Round float32 values to BF16 (nearest, ties to even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("N,alpha,chains,Lc", [(12, 2, 256, 8), (8, 1, 100, 7)])
def test_bf16_gram_matches_cpu_gram_of_rounded_rows(gpu_ok, ref, N, alpha, chains, Lc):
    nh = na = alpha * N
    flat, _, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=21, sigma=0.1)
    keys = random_keys(N, chains * Lc, seed=22)
    ctx = Context(0, 1)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    ctx.set_precision(_lib.PREC_FP32, _lib.PREC_BF16)
    _, O = ctx.logpsi_grad(keys)                  # FP32 network values, reference order
    _, S, _, _ = ctx.grad_from_samples(keys, chains, Lc)
    Ns = len(keys)
    Oc = O - O.mean(axis=0)
    Re = to_bf16(Oc.real.astype(np.float32))
    Im = to_bf16(Oc.imag.astype(np.float32))
    S_cpu = (Re.T @ Re + Im.T @ Im) / Ns
    assert rel_err(S, S_cpu) < 2e-5
    assert np.array_equal(S, S.T)                 # mirrored exactly
    ctx.close()
