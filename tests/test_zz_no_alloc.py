"""nqh_exact_noalloc (exact sums with every buffer in the library's static
scratch memory) against nqh_exact, which allocates, through the C ABI: two
calls in a row, the second on a smaller input, then a ket RBM, and the size
limit refused before any launch."""
import ctypes as C

import numpy as np
import pytest

from _helpers import ising, params
from nqhip import Context, _lib, operators, problems

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


def noalloc(ctx, bits, want_logp=True):
    out = np.empty(4)
    lp = np.empty(1 << bits) if want_logp else None
    st = ctx.lib.nqh_exact_noalloc(ctx.h, out.ctypes.data_as(_dp), lp.ctypes.data_as(_dp) if want_logp else None)
    return st, out, lp


def check_against_exact(ctx, bits):
    st, out, lp = noalloc(ctx, bits)
    assert st == _lib.NQH_OK, ctx.lib.nqh_last_error(ctx.h)
    C_ex, L_ex, logZ, lp_ex = ctx.exact(want_logp=True)
    # the two entry points chunk differently (2^16 vs 2^20 configurations), so
    # the partial sums combine in a different order: rounding-level agreement
    assert abs(out[0] - C_ex) <= 1e-12 * abs(C_ex)
    assert abs(complex(out[1], out[2]) - L_ex) <= 1e-12 * max(abs(L_ex), abs(C_ex))
    assert abs(out[3] - logZ) <= 1e-12 * max(1.0, abs(logZ))
    assert np.max(np.abs(lp - lp_ex)) <= 1e-12 * max(1.0, np.max(np.abs(lp_ex)))


def test_noalloc_exact_matches_exact(gpu_ok):
    ctx = Context(0, 1)
    # 2 x 9 = 18 bits: four chunks of the scratch memory
    flat, _, _ = ising(9)
    ctx.set_problem(flat)
    ctx.set_ndm(9, 9, 9)
    ctx.set_params(params(ctx.np_, seed=3, sigma=0.1))
    check_against_exact(ctx, 18)
    # the second call, smaller (one chunk of 256 configurations)
    flat, _, _ = ising(4)
    ctx.set_problem(flat)
    ctx.set_ndm(4, 8, 4)
    ctx.set_params(params(ctx.np_, seed=4, sigma=0.3))
    check_against_exact(ctx, 8)
    ctx.close()


def test_noalloc_exact_ket_and_limit(gpu_ok):
    N = 6
    ctx = Context(0, 2)
    _, H = problems.ising_hamiltonian(N, g=1.0, V=1.0)
    ctx.set_problem(operators.flatten(H))
    ctx.set_rbm(N, N)
    ctx.set_params(params(ctx.np_, seed=5, sigma=0.1))
    check_against_exact(ctx, N)
    flat, _, _ = ising(16)
    ctx.set_problem(flat)
    ctx.set_ndm(16, 16, 16)
    st, _, _ = noalloc(ctx, 0, want_logp=False)
    assert st == _lib.NQH_E_ARG
    ctx.close()
