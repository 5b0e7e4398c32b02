"""Failure semantics of the SR step, forced from the host.

The reference throws PosDefException from cholesky!(...; check=true) before
any parameter changes (SRDirect.jl:79); here every failure is a status code
and the parameters stay bit-identical: the update kernel reads the
factorization's pivot flag, the finiteness flag of the step and the loss on
the device and applies no step when one is bad, and nqh_update refuses after
a failed nqh_precondition.  A larger shift then recovers (the fallback the
host may take).  A chain that never moves is visible from the per-chain
acceptance counters."""
import numpy as np
import pytest

from _helpers import ising, params, random_keys
from nqhip import Context, NQHError, _lib

pytestmark = pytest.mark.gpu


def ndm_ctx(N, seed, sigma=0.1, flat=None):
    flat = ising(N)[0] if flat is None else flat
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, N, N)
    p = params(ctx.np_, seed=seed, sigma=sigma)
    ctx.set_params(p)
    return ctx, p


def test_zero_s_is_not_positive_definite_and_update_refused(gpu_ok):
    """Four copies of one configuration with all parameters zero: every O row
    is the same and holds only multiples of 1/2 (f'(0) = 1/2), so the column
    means are exact, the centred rows exactly zero, S = 0, and with eps = 0
    the first pivot is 0: NQH_E_NOTPD, no step; nqh_update refuses;
    eps = 0.01 then solves."""
    N = 4
    ctx, p = ndm_ctx(N, 1, sigma=0.0)
    keys = np.full(4, 0b10110101, dtype=np.uint64)
    _, S, F, _ = ctx.grad_from_samples(keys, 2, 2)
    assert not np.any(S)
    with pytest.raises(NQHError) as e:
        ctx.precondition(0.0)
    assert e.value.status == _lib.NQH_E_NOTPD
    with pytest.raises(NQHError) as e:
        ctx.update(0.01)
    assert e.value.status == _lib.NQH_E_STATE
    assert np.array_equal(ctx.get_params(), p)
    ctx.grad_from_samples(keys, 2, 2)
    d = ctx.precondition(0.01)             # the larger shift: S + 0.01 I = 0.01 I
    assert np.allclose(d, F / 0.01, rtol=1e-12, atol=0)
    ctx.update(0.5)
    assert np.allclose(ctx.get_params(), p - 0.5 * d, rtol=0, atol=1e-15)
    ctx.close()


@pytest.mark.parametrize("graph", [False, True])
def test_singular_s_in_the_loop_keeps_the_parameters(gpu_ok, graph):
    """eps = 0 with Ns = 8 < Np = 84: S has rank <= 7, the factorization meets
    a non-positive (or non-finite) pivot, the iteration reports NOTPD (or a
    non-finite step) and the parameters are unchanged, in the eager path and
    in the captured graph; the same context then steps with eps = 0.01."""
    N = 4
    ctx, p = ndm_ctx(N, 2)
    ctx.set_sampler(4, 2, 9, 5)
    ctx.set_iteration(0)
    ctx.set_graph(graph)
    with pytest.raises(NQHError) as e:
        ctx.sr_iteration(0.0, 0.01)
    assert e.value.status in (_lib.NQH_E_NOTPD, _lib.NQH_E_NAN)
    assert np.array_equal(ctx.get_params(), p)
    ctx.sr_iteration(0.01, 0.01)
    assert not np.array_equal(ctx.get_params(), p)
    ctx.close()


def nan_table(N):
    """The dissipative Ising table with one matrix element made NaN: L_loc is
    NaN wherever that connection is reached."""
    flat = ising(N)[0]
    mel = np.array(flat.ent_mel, dtype=np.complex128, copy=True)
    mel[1] = complex(np.nan, 0.0)
    flat.ent_mel = mel
    return flat


@pytest.mark.parametrize("prec", [(_lib.PREC_FP64, _lib.PREC_FP64), (_lib.PREC_FP32, _lib.PREC_BF16)])
@pytest.mark.parametrize("graph", [False, True])
def test_nan_in_local_values_is_caught_before_the_solve(gpu_ok, graph, prec):
    """Also with the FP32 network, whose forces go through the matrix-core
    launch: the non-finite terms reach the exact accumulator's poison word."""
    N = 4
    ctx, p = ndm_ctx(N, 3, flat=nan_table(N))
    ctx.set_precision(*prec)
    keys = random_keys(N, 64, seed=5)
    st, _, _, _ = ctx.grad_from_samples(keys, 8, 8)
    assert not np.isfinite(st[0])
    with pytest.raises(NQHError) as e:
        ctx.precondition(0.01)
    assert e.value.status == _lib.NQH_E_NAN and "no solve" in str(e.value)
    with pytest.raises(NQHError):
        ctx.update(0.01)
    ctx.set_sampler(8, 4, 9, 5)
    ctx.set_iteration(0)
    ctx.set_graph(graph)
    for _ in range(2):
        with pytest.raises(NQHError) as e:
            ctx.sr_iteration(0.01, 0.01)
        assert e.value.status == _lib.NQH_E_NAN
    assert np.array_equal(ctx.get_params(), p)
    ctx.close()


def test_nan_parameters_are_refused(gpu_ok):
    N = 4
    ctx, p = ndm_ctx(N, 4)
    q = p.copy()
    q[3] = np.nan
    with pytest.raises(NQHError) as e:
        ctx.set_params(q)
    assert e.value.status == _lib.NQH_E_NAN
    assert np.array_equal(ctx.get_params(), p)
    ctx.close()


def test_stuck_chains_show_in_the_acceptance_counters(gpu_ok):
    """A bias of 200 on every b_lambda makes the all-up configuration carry
    all the weight (a flip away costs log p by 400): chains kept across
    iterations reach it in the first iteration and then never move again,
    so every chain's counter reads 0; with a plain network they move."""
    N = 4
    ctx, p = ndm_ctx(N, 5)
    ctx.set_sampler(64, 4, 9, 20, _lib.SAMPLER_PERSISTENT)
    ctx.set_iteration(0)
    ctx.sample()
    assert np.all(ctx.chain_acceptance() > 0)
    q = p.copy()
    off = 2 * N + 2 * N * N              # b_la in the flat order b_mu[N] h_mu[N] w_mu[N^2] u_mu[N^2] b_la ...
    q[off:off + N] = 200.0
    ctx.set_params(q)
    for _ in range(3):                    # reach all-up (flips toward it are always accepted):
        ctx.sample()                      # 108 proposals miss a given site with probability 5e-7
    ctx.sample()
    acc = ctx.chain_acceptance()
    assert acc.shape == (64,) and np.all(acc == 0)
    keys = ctx.get_samples()
    assert np.all(keys == np.uint64((1 << (2 * N)) - 1))
    ctx.close()
