"""Checkpoint and resume (nqh_save_state / nqh_load_state with the
parameters): a context restored from the state saved after three SR
iterations continues them bit for bit -- the persistent chains, the Philox
iteration index, and with the learned proposal its weights, Adam moments,
retraining counter and sample ring."""
import numpy as np
import pytest

from _helpers import ising, params
from nqhip import Context, NQHError, _lib

pytestmark = pytest.mark.gpu


def setup(learned, seed=31):
    N = 6
    flat, _, _ = ising(N)
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, N, N)
    ctx.set_params(params(ctx.np_, seed=2, sigma=0.01))
    ctx.set_precision(_lib.PREC_FP32, _lib.PREC_BF16)
    ctx.set_sampler(256, 8, 2 * N + 1, 20, _lib.SAMPLER_PERSISTENT)
    ctx.set_iteration(0)
    if learned:
        # mix 0.5, retrain every 2 iterations, 5 Adam steps of 64, buffer 1024
        ctx.made_setup(0.5, 2, 5, 64, 1e-3, 1024)
    ctx.set_graph(True)
    return ctx


@pytest.mark.parametrize("learned", [False, True])
def test_resume_continues_bit_for_bit(gpu_ok, learned):
    a = setup(learned)
    for _ in range(3):
        a.sr_iteration(0.01, 0.01)
    p3, blob = a.get_params(), a.save_state()
    tail_a = [a.sr_iteration(0.01, 0.01) for _ in range(3)]
    b = setup(learned, seed=31)
    b.set_params(p3)
    b.load_state(blob)
    tail_b = [b.sr_iteration(0.01, 0.01) for _ in range(3)]
    assert np.array_equal(np.array(tail_a), np.array(tail_b))
    assert np.array_equal(a.get_params(), b.get_params())
    assert np.array_equal(a.get_samples(), b.get_samples())
    if learned:
        assert np.array_equal(a.made_get_params(), b.made_get_params())
    a.close()
    b.close()


def test_load_state_refuses_a_mismatched_context(gpu_ok):
    a = setup(False)
    a.sr_iteration(0.01, 0.01)
    blob = a.save_state()
    b = setup(True)                       # has a learned proposal: the blob has none
    with pytest.raises(NQHError) as e:
        b.load_state(blob)
    assert e.value.status == _lib.NQH_E_ARG
    c = setup(False)
    with pytest.raises(NQHError):
        c.load_state(b"\0" * len(blob))   # bad magic
    for x in (a, b, c):
        x.close()
