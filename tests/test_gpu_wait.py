"""Waiting for an iteration: with one iteration in flight the host polls the
status record's done word, which the iteration's last kernel stores after
its other fields; with several in flight (nqh_sr_iteration_async called
again before nqh_wait) it synchronises with the stream.  Both give the
statistics and parameters of iterations run one by one."""
import numpy as np
import pytest

from _helpers import ising, params
from nqhip import Context, _lib

pytestmark = pytest.mark.gpu


def make(ref, prec):
    N = 8
    flat, _, _ = ising(N, periodic=True, g=1.1)
    ctx = Context(0, 17)
    ctx.set_problem(flat)
    ctx.set_ndm(N, N, N)
    ctx.set_params(params(ref.ndm_np(N, N, N), seed=5, sigma=0.05))
    ctx.set_precision(*prec)
    ctx.set_sampler(256, 8, 2 * N + 1, 10, _lib.SAMPLER_PERSISTENT)
    ctx.set_iteration(0)
    return ctx


@pytest.mark.parametrize("prec", [(_lib.PREC_FP64, _lib.PREC_FP64), (_lib.PREC_FP32, _lib.PREC_BF16)])
def test_wait_paths_match_synchronous_iterations(gpu_ok, ref, prec):
    a, b = make(ref, prec), make(ref, prec)
    sa = [a.sr_iteration(0.01, 0.01) for _ in range(5)]
    sb = []
    b.sr_iteration_async(0.01, 0.01)
    sb.append(b.wait())                      # one in flight: the done word
    b.sr_iteration_async(0.01, 0.01)
    b.sr_iteration_async(0.01, 0.01)
    last = b.wait()                          # two in flight: the stream synchronisation
    sb += [None, last]
    sb.append(b.sr_iteration(0.01, 0.01))    # one in flight again
    sb.append(b.sr_iteration(0.01, 0.01))
    for k in (0, 2, 3, 4):
        assert np.array_equal(sb[k], sa[k]), k
    assert np.array_equal(b.get_params(), a.get_params())
    a.close()
    b.close()
