"""Bit-reproducibility.  The force is summed per workgroup in an exact
fixed-point accumulator (integer adds are associative), and every other
reduction runs in a fixed order, so the same inputs give bit-identical F,
S and parameters whatever the order in which waves arrive, whether the
forces launch runs in order or beside the S build on the side stream, and
from one context to the next."""
import numpy as np
import pytest

from _helpers import ising, params, random_keys, xyz
from nqhip import Context, _lib

pytestmark = pytest.mark.gpu


def xyz_ctx(prec, seed=3):
    N, nh, na = 16, 32, 32
    flat, _, _ = xyz(4, 4)
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(params(ctx.np_, seed=17, sigma=0.05))
    ctx.set_precision(*prec)
    return ctx


@pytest.mark.parametrize("prec", [(_lib.PREC_FP64, _lib.PREC_FP64), (_lib.PREC_FP32, _lib.PREC_FP32)])
def test_force_and_s_are_bit_reproducible_at_xyz44_shapes(gpu_ok, prec):
    """4096 x 4 samples of the 4x4 XYZ model (the forces launch runs beside
    the S build on the side stream there): three evaluations, the last with
    per-phase timing on (every launch in order), give identical bits."""
    keys = random_keys(16, 4096 * 4, seed=8)
    out = []
    for i in range(3):
        ctx = xyz_ctx(prec)
        if i == 2:
            ctx.set_profiling(True)
        st, S, F, Ob = ctx.grad_from_samples(keys, 4096, 4)
        out.append((st, S, F, Ob))
        ctx.close()
    for st, S, F, Ob in out[1:]:
        assert np.array_equal(F, out[0][2])
        assert np.array_equal(S, out[0][1])
        assert np.array_equal(st, out[0][0]) and np.array_equal(Ob, out[0][3])


@pytest.mark.parametrize("graph", [False, True])
def test_sr_trajectory_is_bit_reproducible(gpu_ok, graph):
    """Two contexts with the same seed, sampler and parameters take the same
    five SR steps bit for bit (the chosen form: FP32 network, BF16 S build,
    persistent chains), eagerly and as graph replays."""
    prec = (_lib.PREC_FP32, _lib.PREC_BF16)
    runs = []
    for _ in range(2):
        ctx = xyz_ctx(prec, seed=11)
        ctx.set_sampler(4096, 4, 33, 50, _lib.SAMPLER_PERSISTENT)
        ctx.set_iteration(0)
        ctx.set_graph(graph)
        stats = [ctx.sr_iteration(0.01, 0.005) for _ in range(5)]
        runs.append((np.array(stats), ctx.get_params()))
        ctx.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


def test_rbmsplit_force_is_bit_reproducible(gpu_ok):
    N, nh = 6, 12
    flat, _, _ = ising(N)
    keys = random_keys(N, 512 * 8, seed=9)
    Fs = []
    for _ in range(2):
        ctx = Context(0, 5)
        ctx.set_problem(flat)
        ctx.set_rbmsplit(N, nh)
        ctx.set_params(params(ctx.np_, seed=21, sigma=0.05))
        Fs.append(ctx.grad_from_samples(keys, 512, 8)[2])
        ctx.close()
    assert np.array_equal(Fs[0], Fs[1])
