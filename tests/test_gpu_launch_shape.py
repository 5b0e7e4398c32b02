"""Launch shapes (nqh_set_launch_shape): how the work is spread over the
device, not what is computed.  Sampler waves per workgroup leave the chains'
trajectories bit-identical; the estimator's workgroup size and waves regroup
the exact fixed-point force sums (FP64 rounding only); the Gram split-K
regroups FP32 partial sums (BF16 S build)."""
import numpy as np
import pytest

from _helpers import ising, params, rel_err
from nqhip import Batch, Context, NQHError, _lib

pytestmark = pytest.mark.gpu

SHAPES = [(8, 4, 2, 1), (64, 8, 1, 2), (32, 0, 16, 4)]


def ctx_for(ref, N, prec, seed=11):
    flat, _, _ = ising(N, periodic=True, g=1.3)
    c = Context(0, seed)
    c.set_problem(flat)
    c.set_ndm(N, N, N)
    c.set_params(params(ref.ndm_np(N, N, N), seed=seed, sigma=0.05))
    c.set_precision(*prec)
    c.set_sampler(256, 8, 2 * N + 1, 10, 0)
    return c


@pytest.mark.parametrize("shape", SHAPES)
def test_shapes_keep_samples_and_change_sums_by_rounding(gpu_ok, ref, shape):
    N = 9
    base = ctx_for(ref, N, (_lib.PREC_FP64, _lib.PREC_FP64))
    tuned = ctx_for(ref, N, (_lib.PREC_FP64, _lib.PREC_FP64))
    tuned.set_launch_shape(*shape)
    assert tuned.get_launch_shape() == list(shape)
    k0, a0 = base.sample()
    k1, a1 = tuned.sample()
    assert np.array_equal(k0, k1) and a0 == a1
    st0, S0, F0, _ = base.grad_from_samples(k0, 256, 8)
    st1, S1, F1, _ = tuned.grad_from_samples(k0, 256, 8)
    assert rel_err(st1, st0) < 1e-13 and rel_err(F1, F0) < 1e-12 and rel_err(S1, S0) < 1e-13
    # BF16 S build: the split-K regroups FP32 partial sums
    b16 = [ctx_for(ref, N, (_lib.PREC_FP32, _lib.PREC_BF16)) for _ in range(2)]
    b16[1].set_launch_shape(*shape)
    _, Sa, Fa, _ = b16[0].grad_from_samples(k0, 256, 8)
    _, Sb, Fb, _ = b16[1].grad_from_samples(k0, 256, 8)
    assert rel_err(Fb, Fa) < 1e-6 and np.max(np.abs(Sb - Sa)) < 1e-5 * np.max(np.abs(Sa))
    for c in [base, tuned] + b16:
        c.close()


def test_shape_arguments_are_checked(gpu_ok):
    c = Context(0, 1)
    for bad in [(-1, 0, 0, 0), (0, 6, 0, 0), (0, 0, 65, 0), (0, 0, 0, 3)]:
        with pytest.raises(NQHError):
            c.set_launch_shape(*bad)
    c.close()


def test_packed_batch_with_tuned_shapes_matches_the_per_member_graph(gpu_ok, ref):
    """Members with a non-default shape: the packed form uses each member's
    own shape, so it still matches the per-member graph bit for bit."""
    sizes = [(12, 1.1), (7, 0.6), (10, 1.8)]
    prec = (_lib.PREC_FP32, _lib.PREC_BF16)
    mk = lambda: [ctx_for(ref, N, prec, seed=40 + i) for i, (N, _) in enumerate(sizes)]  # noqa: E731
    pk, gr = mk(), mk()
    for c in pk + gr:
        c.set_launch_shape(32, 4, 2, 2)
    bp, bg = Batch(pk), Batch(gr)
    bg.set_mode(_lib.BATCH_GRAPH)
    try:
        for it in range(3):
            assert np.array_equal(bp.sr_iteration(0.01, 0.01), bg.sr_iteration(0.01, 0.01)), it
        assert bp.launches() > 0
        for a, b in zip(pk, gr):
            assert np.array_equal(a.get_params(), b.get_params())
    finally:
        bp.close()
        bg.close()
        for c in pk + gr:
            c.close()
