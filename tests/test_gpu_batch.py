"""Batches of independent problems (nqh_batch_*): every member's SR iteration
captured into one graph and launched once per iteration gives the same
statistics, samples and parameters as the same contexts run one by one."""
import numpy as np
import pytest

from _helpers import ising, params, rel_err
from nqhip import Batch, Context, NQHError, _lib

pytestmark = pytest.mark.gpu

SIZES = [(6, 1.0), (9, 0.6), (7, 1.7), (11, 1.2), (8, 0.9)]


def member(ref, N, g, seed, flags, prec):
    flat, _, _ = ising(N, periodic=True, g=g)
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, N, N)
    ctx.set_params(params(ref.ndm_np(N, N, N), seed=seed, sigma=0.05))
    ctx.set_precision(*prec)
    ctx.set_sampler(64, 8, 2 * N + 1, 10, flags)
    ctx.set_iteration(0)
    return ctx


@pytest.mark.parametrize("flags", [0, _lib.SAMPLER_PERSISTENT])
@pytest.mark.parametrize("prec", [(_lib.PREC_FP64, _lib.PREC_FP64), (_lib.PREC_FP32, _lib.PREC_BF16)])
def test_batch_matches_members_run_alone(gpu_ok, ref, flags, prec):
    alone = [member(ref, N, g, 100 + i, flags, prec) for i, (N, g) in enumerate(SIZES)]
    inb = [member(ref, N, g, 100 + i, flags, prec) for i, (N, g) in enumerate(SIZES)]
    batch = Batch(inb)
    for it in range(4):
        st = batch.sr_iteration(0.01, 0.01)
        for i, c in enumerate(alone):
            sa = c.sr_iteration(0.01, 0.01)
            assert rel_err(st[i], sa) < 1e-12, (it, i)
            assert np.array_equal(inb[i].get_samples(), c.get_samples()), (it, i)
    for a, b in zip(alone, inb):
        assert rel_err(b.get_params(), a.get_params()) < 1e-12
    batch.close()
    for c in alone + inb:
        c.close()


def test_batch_member_keeps_working_alone(gpu_ok, ref):
    """A member's own iterations interleave with the batch's in order."""
    prec = (_lib.PREC_FP64, _lib.PREC_FP64)
    a = [member(ref, 6, 1.0, 7, _lib.SAMPLER_PERSISTENT, prec), member(ref, 8, 1.4, 8, _lib.SAMPLER_PERSISTENT, prec)]
    b = [member(ref, 6, 1.0, 7, _lib.SAMPLER_PERSISTENT, prec), member(ref, 8, 1.4, 8, _lib.SAMPLER_PERSISTENT, prec)]
    batch = Batch(b)
    for it in range(3):
        batch.sr_iteration(0.01, 0.01)
        for c in a:
            c.sr_iteration(0.01, 0.01)
        sb = b[1].sr_iteration(0.01, 0.01)       # member 1 alone, between batch iterations
        sa = a[1].sr_iteration(0.01, 0.01)
        assert rel_err(sb, sa) < 1e-12, it
        a[0].sr_iteration(0.01, 0.01)
        b[0].sr_iteration(0.01, 0.01)
    for x, y in zip(a, b):
        assert rel_err(y.get_params(), x.get_params()) < 1e-12
    batch.close()
    for c in a + b:
        c.close()


def test_batch_rejects_learned_proposals(gpu_ok, ref):
    prec = (_lib.PREC_FP32, _lib.PREC_BF16)
    c0 = member(ref, 6, 1.0, 1, 0, prec)
    c1 = member(ref, 6, 1.0, 2, 0, prec)
    c1.made_setup(0.5, 0, 0, 64, 1e-3, 0)
    batch = Batch([c0, c1])
    with pytest.raises(NQHError, match="member 1"):
        batch.sr_iteration(0.01, 0.01)
    batch.close()
    c0.close()
    c1.close()


def test_batch_create_rejects_duplicates(gpu_ok):
    c = Context(0, 1)
    with pytest.raises(NQHError):
        Batch([c, c])
    c.close()


def test_batch_recaptures_after_a_member_is_reconfigured(gpu_ok, ref):
    """A member given a new sampler (other chain count: its buffers are
    regrown), new parameters and a new iteration number between batch
    iterations: the next batch iteration captures the graph again and the
    member matches the same context run alone."""
    prec = (_lib.PREC_FP64, _lib.PREC_FP64)
    alone = [member(ref, N, g, 300 + i, 0, prec) for i, (N, g) in enumerate(SIZES[:3])]
    inb = [member(ref, N, g, 300 + i, 0, prec) for i, (N, g) in enumerate(SIZES[:3])]
    batch = Batch(inb)

    def step(it):
        st = batch.sr_iteration(0.01, 0.01)
        for i, c in enumerate(alone):
            sa = c.sr_iteration(0.01, 0.01)
            assert rel_err(st[i], sa) < 1e-12, (it, i)
            assert np.array_equal(inb[i].get_samples(), c.get_samples()), (it, i)

    step(0)
    step(1)
    for c in (alone[1], inb[1]):
        N = SIZES[1][0]
        c.set_sampler(96, 6, 2 * N + 1, 10, 0)
        c.set_params(params(ref.ndm_np(N, N, N), seed=77, sigma=0.05))
        c.set_iteration(40)
    step(2)
    step(3)
    for a, b in zip(alone, inb):
        assert rel_err(b.get_params(), a.get_params()) < 1e-12
    batch.close()
    for c in alone + inb:
        c.close()


UNEVEN = [(16, 1.3), (6, 0.7), (12, 1.9), (9, 0.5), (14, 1.1), (7, 1.6), (11, 0.8), (8, 1.2)]


@pytest.mark.parametrize("flags", [0, _lib.SAMPLER_PERSISTENT])
@pytest.mark.parametrize("prec", [(_lib.PREC_FP64, _lib.PREC_FP64), (_lib.PREC_FP32, _lib.PREC_BF16),
                                  (_lib.PREC_FP32, _lib.PREC_FP32)])
def test_packed_batch_matches_the_per_member_graph(gpu_ok, ref, flags, prec):
    """Members of uneven size (N = 6 .. 16) in the packed form, every kernel
    launched once for all members, against the same members in the
    per-member graph: bit-identical statistics, samples and parameters over
    four iterations."""
    pk = [member(ref, N, g, 500 + i, flags, prec) for i, (N, g) in enumerate(UNEVEN)]
    gr = [member(ref, N, g, 500 + i, flags, prec) for i, (N, g) in enumerate(UNEVEN)]
    bp, bg = Batch(pk), Batch(gr)
    bg.set_mode(_lib.BATCH_GRAPH)
    try:
        for it in range(4):
            sp = bp.sr_iteration(0.01, 0.01)
            sg = bg.sr_iteration(0.01, 0.01)
            assert np.array_equal(sp, sg), it
            for i in range(len(UNEVEN)):
                assert np.array_equal(pk[i].get_samples(), gr[i].get_samples()), (it, i)
        if flags & _lib.SAMPLER_PERSISTENT:
            assert bp.launches() > 0 and bg.launches() == 0
        for a, b in zip(pk, gr):
            assert np.array_equal(a.get_params(), b.get_params())
    finally:
        bp.close()
        bg.close()
        for c in pk + gr:
            c.close()


def test_packed_batch_falls_back_for_members_it_does_not_cover(gpu_ok, ref):
    """A member with more than 32 hidden units (two unit rows per lane) keeps
    the batch on the per-member graph (no packed form), and the batch still
    matches members run alone."""
    prec = (_lib.PREC_FP64, _lib.PREC_FP64)

    def wide(seed):
        ctx = member(ref, 6, 1.0, seed, 0, prec)
        ctx.set_ndm(6, 40, 6)
        ctx.set_params(params(ref.ndm_np(6, 40, 6), seed=seed, sigma=0.05))
        ctx.set_sampler(64, 8, 13, 10, 0)
        return ctx

    alone = [member(ref, 7, 1.7, 700, 0, prec), wide(701)]
    inb = [member(ref, 7, 1.7, 700, 0, prec), wide(701)]
    batch = Batch(inb)
    try:
        for it in range(2):
            st = batch.sr_iteration(0.01, 0.01)
            for i, c in enumerate(alone):
                assert np.array_equal(st[i], c.sr_iteration(0.01, 0.01)), (it, i)
        assert batch.launches() == 0
    finally:
        batch.close()
        for c in alone + inb:
            c.close()
