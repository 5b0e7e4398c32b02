"""The host-side mirror of the reference's driver API, run as the reference's
example runs it (examples/dissipative_ising1d.jl: BatchedSampler, sample!,
compute_observables, precondition!, Optimisers.update!), against the same
steps made directly through the C ABI with the same seeds.  Both paths issue
the same kernels in the same order; the force sums its per-sample terms with
LDS atomics, whose order varies from run to run, so the step agrees to
cond(S) x 1e-16 and everything after it to 1e-10."""
import numpy as np
import pytest

from nqhip import (NDM, SR, BatchedSampler, Context, Descent, LocalRule, MetropolisSampler, _lib,
                   add_observable_, compute_observables, hilbert, operators, precondition_, sample_, update_)

pytestmark = pytest.mark.gpu


def example_problem(N, g=0.4, V=2.0):
    h = hilbert.HomogeneousSpin(N)
    H = operators.LocalOperator(h)
    Sx = operators.LocalOperator(h)
    Sz = operators.LocalOperator(h)
    jumps = []
    for i in range(1, N + 1):
        H = H + (g / 2.0) * operators.sigmax(h, i)
        H = H + (V / 4.0) * (operators.sigmaz(h, i) * operators.sigmaz(h, i % N + 1))
        Sx = Sx + (1.0 / N) * operators.sigmax(h, i)
        Sz = Sz + (1.0 / N) * operators.sigmaz(h, i)
        jumps.append(operators.sigmam(h, i))
    return operators.liouvillian(H, jumps), {"Sx": Sx, "Sz": Sz}


@pytest.mark.parametrize("precondition_type", ["shift", "multiplicative"])
def test_driver_matches_the_abi(gpu_ok, precondition_type):
    N, batch, L, passes, burn, seed = 5, 16, 25, 5, 20, 77
    liouv, obs = example_problem(N)
    net = NDM(N, 1, 1).init_random_pars(np.random.default_rng(3), sigma=0.01)
    p0 = net.flat.copy()
    sampl = MetropolisSampler(LocalRule(), L, passes, burn=burn, seed=seed)
    algo = SR(0.001, precondition_type=precondition_type, lambda0=1.0, b=0.9, lambda_min=1e-3)
    opt = Descent(0.01)
    is_ = BatchedSampler(net, sampl, liouv, algo, batch_sz=batch)
    for name, op in obs.items():
        add_observable_(is_, name, op)

    ctx = Context(0, seed)
    ctx.set_problem(operators.flatten(liouv))
    ctx.set_ndm(N, N, N)
    ctx.set_precision(_lib.PREC_FP64, _lib.PREC_FP64)
    ctx.set_params(p0)
    ctx.set_sampler(batch, L, passes, burn, 0)
    flat_obs = {name: operators.flatten(op) for name, op in obs.items()}
    try:
        for i in range(1, 4):
            ldata, prec = sample_(is_)
            ob = compute_observables(is_)
            grad = precondition_(prec, algo, i)
            update_(opt, is_, grad)

            st = ctx.sample_grad()
            ref_ob = {name: ctx.observable(op, batch, L, passes, burn) for name, op in flat_obs.items()}
            eps = 0.001 if precondition_type == "shift" else max(1.0 * 0.9 ** i, 1e-3)
            kind = _lib.SHIFT_ADD if precondition_type == "shift" else _lib.SHIFT_DIAG
            d = ctx.precondition(eps, kind)
            ctx.update(0.01)

            assert abs(ldata.mean - st[0]) <= 1e-10 * abs(st[0]) and abs(ldata.error - st[1]) <= 1e-10 * st[1]
            for name, r in ref_ob.items():
                assert abs(ob[name].mean - complex(r[0], r[5])) <= 1e-10 and abs(ob[name].error - r[1]) <= 1e-10
                assert abs(ob[name].mean.real) <= 1.0 + 1e-12
            np.testing.assert_allclose(grad, d, rtol=0, atol=1e-10 * np.abs(d).max())
            np.testing.assert_allclose(net.flat, ctx.get_params(), rtol=0, atol=1e-12)
        assert not np.array_equal(net.flat, p0)
    finally:
        is_.close()
        ctx.close()


def test_fused_step_follows_the_schedule(gpu_ok):
    """`step` (one fused SR iteration) applies the multiplicative schedule
    from iteration 0 on, as three eager precondition calls with n = 0, 1, 2
    (the fused call may order reductions differently: 1e-12)."""
    N, batch, L, passes, burn, seed = 4, 16, 16, 5, 10, 5
    liouv, _ = example_problem(N)
    net = NDM(N, 1, 1).init_random_pars(np.random.default_rng(4), sigma=0.01)
    p0 = net.flat.copy()
    algo = SR(precondition_type="multiplicative", lambda0=2.0, b=0.5, lambda_min=0.3)
    a = BatchedSampler(net, MetropolisSampler(LocalRule(), L, passes, burn=burn, seed=seed), liouv, algo,
                       batch_sz=batch)
    net_b = NDM(N, 1, 1)
    net_b.flat[:] = p0
    b = BatchedSampler(net_b, MetropolisSampler(LocalRule(), L, passes, burn=burn, seed=seed), liouv, algo,
                       batch_sz=batch)
    try:
        for n in range(3):
            la = a.step(Descent(0.02))
            lb, prec = sample_(b)
            precondition_(prec, algo, n)
            update_(Descent(0.02), b)
            assert abs(la.mean - lb.mean) <= 1e-12 * abs(lb.mean)
        a.net.flat[:] = a.ctx.get_params()
        np.testing.assert_allclose(a.net.flat, net_b.flat, rtol=0, atol=1e-12 * np.abs(net_b.flat).max())
    finally:
        a.close()
        b.close()
