"""The host-side mirror of the reference's VMC API, run as the reference's
example runs it (examples/dissipative_ising1d.jl:33-55: BatchedSampler,
add_observable!, sample!, compute_observables, precondition!,
Optimisers.update!(optimizer, net, grad)), against the same steps made
directly through the C ABI with the same seeds.  Both paths issue the same
kernels in the same order and every reduction is order-independent or
fixed-order, so the statistics, observables, steps and parameters agree bit
for bit."""
import numpy as np
import pytest

from nqhip import (NDM, SR, BatchedSampler, Context, Descent, LocalRule, MetropolisSampler, _lib,
                   add_observable_, compute_observables, hilbert, operators, precondition_, sample_, update_)

pytestmark = pytest.mark.gpu


def example_problem(N, g=0.4, V=2.0):
    h = hilbert.HomogeneousSpin(N)
    H = operators.LocalOperator(h)
    Sx = operators.LocalOperator(h)
    Sy = operators.LocalOperator(h)
    Sz = operators.LocalOperator(h)
    jumps = []
    for i in range(1, N + 1):
        H = H + (g / 2.0) * operators.sigmax(h, i)
        H = H + (V / 4.0) * (operators.sigmaz(h, i) * operators.sigmaz(h, i % N + 1))
        Sx = Sx + (1.0 / N) * operators.sigmax(h, i)
        Sy = Sy + (1.0 / N) * operators.sigmay(h, i)
        Sz = Sz + (1.0 / N) * operators.sigmaz(h, i)
        jumps.append(operators.sigmam(h, i))
    return operators.liouvillian(H, jumps), {"Sx": Sx, "Sy": Sy, "Sz": Sz}


# (N, batch_sz, chain_length, passes, burn, precondition, iterations); "example" is
# examples/dissipative_ising1d.jl as written: N = 7, MetropolisSampler(LocalRule(),
# 125, N, burn=100), SR(ϵ=0.001, algorithm=sr_cholesky), batch_sz = 16
CASES = {"shift": (5, 16, 25, 5, 20, "shift", 3), "multiplicative": (5, 16, 25, 5, 20, "multiplicative", 3),
         "example": (7, 16, 125, 7, 100, "shift", 4)}


@pytest.mark.parametrize("case", sorted(CASES))
def test_driver_matches_the_abi(gpu_ok, case):
    N, batch, L, passes, burn, precondition_type, iters = CASES[case]
    seed = 77
    liouv, obs = example_problem(N)
    net = NDM(N, 1, 1, act="softplus").init_random_pars(np.random.default_rng(3), sigma=0.01)
    p0 = net.flat.copy()
    sampl = MetropolisSampler(LocalRule(), L, passes, burn=burn, seed=seed)
    algo = SR(0.001, algorithm="cholesky", precondition_type=precondition_type, lambda0=1.0, b=0.9,
              lambda_min=1e-3)
    optimizer = Descent(0.01)
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
        for i in range(1, iters + 1):
            ldata, prec = sample_(is_)
            ob = compute_observables(is_)
            grad = precondition_(prec, algo, i)
            update_(optimizer, net, grad)

            st = ctx.sample_grad()
            ref_ob = {name: ctx.observable(op, batch, L, passes, burn) for name, op in flat_obs.items()}
            eps = 0.001 if precondition_type == "shift" else max(1.0 * 0.9 ** i, 1e-3)
            kind = _lib.SHIFT_ADD if precondition_type == "shift" else _lib.SHIFT_DIAG
            d = ctx.precondition(eps, kind)
            ctx.update(0.01)

            assert ldata.mean == st[0] and ldata.error == st[1]
            for name, r in ref_ob.items():
                assert ob[name].mean == complex(r[0], r[5]) and ob[name].error == r[1]
                assert abs(ob[name].mean.real) <= 1.0 + 1e-12
            np.testing.assert_array_equal(grad, d)
            np.testing.assert_array_equal(net.flat, ctx.get_params())
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


def test_hamiltonian_energy_as_the_reference_test(gpu_ok):
    """test/solvers/ising1d_closed.jl: a ket RBM(N, 2) on the closed Ising
    chain, BatchedSampler(net, sampler, H, SR(); batch_sz=32), one sample!:
    the sampled energy agrees with the exact <psi|H|psi>/<psi|psi> of the
    same network (dense H from the operator algebra, psi from the oracle's
    RBM), here within 4 standard errors of the sampled statistics; then a
    few SR steps lower the energy."""
    from nqhip import RBM
    from oracle import rbm_ref
    N, g, V = 4, 0.7, 2.0
    h = hilbert.HomogeneousSpin(N)
    H = operators.LocalOperator(h)
    for i in range(1, N + 1):
        H = H + (g / 2.0) * operators.sigmax(h, i)
        H = H + (V / 4.0) * (operators.sigmaz(h, i) * operators.sigmaz(h, i % N + 1))
    net = RBM(N, 2).init_random_pars(np.random.default_rng(6), sigma=0.01)
    with pytest.warns(UserWarning):
        sampl = MetropolisSampler(LocalRule(), 300, N, burn=30, seed=11)
    algo = SR(0.001)
    is_ = BatchedSampler(net, sampl, H, algo, batch_sz=32)
    try:
        res, prec = sample_(is_)
        keys = np.arange(2 ** N, dtype=np.uint64)
        psi = np.exp(rbm_ref.ket_logpsi(net.flat, keys, N, net.nh))
        Hm = operators.to_matrix(H)
        E_ex = (np.vdot(psi, Hm @ psi) / np.vdot(psi, psi)).real
        assert 0 < res.error < 0.1
        assert abs(res.mean - E_ex) < 4 * res.error
        first = res.mean
        for it in range(1, 31):
            grad = precondition_(prec, algo, it)
            assert np.iscomplexobj(grad) and np.all(np.isfinite(grad))
            update_(Descent(0.05), is_, grad)
            res, prec = sample_(is_)
        assert res.mean < first - 0.1
    finally:
        is_.close()


def test_rbmsplit_driver_matches_the_abi(gpu_ok):
    """RBMSplit (RBMSplit.jl) through BatchedSampler / sample! / precondition! /
    update! against the same steps through the C ABI (complex SR: the step
    comes back complex)."""
    from nqhip import RBMSplit
    N, batch, L, passes, burn, seed = 4, 16, 16, 5, 10, 21
    liouv, _ = example_problem(N)
    net = RBMSplit(N, 1).init_random_pars(np.random.default_rng(8), sigma=0.01)
    p0 = net.flat.copy()
    algo = SR(0.01)
    is_ = BatchedSampler(net, MetropolisSampler(LocalRule(), L, passes, burn=burn, seed=seed), liouv, algo,
                         batch_sz=batch)
    ctx = Context(0, seed)
    ctx.set_problem(operators.flatten(liouv))
    ctx.set_rbmsplit(N, N)
    ctx.set_precision(_lib.PREC_FP64, _lib.PREC_FP64)
    ctx.set_params(p0.view(np.float64))
    ctx.set_sampler(batch, L, passes, burn, 0)
    try:
        for i in range(1, 3):
            ldata, prec = sample_(is_)
            grad = precondition_(prec, algo, i)
            update_(Descent(0.01), is_, grad)
            st = ctx.sample_grad()
            d = ctx.precondition(0.01)
            ctx.update(0.01)
            assert np.iscomplexobj(grad) and grad.shape == (net.n_params,)
            assert abs(ldata.mean - st[0]) <= 1e-10 * abs(st[0])
            np.testing.assert_allclose(grad, d, rtol=0, atol=1e-10 * np.abs(d).max())
            np.testing.assert_allclose(net.flat, ctx.get_params().view(np.complex128), rtol=0, atol=1e-12)
        assert not np.array_equal(net.flat, p0)
    finally:
        is_.close()
        ctx.close()
