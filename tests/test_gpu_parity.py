"""Device parity against the oracle: log rho, O, connections, the local
estimator, the gradient post-processing, the sampler trajectory and the SR
solve, on identical host-fed inputs.

Tolerance (north star): 1e-10 relative for the FP64 kernels; configuration
and connection indexing bit-exact."""
import numpy as np
import pytest
import scipy.linalg

from _helpers import ising, params, random_keys, rel_err, xyz
from nqhip import Context, _lib
from oracle import nqref

pytestmark = pytest.mark.gpu
TOL = 1e-10


def make_ctx(flat, N, nh, na, p, seed=7):
    ctx = Context(0, seed)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    return ctx


@pytest.mark.parametrize("N,nh,na", [(3, 3, 3), (4, 8, 4), (4, 4, 8)])
def test_logpsi_grad_all_states(gpu_ok, ref, N, nh, na):
    flat, _, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=N * 100 + nh)
    keys = nqref.all_keys(N)
    ctx = make_ctx(flat, N, nh, na, p)
    lp, O = ctx.logpsi_grad(keys)
    lp_r, O_r = ref.logpsi_grad(N, nh, na, p, keys)
    assert rel_err(lp, lp_r) < TOL
    assert rel_err(O, O_r) < TOL
    assert rel_err(ctx.logpsi(keys), lp_r) < TOL


def test_logpsi_grad_xyz44_shapes(gpu_ok, ref):
    N, nh, na = 16, 32, 32
    flat, _, _ = xyz(4, 4)
    p = params(ref.ndm_np(N, nh, na), seed=11, sigma=0.05)
    keys = random_keys(N, 1024, seed=3)
    ctx = make_ctx(flat, N, nh, na, p)
    lp, O = ctx.logpsi_grad(keys)
    lp_r, O_r = ref.logpsi_grad(N, nh, na, p, keys)
    assert rel_err(lp, lp_r) < TOL
    assert rel_err(O, O_r) < TOL


@pytest.mark.parametrize("model", ["ising4", "ising5open", "xyz22", "xyz44"])
def test_connections_bit_exact(gpu_ok, ref, model):
    if model == "ising4":
        flat, tab, _ = ising(4)
        keys = nqref.all_keys(4)
    elif model == "ising5open":
        flat, tab, _ = ising(5, periodic=False)
        keys = random_keys(5, 300, seed=5)
    elif model == "xyz22":
        flat, tab, _ = xyz(2, 2)
        keys = nqref.all_keys(4)
    else:
        flat, tab, _ = xyz(4, 4)
        keys = random_keys(16, 200, seed=9)
    N = flat.n_sites
    ctx = Context(0, 1)
    ctx.set_problem(flat)
    cap = flat.max_connections()
    cnt, ko, mo = ctx.connections(keys, cap)
    for b, k in enumerate(keys):
        rk, rm = ref.connections(tab, int(k))
        assert cnt[b] == len(rk)
        assert np.array_equal(ko[b, :cnt[b]], rk)
        assert np.array_equal(mo[b, :cnt[b]], rm)
    assert N == tab.n_sites


@pytest.mark.parametrize("N,nh,na,periodic", [(3, 3, 3, True), (4, 4, 8, True), (4, 8, 4, False)])
def test_local_estimator_all_states(gpu_ok, ref, N, nh, na, periodic):
    flat, tab, _ = ising(N, periodic=periodic)
    p = params(ref.ndm_np(N, nh, na), seed=N + nh)
    keys = nqref.all_keys(N)
    ctx = make_ctx(flat, N, nh, na, p)
    L, G = ctx.local_estimator(keys, grad=True)
    L_r, G_r = ref.local_estimator(tab, nh, na, p, keys)
    assert rel_err(L, L_r) < TOL
    assert rel_err(G, G_r) < TOL


def test_local_estimator_xyz44(gpu_ok, ref):
    N, nh, na = 16, 32, 32
    flat, tab, _ = xyz(4, 4)
    p = params(ref.ndm_np(N, nh, na), seed=21, sigma=0.05)
    keys = random_keys(N, 64, seed=4)
    ctx = make_ctx(flat, N, nh, na, p)
    L, G = ctx.local_estimator(keys, grad=True)
    L_r, G_r = ref.local_estimator(tab, nh, na, p, keys)
    assert rel_err(L, L_r) < TOL
    assert rel_err(G, G_r) < TOL


# 4096 samples and more take the estimator's two-launch path (walk, then
# forces beside the S build); fewer take the single launch
@pytest.mark.parametrize("N,nh,na,chains,Lc", [(4, 4, 4, 16, 8), (6, 6, 6, 16, 20), (4, 4, 4, 512, 8)])
def test_grad_from_samples(gpu_ok, ref, N, nh, na, chains, Lc):
    flat, tab, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=31)
    keys = random_keys(N, chains * Lc, seed=6)
    ctx = make_ctx(flat, N, nh, na, p)
    st, S, F, Ob = ctx.grad_from_samples(keys, chains, Lc)
    lp, O = ref.logpsi_grad(N, nh, na, p, keys)
    L, dL = ref.local_estimator(tab, nh, na, p, keys)
    S_r, F_r, Ob_r, st_r = ref.grad_step(chains, Lc, O, L, dL)
    assert rel_err(st, st_r) < TOL
    assert rel_err(Ob, Ob_r) < TOL
    assert rel_err(F, F_r) < TOL
    assert rel_err(S, S_r) < TOL


@pytest.mark.parametrize("compat", [False, True])
def test_sampler_trajectory_bit_exact(gpu_ok, ref, compat):
    N, nh, na = 4, 4, 4
    flat, _, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=41, sigma=0.3)
    chains, Lc, passes, burn, seed = 16, 10, 5, 7, 1234
    ctx = make_ctx(flat, N, nh, na, p, seed=seed)
    ctx.set_sampler(chains, Lc, passes, burn, flags=1 if compat else 0)
    ctx.set_iteration(3)
    keys, acc = ctx.sample()
    keys_r, acc_r = ref.metropolis(N, nh, na, p, chains, Lc, passes, burn, seed, it=3, compat=compat)
    assert np.array_equal(keys, keys_r)
    assert acc == acc_r


def test_sr_solve_matches_lapack(gpu_ok, ref):
    N, nh, na = 6, 6, 6
    chains, Lc = 16, 40
    flat, tab, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=51)
    keys = random_keys(N, chains * Lc, seed=8)
    ctx = make_ctx(flat, N, nh, na, p)
    st, S, F, _ = ctx.grad_from_samples(keys, chains, Lc)
    delta = ctx.precondition(0.01)
    delta_r = nqref.sr_solve(S, F, 0.01)
    assert rel_err(delta, delta_r) < 1e-9


@pytest.mark.parametrize("kind", [_lib.SHIFT_ADD, _lib.SHIFT_DIAG])
def test_split_solve_matches_lapack_c2_shapes(gpu_ok, ref, kind):
    """The NDM solve factors [lambda | Pi] and [mu | Pi] side by side and then
    the Pi Schur complement (S's lambda-mu block is zero).  At the N = 12,
    alpha = 2 shapes (nl = 324, npi = 600: ragged panels in both phases) it
    matches LAPACK on the full S, and the unsplit path (taken with
    refinement) gives the same step."""
    N, nh, na = 12, 24, 24
    chains, Lc = 128, 16
    flat, tab, _ = ising(N)
    p = params(ref.ndm_np(N, nh, na), seed=52, sigma=0.05)
    keys = random_keys(N, chains * Lc, seed=9)
    ctx = make_ctx(flat, N, nh, na, p)
    _, S, F, _ = ctx.grad_from_samples(keys, chains, Lc)
    Ssh = S + (0.01 * np.eye(len(F)) if kind == _lib.SHIFT_ADD else 0.01 * np.diag(np.diag(S)))
    delta_r = scipy.linalg.cho_solve(scipy.linalg.cho_factor(Ssh, lower=True), F)
    d_split = ctx.precondition(0.01, kind, 0)
    assert rel_err(d_split, delta_r) < 1e-9
    d_full = ctx.precondition(0.01, kind, 1)
    assert rel_err(d_split, d_full) < 1e-9
    ctx.close()


@pytest.mark.parametrize("kind", [_lib.SHIFT_ADD, _lib.SHIFT_DIAG])
def test_split_solve_matches_lapack_c3_shapes(gpu_ok, kind):
    """The benchmark's shapes: 4x4 XYZ, NDM alpha = 2 (Np = 2176: nl = 560,
    npi = 1056, 18 + 33 panels, the look-ahead tile's three-wave diagonal
    factor, the Schur update and both folded first-block factors), S and F
    of 4096 x 4 samples in FP64; the split step against LAPACK on the full
    shifted S."""
    N, nh, na = 16, 32, 32
    flat, _, _ = xyz(4, 4)
    ctx = Context(0, 5)
    ctx.set_problem(flat)
    ctx.set_ndm(N, nh, na)
    ctx.set_params(params(ctx.np_, seed=53, sigma=0.05))
    keys = random_keys(N, 4096 * 4, seed=10)
    _, S, F, _ = ctx.grad_from_samples(keys, 4096, 4)
    Ssh = S + (0.01 * np.eye(len(F)) if kind == _lib.SHIFT_ADD else 0.01 * np.diag(np.diag(S)))
    delta_r = scipy.linalg.cho_solve(scipy.linalg.cho_factor(Ssh, lower=True), F)
    assert rel_err(ctx.precondition(0.01, kind, 0), delta_r) < 1e-9
    ctx.close()
