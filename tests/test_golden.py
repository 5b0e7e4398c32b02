"""The committed golden fixtures (tests/golden/*.npz, written and cross-checked
by tests/golden/make_golden.py): the oracle and the host operator algebra
reproduce them on the CPU; the device path reproduces them on the GPU."""
import glob
import os

import numpy as np
import pytest

from _helpers import rel_err
from nqhip import Context, hilbert, operators, problems
from oracle import made_ref, nqref, ops_ref

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = sorted(glob.glob(os.path.join(HERE, "golden", "ising_*.npz")))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def product_liouvillian(g):
    _, _, _, L = problems.dissipative_ising_chain(int(g["N"]), V=float(g["V"]), g=float(g["g"]),
                                                   gamma=float(g["gamma"]), periodic=bool(g["periodic"]))
    return L


def product_table(g):
    return operators.flatten(product_liouvillian(g))


def split_conns(g):
    off = np.concatenate([[0], np.cumsum(g["conn_count"])])
    return [(g["conn_keys"][a:b], g["conn_mel"][a:b]) for a, b in zip(off[:-1], off[1:])]


def test_fixtures_present():
    assert len(CASES) == 3
    assert os.path.exists(os.path.join(HERE, "golden", "made_d8.npz"))


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_oracle_reproduces_golden(ref, path):
    g = load(path)
    N, nh, na, p, keys = int(g["N"]), int(g["nh"]), int(g["na"]), g["params"], g["keys"]
    model = ops_ref.ising_chain_model(N, V=float(g["V"]), g=float(g["g"]), gamma=float(g["gamma"]),
                                      periodic=bool(g["periodic"]))
    table = nqref.Table(ops_ref.table_from_rules(model))
    lp, O = ref.logpsi_grad(N, nh, na, p, keys)
    L, dL = ref.local_estimator(table, nh, na, p, keys)
    assert rel_err(lp, g["logpsi"]) < 1e-13
    assert rel_err(O, g["O"]) < 1e-13
    assert rel_err(L, g["Lloc"]) < 1e-12
    assert rel_err(dL, g["dLloc"]) < 1e-12
    S, F, Ob, st = ref.grad_step(len(keys) // 4, 4, O, L, dL)
    assert rel_err(st, g["stats"]) < 1e-12
    assert rel_err(S, g["S"]) < 1e-12
    assert rel_err(F, g["F"]) < 1e-12
    assert rel_err(nqref.sr_solve(S, F, float(g["sr_eps"])), g["sr_delta"]) < 1e-10


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_host_operator_tables_reproduce_golden_connections(path):
    """The product's KLocalLiouvillian enumerates the golden connection lists
    (targets bit-exact, in order; mels to 1e-15)."""
    g = load(path)
    L = product_liouvillian(g)
    N = int(g["N"])
    for k, (kc, mc) in zip(g["keys"][:4], split_conns(g)):
        s, e = hilbert.states_from_key(int(k), N)
        got = operators.enumerate_connections(L, s, e)
        tk = [hilbert.key_from_states(operators._apply(s, rc), operators._apply(e, cc)) for _, rc, cc in got]
        assert np.array_equal(np.array(tk, dtype=np.uint64), kc)
        assert np.allclose(np.array([c[0] for c in got]), mc, rtol=0, atol=1e-15)


def test_made_oracle_reproduces_golden():
    with np.load(os.path.join(HERE, "golden", "made_d8.npz"), allow_pickle=False) as z:
        D, H, th = int(z["D"]), int(z["H"]), z["theta"]
        assert rel_err(made_ref.log_q(th, z["keys"], D, H), z["logq"]) < 1e-13
        nll, gr = made_ref.nll_and_grad(th, z["batch"], D, H)
        assert abs(nll - float(z["nll"])) < 1e-12
        assert rel_err(gr, z["grad"]) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_device_reproduces_golden(gpu_ok, path):
    """FP64 device path against the golden vectors: log rho, O, Lloc, dLloc,
    S, F, O-bar within 1e-10 relative, connections bit-exact, the shifted
    Cholesky solve within 1e-9 relative."""
    g = load(path)
    N, nh, na, p, keys = int(g["N"]), int(g["nh"]), int(g["na"]), g["params"], g["keys"]
    ctx = Context(0, 1)
    ctx.set_problem(product_table(g))
    ctx.set_ndm(N, nh, na)
    ctx.set_params(p)
    lp, O = ctx.logpsi_grad(keys)
    assert rel_err(lp, g["logpsi"]) < 1e-10
    assert rel_err(O, g["O"]) < 1e-10
    cnt, ko, mo = ctx.connections(keys[:4], 256)
    for i, (kc, mc) in enumerate(split_conns(g)):
        assert cnt[i] == len(kc)
        assert np.array_equal(ko[i, :cnt[i]], kc)
        assert rel_err(mo[i, :cnt[i]], mc) < 1e-15
    L, dL = ctx.local_estimator(keys, grad=True)
    assert rel_err(L, g["Lloc"]) < 1e-10
    assert rel_err(dL, g["dLloc"]) < 1e-10
    st, S, F, Ob = ctx.grad_from_samples(keys, len(keys) // 4, 4)
    assert rel_err(S, g["S"]) < 1e-10
    assert rel_err(F, g["F"]) < 1e-10
    assert rel_err(Ob, g["Obar"]) < 1e-10
    assert rel_err(st, g["stats"]) < 1e-10
    d = ctx.precondition(float(g["sr_eps"]))
    assert rel_err(d, g["sr_delta"]) < 1e-9
    ctx.close()
