"""The C-ABI library loads and exports every entry point include/nqhip.h
declares (CPU: no compute calls), and the host API mirror behaves like the
reference's constructors."""
import os
import re
import subprocess
import warnings

import pytest

import nqhip
from nqhip import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nqhip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nqh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "nqh_sr_iteration" in names and "nqh_local_estimator" in names
    assert set(names) == set(_lib.SIGNATURES), "binding and header disagree"


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libnqhip.so is not built (run __graft_entry__.build())")
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (nqh_[a-z_0-9]+)$", out, flags=re.M))
    missing = [n for n in declared() if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_abi_version():
    lib = _lib.load()
    assert lib.nqh_abi_version() == 1


def test_metropolis_even_passes_bumped_to_odd():
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        s = nqhip.MetropolisSampler(nqhip.LocalRule(), 10, 4, burn=2, seed=1)
    assert s.passes == 5 and any("non-ergodic" in str(x.message) for x in w)


def test_ndm_flat_layout_and_init():
    import numpy as np
    net = nqhip.NDM(4, 2, 1)
    assert net.n_params == 2 * 4 + 2 * 8 + 4 + 2 * 8 * 4 + 2 * 4 * 4
    net.flat[:] = np.arange(net.n_params)
    assert net.b_mu[0] == 0 and net.h_mu[0] == 4 and net.w_mu[1, 0] == 13 and net.w_mu[0, 1] == 20
    assert net.u_la[-1, -1] == net.n_params - 1
    net.init_random_pars(np.random.default_rng(0), sigma=1e-4)
    assert abs(np.std(net.flat) - 1e-2) < 3e-3


def test_sr_mirror_algorithms():
    """SR(algorithm=...) accepts the solvers built (sr_cholesky, sr_cg, sr_minres, sr_qlp, sr_lsq)
    and maps them to the ABI's solver codes; others are refused."""
    from nqhip import SR
    assert SR(0.01).solver == _lib.SOLVER_CHOLESKY
    cg = SR(0.01, algorithm="cg", precision=1e-6)
    assert cg.solver == _lib.SOLVER_CG and cg.precision == 1e-6 and cg.maxiter == 0
    assert SR(0.01, algorithm="minres").solver == _lib.SOLVER_MINRES
    assert SR(0.01, algorithm="qlp").solver == _lib.SOLVER_MINRES_QLP
    assert SR(0.01, algorithm="lsq").solver == _lib.SOLVER_LSQR
    assert SR(0.01, algorithm="div").solver == SR(0.01, algorithm="diag").solver == _lib.SOLVER_CHOLESKY
    with pytest.raises(NotImplementedError):
        SR(0.01, algorithm="pivcholesky")


def test_sr_mirror_shift_schedules():
    """precondition_type (SR.jl:4, SRDirect.jl:56-70): none, shift and the
    multiplicative lambda = max(lambda0 b^n, lambda_min) on diag(S)."""
    from nqhip import SR
    assert SR(precondition_type="none").shift(5) == (0.0, _lib.SHIFT_ADD)
    assert SR(0.02).shift(5) == (0.02, _lib.SHIFT_ADD)
    m = SR(precondition_type="multiplicative")     # the reference's defaults 100, 0.95, 1e-4
    assert m.shift(0) == (100.0, _lib.SHIFT_DIAG)
    assert abs(m.shift(10)[0] - 100.0 * 0.95 ** 10) < 1e-12
    assert m.shift(1000) == (1e-4, _lib.SHIFT_DIAG)
    with pytest.raises(ValueError):
        SR(precondition_type="sr_shift")
