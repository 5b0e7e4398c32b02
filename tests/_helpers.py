"""Shared builders for the parity tests: models in both the product's and the
oracle's construction, seeded parameters and configurations."""
import numpy as np

from nqhip import operators, problems
from oracle import nqref, ops_ref


def ising(n, periodic=True, V=2.0, g=1.0, gamma=1.0):
    """(product flat table, oracle table, oracle model) of the dissipative
    Ising chain."""
    _, _, _, L = problems.dissipative_ising_chain(n, V=V, g=g, gamma=gamma, periodic=periodic)
    m = ops_ref.ising_chain_model(n, V=V, g=g, gamma=gamma, periodic=periodic)
    return operators.flatten(L), nqref.Table(ops_ref.table_from_rules(m)), m


def xyz(lx, ly, gamma=1.0):
    _, _, _, L = problems.dissipative_xyz(lx, ly, gamma=gamma)
    m = ops_ref.xyz_model(lx, ly, gamma=gamma)
    return operators.flatten(L), nqref.Table(ops_ref.table_from_rules(m)), m


def params(n_params, seed, sigma=0.1):
    return np.random.default_rng(seed).normal(0.0, sigma, n_params)


def random_keys(n, count, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2 ** (2 * n), size=count, dtype=np.uint64)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
