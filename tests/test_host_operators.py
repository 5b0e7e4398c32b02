"""Host-side operator algebra and indexing (CPU): the product's flat
connection tables equal the oracle's rule restatement bit for bit, and the
dense matrices they encode equal the textbook Kronecker Liouvillian
(test/Operators/ising.jl:28-41 pattern, test/Hilbert/space.jl:5-50 pattern)."""
import itertools

import numpy as np
import pytest

from nqhip import hilbert, operators, problems
from oracle import ops_ref


def assert_tables_equal(flat, ref):
    for f in ("term_kind", "term_op_l", "term_op_r", "op_site_off", "op_sites", "op_row_off",
              "row_ent_off", "ent_flip"):
        a, b = getattr(flat, f), ref[f]
        assert a.shape == b.shape and np.array_equal(a, b), f
    assert np.array_equal(flat.ent_mel, ref["ent_mel"])


@pytest.mark.parametrize("n,periodic", [(2, True), (3, True), (4, True), (4, False), (6, False), (12, True)])
def test_ising_tables_bit_exact(n, periodic):
    _, _, _, L = problems.dissipative_ising_chain(n, V=2.0, g=1.0, gamma=1.0, periodic=periodic)
    assert_tables_equal(operators.flatten(L), ops_ref.table_from_rules(
        ops_ref.ising_chain_model(n, V=2.0, g=1.0, periodic=periodic)))


@pytest.mark.parametrize("lx,ly", [(2, 2), (4, 4)])
def test_xyz_tables_bit_exact(lx, ly):
    _, _, _, L = problems.dissipative_xyz(lx, ly)
    assert_tables_equal(operators.flatten(L), ops_ref.table_from_rules(ops_ref.xyz_model(lx, ly)))


def test_ising_reference_test_case_dense():
    """test/Operators/ising.jl: N = 4 periodic, g = 0.7, V = 2, fields first,
    then bonds; H, the coherent part alone, the dissipator alone and the
    full Liouvillian against the Kronecker forms."""
    n, g, V = 4, 0.7, 2.0
    hilb, H = problems.ising_hamiltonian(n, g=g, V=V, periodic=True)
    m = ops_ref.ising_chain_model(n, V=V, g=g, fields_first=True)
    Hd = sum(ops_ref.full_op(n, s, M) for s, M in m.h_terms)
    assert np.max(np.abs(operators.to_matrix(H) - Hd)) < 1e-14
    ops = [operators.sigmam(hilb, i) for i in range(1, n + 1)]
    Lfull = operators.liouvillian(H, ops)
    assert np.max(np.abs(operators.to_matrix(Lfull) - ops_ref.dense_liouvillian_kron(m))) < 1e-13
    m0 = ops_ref.Model(n, m.h_terms, [])
    Lh = operators.liouvillian(H, [])
    assert np.max(np.abs(operators.to_matrix(Lh) - ops_ref.dense_liouvillian_kron(m0))) < 1e-13
    md = ops_ref.Model(n, [], m.jumps)
    Ld = operators.liouvillian(operators.LocalOperator(hilb), ops)
    assert np.max(np.abs(operators.to_matrix(Ld) - ops_ref.dense_liouvillian_kron(md))) < 1e-13


def test_row_rules_threshold_order_diagonal():
    hilb = hilbert.HomogeneousSpin(2)
    M = np.array([[0.0, 1e-6, 2.0, 0.0],
                  [3.0, 0.5, 0.0, 2e-5],
                  [0.0, 0.0, 0.0, 0.0],
                  [1.0, 9e-6, 1e-5, 4.0]])
    op = operators.KLocalOperatorRow(hilb, [1, 2], M)
    assert [m for m, _ in op.conns[0]] == [0.0, 2.0]                 # diagonal first, 1e-6 dropped
    assert [m for m, _ in op.conns[1]] == [0.5, 3.0, 2e-5]
    assert [m for m, _ in op.conns[2]] == [0.0]                      # zero diagonal kept
    assert [m for m, _ in op.conns[3]] == [4.0, 1.0, 1e-5]           # 9e-6 < 1e-5 dropped
    assert op.conns[3][1][1] == ((1, -1.0), (2, -1.0))


def test_same_site_sum_merges_and_products_use_kron():
    hilb = hilbert.HomogeneousSpin(3)
    a = 0.3 * operators.sigmax(hilb, 2)
    b = operators.sigmaz(hilb, 2)
    s = a + b
    assert isinstance(s, operators.KLocalOperator) and s.sites == [2]
    assert np.allclose(s.mat, 0.3 * ops_ref.SX + ops_ref.SZ)
    p = operators.sigmax(hilb, 3) * operators.sigmay(hilb, 1)
    assert p.sites == [1, 3]
    assert np.allclose(p.mat, np.kron(ops_ref.SX, ops_ref.SY))
    total = operators.LocalOperator(hilb) + a + operators.sigmaz(hilb, 1) + b
    assert [o.sites for o in operators.operators(total)] == [[2], [1]]


def test_indexing_round_trips():
    for n in range(1, 7):
        h = hilbert.HomogeneousSpin(n)
        for i in range(1, 2 ** n + 1):
            assert hilbert.toint(hilbert.state_from_int(h, i), h) == i
        for k in range(4 ** n):
            s, e = hilbert.states_from_key(k, n)
            assert hilbert.key_from_states(s, e) == k
            # doubled index = (toint(e)-1) 2^N + toint(s) = key + 1 (SuperHilbert.jl:50-56)
            assert (hilbert.toint(e, h) - 1) * 2 ** n + hilbert.toint(s, h) == k + 1
    keys = np.arange(4 ** 4, dtype=np.uint64)
    s, e = hilbert.keys_to_states(keys, 4)
    assert np.array_equal(hilbert.states_to_keys(s, e), keys)


def test_local_index_matches_table_rows():
    n = 5
    for sites in itertools.permutations(range(1, n + 1), 2):
        for k in range(2 ** n):
            s = np.array([1.0 if (k >> i) & 1 else -1.0 for i in range(n)])
            want = 1 + sum(((k >> (site - 1)) & 1) << m for m, site in enumerate(sites))
            assert hilbert.local_index(s, sites) == want


def test_flipat_doubled_sites():
    n = 3
    assert hilbert.flipat(0, 1, n) == 1
    assert hilbert.flipat(0, n + 1, n) == 1 << n
    with pytest.raises(IndexError):
        hilbert.flipat(0, 2 * n + 1, n)


def test_square_lattice_edges():
    es = problems.square_edges(4, 4)
    assert len(es) == 32 and es[:4] == [(1, 2), (1, 4), (1, 5), (1, 13)]
    assert all(i < j for i, j in es) and es == sorted(es)
    deg = np.zeros(17, dtype=int)
    for i, j in es:
        deg[i] += 1
        deg[j] += 1
    assert np.all(deg[1:] == 4)
