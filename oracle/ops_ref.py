"""ORACLE (test infrastructure only) -- connection tables of the Liouvillian.

This is the CPU restatement the parity tests check the product against; only
`tests/`, `__graft_entry__.smoke()` and `bench.py`'s CPU-baseline leg may use
it.  It is written independently of the product's operator algebra: it works
directly on dense local matrices and restates the reference's *rules*:

* rows of a local operator: diagonal first (always present, even if zero),
  then off-diagonal entries in increasing column order with |mel| >= 1e-5
  (src/Operators/Operators/KLocalOperator.jl:67-109, threshold `10e-6` :86);
* the row index of a configuration on sites [s_1..s_k] is
  sum_m bit(s_m) 2^(m-1) (src/Hilbert/HomogeneousSpin.jl:171-179), bit = 1
  for v = +1;
* HnH = H + sum_L -0.5i L^dag L, same-site terms summed, kept in order of
  first appearance (KLocalOperatorSum.jl:74-85, KLocalLiouvillian.jl:27-33);
* enumeration: for every HnH term the row-side rows of -i*M, then for every
  HnH term the column-side rows of +i*conj(M^T), then for every jump operator
  the nested loop rows(L) x rows(conj L) with product mels
  (KLocalLiouvillian.jl:9-19,46-52; KLocalOperatorTensor.jl:129-156);
* exact-zero mels are skipped by the accumulator (AccumulatorObsGrad.jl:67).

Two pins live here as well:
* `dense_liouvillian_kron`: the textbook Lindblad superoperator in the
  reference's vectorisation, built with Kronecker products of full-space
  matrices (no connection logic at all);
* `table_from_rules`: the flat table format of include/nqhip.h built from the
  rules above, for the C oracle.

Parity status: pinned by the Kronecker closed form (values and targets) and by
the rule restatement (order); the reference has no golden connection lists.
"""
from __future__ import annotations

import numpy as np

THRESH = 1e-5

SX = np.array([[0, 1], [1, 0]], dtype=complex)
SY = np.array([[0, -1j], [1j, 0]], dtype=complex)
SZ = np.array([[1, 0], [0, -1]], dtype=complex)
SM = np.array([[0, 0], [1, 0]], dtype=complex)   # SimpleOperators.jl:68-77: mat[2,1] = 1
SP = SM.T.copy()
I2 = np.eye(2, dtype=complex)


def pair_matrix(A, B):
    """Local matrix of A on site a and B on site b (a < b), local bit 0 = site a:
    M[(ra, rb), (ca, cb)] = A[ra, ca] B[rb, cb] with index r = ra + 2 rb."""
    M = np.zeros((4, 4), dtype=complex)
    for r in range(4):
        for c in range(4):
            M[r, c] = A[r & 1, c & 1] * B[r >> 1, c >> 1]
    return M


class Model:
    """H given as ordered (sites, local matrix) terms (1-based sites, sorted
    within a term), plus jump operators as (sites, local matrix)."""

    def __init__(self, n, h_terms, jumps):
        self.n = n
        self.h_terms = [(tuple(s), np.array(m, dtype=complex)) for s, m in h_terms]
        self.jumps = [(tuple(s), np.array(m, dtype=complex)) for s, m in jumps]

    def hnh_terms(self):
        order = []
        mats = {}
        for s, m in self.h_terms:
            if s not in mats:
                order.append(s)
                mats[s] = m.copy()
            else:
                mats[s] = mats[s] + m
        for s, L in self.jumps:
            t = -0.5j * (L.conj().T @ L)
            if s not in mats:
                order.append(s)
                mats[s] = t
            else:
                mats[s] = mats[s] + t
        return [(s, mats[s]) for s in order]


def ising_chain_model(n, V=2.0, g=1.0, gamma=1.0, periodic=True, fields_first=False):
    """Dissipative Ising chain in the construction order of
    examples/dissipative_ising1d.jl:17-26 (or all fields first when
    `fields_first`, as test/Operators/ising.jl:16-26 does)."""
    h = []
    bonds = []
    for i in range(1, n + 1):
        field = ((i,), (g / 2.0) * SX)
        if i < n:
            bond = ((i, i + 1), (V / 4.0) * pair_matrix(SZ, SZ))
        elif periodic and n >= 2:
            bond = ((1, n), (V / 4.0) * pair_matrix(SZ, SZ))
        else:
            bond = None
        if fields_first:
            h.append(field)
            if bond:
                bonds.append(bond)
        else:
            h.append(field)
            if bond:
                h.append(bond)
    h.extend(bonds)
    jumps = [((i,), np.sqrt(gamma) * SM) for i in range(1, n + 1)]
    return Model(n, h, jumps)


def square_edges(lx, ly):
    es = set()
    for y in range(ly):
        for x in range(lx):
            v = 1 + x + lx * y
            for nx, ny in (((x + 1) % lx, y), (x, (y + 1) % ly)):
                w = 1 + nx + lx * ny
                if w != v:
                    es.add((min(v, w), max(v, w)))
    return sorted(es)


def xyz_model(lx, ly, Jx=0.9, Jy=1.0, Jz=1.1, gamma=1.0):
    h = []
    for (i, j) in square_edges(lx, ly):
        h.append(((i, j), Jx * pair_matrix(SX, SX)))
        h.append(((i, j), Jy * pair_matrix(SY, SY)))
        h.append(((i, j), Jz * pair_matrix(SZ, SZ)))
    jumps = [((i,), np.sqrt(gamma) * SM) for i in range(1, lx * ly + 1)]
    return Model(lx * ly, h, jumps)


def rows_of(M):
    """List (per local row r) of (mel, local flip mask)."""
    d = M.shape[0]
    out = []
    for r in range(d):
        row = [(complex(M[r, r]), 0)]
        for c in range(d):
            if c != r and abs(M[r, c]) >= THRESH:
                row.append((complex(M[r, c]), r ^ c))
        out.append(row)
    return out


def _local_row(key_bits, sites):
    r = 0
    for m, s in enumerate(sites):
        r |= ((key_bits >> (s - 1)) & 1) << m
    return r


def _global_mask(local_mask, sites):
    g = 0
    for m, s in enumerate(sites):
        if (local_mask >> m) & 1:
            g |= 1 << (s - 1)
    return g


def terms(model):
    """Enumeration terms as (kind, sites_l, rows_l, sites_r, rows_r)."""
    hnh = model.hnh_terms()
    out = []
    for s, M in hnh:
        out.append((0, s, rows_of(-1j * M), None, None))
    for s, M in hnh:
        out.append((1, None, None, s, rows_of(1j * np.conj(M.T))))
    for s, L in model.jumps:
        out.append((2, s, rows_of(L), s, rows_of(np.conj(L))))
    return out


def connections(model, key, skip_zero=True, _terms=None):
    """Connections of the packed doubled key in reference order:
    list of (mel, new_key)."""
    n = model.n
    ks = key & ((1 << n) - 1)
    ke = key >> n
    out = []
    for kind, sl, rl, sr, rr in (_terms or terms(model)):
        if kind == 0:
            for mel, fm in rl[_local_row(ks, sl)]:
                out.append((mel, key ^ _global_mask(fm, sl)))
        elif kind == 1:
            for mel, fm in rr[_local_row(ke, sr)]:
                out.append((mel, key ^ (_global_mask(fm, sr) << n)))
        else:
            for ml, fl in rl[_local_row(ks, sl)]:
                for mr, fr in rr[_local_row(ke, sr)]:
                    out.append((ml * mr, key ^ _global_mask(fl, sl) ^ (_global_mask(fr, sr) << n)))
    if skip_zero:
        out = [c for c in out if c[0] != 0]
    return out


def full_op(n, sites, M):
    """Full 2^n matrix of a local operator; site 1 is the fastest index."""
    d = 2 ** n
    F = np.zeros((d, d), dtype=complex)
    k = len(sites)
    for i in range(d):
        ri = _local_row(i, sites)
        rest = i
        for s in sites:
            rest &= ~(1 << (s - 1))
        for c in range(2 ** k):
            val = M[ri, c]
            if val == 0:
                continue
            j = rest | _global_mask(c, sites)
            F[i, j] += val
    return F


def dense_liouvillian_kron(model):
    """L = -i kron(I, HnH) + i kron(HnH', I) + sum_j kron(conj(L_j), L_j) on
    vec(rho) column-major (row index fastest).  With the reference's HnH'
    (= conj(HnH^T)) in the middle term; equal to the physical conj(HnH)
    whenever HnH is symmetric, as for every model here."""
    n = model.n
    d = 2 ** n
    H = np.zeros((d, d), dtype=complex)
    for s, M in model.h_terms:
        H += full_op(n, s, M)
    HnH = H.copy()
    Ls = [full_op(n, s, L) for s, L in model.jumps]
    for L in Ls:
        HnH += -0.5j * (L.conj().T @ L)
    I = np.eye(d, dtype=complex)
    Lv = -1j * np.kron(I, HnH) + 1j * np.kron(HnH.conj().T, I)
    for L in Ls:
        Lv += np.kron(L.conj(), L)
    return Lv


def dense_from_connections(model):
    n = model.n
    D = 4 ** n
    M = np.zeros((D, D), dtype=complex)
    ts = terms(model)
    for key in range(D):
        for mel, k2 in connections(model, key, skip_zero=False, _terms=ts):
            M[key, k2] += mel
    return M


def table_from_rules(model):
    """Flat table (same layout as include/nqhip.h `nqh_optable`) from the
    rules restated above, as a dict of numpy arrays."""
    ops = []            # (sites, rows)
    kinds, ls, rs = [], [], []
    for kind, sl, rl, sr, rr in terms(model):
        kinds.append(kind)
        if sl is not None:
            ops.append((sl, rl)); ls.append(len(ops) - 1)
        else:
            ls.append(-1)
        if sr is not None:
            ops.append((sr, rr)); rs.append(len(ops) - 1)
        else:
            rs.append(-1)
    site_off, sites, row_off, ent_off, mel, flip = [0], [], [0], [0], [], []
    for s, rows in ops:
        sites.extend(x - 1 for x in s)
        site_off.append(len(sites))
        for row in rows:
            for m, fm in row:
                mel.append(m)
                flip.append(_global_mask(fm, s))
            ent_off.append(len(mel))
        row_off.append(len(ent_off) - 1)
    return dict(
        n_sites=model.n,
        term_kind=np.array(kinds, dtype=np.int32),
        term_op_l=np.array(ls, dtype=np.int32),
        term_op_r=np.array(rs, dtype=np.int32),
        op_site_off=np.array(site_off, dtype=np.int32),
        op_sites=np.array(sites, dtype=np.int32),
        op_row_off=np.array(row_off, dtype=np.int32),
        row_ent_off=np.array(ent_off, dtype=np.int32),
        ent_mel=np.array(mel, dtype=np.complex128),
        ent_flip=np.array(flip, dtype=np.uint64),
    )
