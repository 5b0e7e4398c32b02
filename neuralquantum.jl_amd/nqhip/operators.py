"""k-local operators, sums, tensors and the Liouvillian -- host-side builder.

Construction semantics follow the reference exactly, because they fix the
connection values, their order and which entries exist:

* `KLocalOperatorRow` (src/Operators/Operators/KLocalOperator.jl:54-112):
  every row stores its diagonal first, always (even when zero), then the
  off-diagonal entries in increasing column order, dropping |mel| < 1e-5
  (threshold `10e-6`, KLocalOperator.jl:86).
* `scalar * op`, `-op`, `transpose` and same-site products rebuild the rows
  from the new matrix (KLocalOperator.jl:114-115,207,251-252,266-275).
* same-site sums merge matrices and connection lists: an entry of the right
  operand whose state change already exists in the left operand's row is
  added to it, otherwise appended (KLocalOperator.jl:211-249,
  OpConnection.jl:115-125).  Sums keep insertion order
  (KLocalOperatorSum.jl:74-85).
* products of single-site operators on sites i < j live on [i, j] with
  matrix kron(B, A) (KLocalOperator.jl:292-306).
* `liouvillian(H, Ls)` builds HnH = H + sum_L -0.5i L'L, then
  HnH_l = -i (HnH (x) I), HnH_r = +i (I (x) HnH'), LLdag = sum_L L (x) conj(L)
  (KLocalLiouvillian.jl:9-33); connections are enumerated HnH_l, HnH_r,
  LLdag, operators in insertion order, tensor products as nested loops
  (KLocalLiouvillian.jl:46-52, KLocalOperatorTensor.jl:129-156).

`flatten(liouv)` turns the result into the flat connection table handed to
the device library (`nqh_optable` in include/nqhip.h).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .hilbert import HomogeneousSpin, SuperOpSpace, local_index, state_from_int

__all__ = [
    "KLocalOperator", "KLocalOperatorRow", "KLocalOperatorSum", "KLocalOperatorTensor",
    "KLocalIdentity", "KLocalOperatorZero", "KLocalLiouvillian", "LocalOperator",
    "sigmax", "sigmay", "sigmaz", "sigmam", "sigmap", "liouvillian", "operators",
    "sites", "to_matrix", "flatten", "FlatOpTable", "enumerate_connections",
    "MEL_THRESHOLD",
]

MEL_THRESHOLD = 10e-6  # KLocalOperator.jl:86 -- entries with abs(mel) < 1e-5 are dropped


def _changes_key(ch):
    """StateChanges equality is insensitive to the order of the changed sites
    (src/States/StateChanges.jl:83-89)."""
    return tuple(sorted(ch))


class KLocalOperator:
    """Operator on `sites` (1-based) stored by row: conns[r] is the list of
    [mel, changes] of local row r (0-based), changes = tuple of (site, new_value)."""

    def __init__(self, hilb, sites, mat, conns):
        self.hilb = hilb
        self.sites = list(sites)
        self.mat = np.array(mat, dtype=np.complex128)
        self.conns = conns

    def duplicate(self):
        return KLocalOperator(self.hilb, list(self.sites), self.mat.copy(),
                              [[[m, ch] for (m, ch) in row] for row in self.conns])

    # -- algebra -------------------------------------------------------
    def __neg__(self):
        return KLocalOperatorRow(self.hilb, self.sites, -self.mat)

    def __rmul__(self, a):
        if isinstance(a, (int, float, complex, np.number)):
            return KLocalOperatorRow(self.hilb, self.sites, a * self.mat)
        return NotImplemented

    def __mul__(self, other):
        if isinstance(other, (int, float, complex, np.number)):
            return KLocalOperatorRow(self.hilb, self.sites, other * self.mat)
        if isinstance(other, KLocalOperatorZero):
            return other
        if isinstance(other, KLocalOperatorSum):
            return _op_times_sum(self, other)
        if isinstance(other, KLocalOperator):
            return _kop_kop_prod(self, other)
        return NotImplemented

    def __truediv__(self, a):
        return KLocalOperatorRow(self.hilb, self.sites, self.mat / a)

    def __add__(self, other):
        if isinstance(other, KLocalOperatorZero):
            return self.duplicate()
        if isinstance(other, KLocalOperator):
            if self.sites == other.sites:
                return _add_samesite(self.duplicate(), other)
            return KLocalOperatorSum(self.hilb, [self]) + other
        if isinstance(other, KLocalOperatorSum):
            return other + self
        return NotImplemented

    def __sub__(self, other):
        if isinstance(other, KLocalOperatorZero):
            return -self
        return self + (-other)

    def transpose(self):
        return KLocalOperatorRow(self.hilb, self.sites, self.mat.T.copy())

    def conj(self):
        op = self.duplicate()
        op.mat = np.conj(op.mat)
        for row in op.conns:
            for e in row:
                e[0] = np.conj(e[0])
        return op

    def adjoint(self):
        return self.transpose().conj()

    @property
    def H(self):  # noqa: N802 -- Julia's op'
        return self.adjoint()

    def __repr__(self):
        return f"KLocalOperator(sites={self.sites}, mat={self.mat.tolist()})"


def KLocalOperatorRow(hilb, sites, mat) -> KLocalOperator:  # noqa: N802
    """Row-stored connections of `mat` on `sites` (KLocalOperator.jl:54-112)."""
    mat = np.array(mat, dtype=np.complex128)
    k = len(sites)
    dim = 2 ** k
    if mat.shape != (dim, dim):
        raise ValueError("operator matrix does not match 2^k x 2^k")
    conns = []
    for r in range(dim):
        row = [[mat[r, r], ()]]                     # diagonal first, always present
        st = [((r >> m) & 1) * 2 - 1 for m in range(k)]
        for c in range(dim):
            if c == r:
                continue
            val = mat[r, c]
            if abs(val) < MEL_THRESHOLD:
                continue
            st1 = [((c >> m) & 1) * 2 - 1 for m in range(k)]
            ch = tuple((sites[m], float(st1[m])) for m in range(k) if st[m] != st1[m])
            row.append([val, ch])
        conns.append(row)
    return KLocalOperator(hilb, sites, mat, conns)


def _add_samesite(op_l: KLocalOperator, op_r: KLocalOperator) -> KLocalOperator:
    """In-place merge of op_r into op_l (KLocalOperator.jl:211-249 +
    OpConnection.jl:115-125)."""
    assert op_l.sites == op_r.sites
    op_l.mat = op_l.mat + op_r.mat
    for row_l, row_r in zip(op_l.conns, op_r.conns):
        for mel, ch in row_r:
            key = _changes_key(ch)
            for e in row_l:
                if _changes_key(e[1]) == key:
                    e[0] = e[0] + mel
                    break
            else:
                row_l.append([mel, ch])
    return op_l


def _kop_kop_prod(opl: KLocalOperator, opr: KLocalOperator):
    if opl.sites == opr.sites:
        return KLocalOperatorRow(opl.hilb, opl.sites, opl.mat @ opr.mat)
    if not set(opl.sites) & set(opr.sites):
        if len(opl.sites) == 1 and len(opr.sites) == 1:
            if opl.sites[0] > opr.sites[0]:
                opl, opr = opr, opl
            mat = np.kron(opr.mat, opl.mat)          # KLocalOperator.jl:305
            return KLocalOperatorRow(opl.hilb, [opl.sites[0], opr.sites[0]], mat)
        raise NotImplementedError(
            "tensor product of disjoint operators with more than one site each "
            "(not implemented in the reference either, KLocalOperator.jl:311)")
    # overlapping support, one operand single-site (KLocalOperator.jl:316-338)
    if len(opl.sites) == 1 or len(opr.sites) == 1:
        reversed_ = False
        if len(opl.sites) == 1:
            opl, opr = opr, opl
            reversed_ = True
        sr = opr.sites[0]
        r_index = opl.sites.index(sr)
        mats = [np.eye(2, dtype=np.complex128) for _ in opl.sites]
        mats[r_index] = opr.mat
        mat_r = mats[-1]
        for m in reversed(mats[:-1]):
            mat_r = np.kron(mat_r, m)
        prod = mat_r @ opl.mat if reversed_ else opl.mat @ mat_r
        return KLocalOperatorRow(opl.hilb, opl.sites, prod)
    raise NotImplementedError("product of overlapping multi-site operators "
                              "(not implemented in the reference either)")


class KLocalOperatorZero:
    """`LocalOperator(hilb)`: the empty operator (KLocalOperatorZero.jl:7-56)."""

    def __init__(self, hilb):
        self.hilb = hilb
        self.sites = []

    def duplicate(self):
        return KLocalOperatorZero(self.hilb)

    def __add__(self, other):
        if isinstance(other, KLocalOperator):
            return other.duplicate()
        if isinstance(other, KLocalOperatorZero):
            return self.duplicate()
        if isinstance(other, KLocalOperatorSum):
            return other + self
        return NotImplemented

    def __sub__(self, other):
        return -other

    def __rmul__(self, a):
        return self

    def __mul__(self, other):
        return self

    def adjoint(self):
        return self.duplicate()

    def transpose(self):
        return self.duplicate()

    def conj(self):
        return self.duplicate()


def LocalOperator(hilb):  # noqa: N802
    return KLocalOperatorZero(hilb)


class KLocalOperatorSum:
    """Sum of operators, insertion-ordered, same sites merged
    (KLocalOperatorSum.jl:7-111)."""

    def __init__(self, hilb, ops):
        self.hilb = hilb
        self.operators = list(ops)

    @property
    def sites(self):
        return [sites(o) for o in self.operators]

    def duplicate(self):
        return KLocalOperatorSum(self.hilb, [o.duplicate() for o in self.operators])

    def _add(self, op):
        s = sites(op)
        for i, o in enumerate(self.operators):
            if sites(o) == s:
                if isinstance(o, KLocalOperatorTensor):
                    self.operators[i] = _tensor_add_samesite(o, op)
                else:
                    _add_samesite(o, op)
                return self
        self.operators.append(op)
        return self

    def __add__(self, other):
        out = self.duplicate()
        if isinstance(other, KLocalOperatorSum):
            for op in other.operators:
                out._add(op.duplicate())
            return out
        return out._add(other.duplicate())

    __radd__ = __add__

    def __sub__(self, other):
        if isinstance(other, KLocalOperatorSum):
            out = self.duplicate()
            for op in other.operators:
                out._add(-op)
            return out
        return self.duplicate()._add(-other)

    def __neg__(self):
        return KLocalOperatorSum(self.hilb, [-o for o in self.operators])

    def __rmul__(self, a):
        if isinstance(a, (int, float, complex, np.number)):
            return _sum_list([a * op for op in self.operators])
        return NotImplemented

    def __mul__(self, other):
        if isinstance(other, (int, float, complex, np.number)):
            return _sum_list([other * op for op in self.operators])
        if isinstance(other, KLocalOperator):
            out = self.duplicate()
            out.operators = [op * other for op in out.operators]
            return out
        return NotImplemented

    def __truediv__(self, a):
        return _sum_list([op / a for op in self.operators])

    def transpose(self):
        return KLocalOperatorSum(self.hilb, [op.transpose() for op in self.operators])

    def conj(self):
        return KLocalOperatorSum(self.hilb, [op.conj() for op in self.operators])

    def adjoint(self):
        return self.transpose().conj()

    @property
    def H(self):  # noqa: N802
        return self.adjoint()


def _op_times_sum(opl, opr_sum):
    out = opr_sum.duplicate()
    out.operators = [opl * op for op in out.operators]
    return out


def _sum_list(ops):
    """Julia's `sum(vector_of_ops)`: a left fold with `+`; one element is
    returned as itself."""
    acc = ops[0]
    for op in ops[1:]:
        acc = acc + op
    return acc


class KLocalIdentity:
    def __init__(self, hilb):
        self.hilb = hilb


class KLocalOperatorTensor:
    """op_l (x) op_r on the doubled space; either side may be the identity
    (KLocalOperatorTensor.jl:7-43)."""

    def __init__(self, op_l, op_r):
        self.op_l = op_l
        self.op_r = op_r

    @property
    def sites(self):
        sl = [] if isinstance(self.op_l, KLocalIdentity) else list(self.op_l.sites)
        sr = [] if isinstance(self.op_r, KLocalIdentity) else list(self.op_r.sites)
        return (sl, sr)

    def duplicate(self):
        dl = self.op_l if isinstance(self.op_l, KLocalIdentity) else self.op_l.duplicate()
        dr = self.op_r if isinstance(self.op_r, KLocalIdentity) else self.op_r.duplicate()
        return KLocalOperatorTensor(dl, dr)

    def __add__(self, other):
        if isinstance(other, KLocalOperatorTensor):
            if self.sites == other.sites:
                return _tensor_add_samesite(self.duplicate(), other)
            return KLocalOperatorSum(None, [self]) + other
        if isinstance(other, KLocalOperatorSum):
            return other + self
        return NotImplemented

    def __rmul__(self, a):
        # KLocalOperatorTensor.jl:195-204
        if isinstance(self.op_l, KLocalIdentity):
            return KLocalOperatorTensor(self.op_l, a * self.op_r)
        if isinstance(self.op_r, KLocalIdentity):
            return KLocalOperatorTensor(a * self.op_l, self.op_r)
        s = np.sqrt(complex(a))
        return KLocalOperatorTensor(s * self.op_l, s * self.op_r)


def _tensor_add_samesite(t_l, t_r):
    """KLocalOperatorTensor.jl:160-169 merges the two sides independently."""
    if isinstance(t_l.op_l, KLocalIdentity) or isinstance(t_l.op_r, KLocalIdentity):
        raise NotImplementedError("same-site sum of tensors with an identity side "
                                  "has no method in the reference")
    _add_samesite(t_l.op_l, t_r.op_l.duplicate())
    _add_samesite(t_l.op_r, t_r.op_r.duplicate())
    return t_l


def _tensor(op_l, op_r):
    """KLocalOperatorTensor(op_l, op_r) with sums distributed
    (KLocalOperatorTensor.jl:35-43)."""
    if isinstance(op_l, KLocalOperatorSum) and isinstance(op_r, KLocalIdentity):
        return _sum_list([KLocalOperatorTensor(op, op_r) for op in op_l.operators])
    if isinstance(op_l, KLocalIdentity) and isinstance(op_r, KLocalOperatorSum):
        return _sum_list([KLocalOperatorTensor(op_l, op) for op in op_r.operators])
    return KLocalOperatorTensor(op_l, op_r)


class KLocalLiouvillian:
    """KLocalLiouvillian.jl:1-33."""

    def __init__(self, HnH, Lops):
        hilb = HnH.hilb
        self.phys = hilb
        self.hilb = SuperOpSpace(hilb)
        self.HnH = HnH
        self.HnH_l = (-1j) * _tensor(HnH, KLocalIdentity(hilb))
        self.HnH_r = (1j) * _tensor(KLocalIdentity(hilb), HnH.adjoint())
        ll = [KLocalOperatorTensor(L, L.conj()) for L in Lops]
        self.LLdag = _sum_list(ll) if ll else KLocalOperatorZero(self.hilb)


def liouvillian(H, Lops):
    """`liouvillian(H, ops)` (KLocalLiouvillian.jl:27-33)."""
    HnH = H.duplicate()
    for L in Lops:
        HnH = HnH + ((-0.5j) * L.adjoint()) * L
    return KLocalLiouvillian(HnH, list(Lops))


def operators(op):
    """Operators of a sum in enumeration order; zero operators (which a sum
    may hold, KLocalOperatorSum.jl:74-85) have no connections."""
    if isinstance(op, KLocalOperatorSum):
        return [o for o in op.operators if not isinstance(o, KLocalOperatorZero)]
    if isinstance(op, KLocalOperatorZero):
        return []
    return [op]


def sites(op):
    return op.sites


# -- single-site operators (src/Operators/SimpleOperators.jl:11-96) -----------
def _check_site(h, i):
    if not 1 <= i <= h.n_sites:
        raise IndexError("site index out of range")


def sigmax(h: HomogeneousSpin, i: int):
    _check_site(h, i)
    return KLocalOperatorRow(h, [i], [[0, 1], [1, 0]])


def sigmay(h: HomogeneousSpin, i: int):
    _check_site(h, i)
    return KLocalOperatorRow(h, [i], [[0, -1j], [1j, 0]])


def sigmaz(h: HomogeneousSpin, i: int):
    _check_site(h, i)
    return KLocalOperatorRow(h, [i], [[1, 0], [0, -1]])


def sigmam(h: HomogeneousSpin, i: int):
    _check_site(h, i)
    return KLocalOperatorRow(h, [i], [[0, 0], [1, 0]])


def sigmap(h: HomogeneousSpin, i: int):
    _check_site(h, i)
    return KLocalOperatorRow(h, [i], [[0, 1], [0, 0]])


# -- connection enumeration (host reference of the device enumeration) ------
def _tensor_terms(op):
    return operators(op)


def enumerate_connections(liouv: KLocalLiouvillian, s, e, skip_zero=True):
    """All (mel, row changes, column changes) of one doubled configuration in
    the reference's order.  `skip_zero` drops exact-zero mels as the
    accumulator does (AccumulatorObsGrad.jl:67)."""
    out = []
    for t in _tensor_terms(liouv.HnH_l):
        r = local_index(s, t.op_l.sites) - 1
        for mel, ch in t.op_l.conns[r]:
            out.append((mel, ch, ()))
    for t in _tensor_terms(liouv.HnH_r):
        r = local_index(e, t.op_r.sites) - 1
        for mel, ch in t.op_r.conns[r]:
            out.append((mel, (), ch))
    for t in _tensor_terms(liouv.LLdag):
        rl = local_index(s, t.op_l.sites) - 1
        rr = local_index(e, t.op_r.sites) - 1
        for ml, cl in t.op_l.conns[rl]:
            for mr, cr in t.op_r.conns[rr]:
                out.append((ml * mr, cl, cr))
    if skip_zero:
        out = [c for c in out if c[0] != 0]
    return out


def _apply(v, ch):
    v = v.copy()
    for site, val in ch:
        v[site - 1] = val
    return v


def to_matrix(op) -> np.ndarray:
    """Dense matrix of an operator or Liouvillian from its connections
    (src/Operators/OpConversion.jl:6-24): mat[i, j] += mel for the connection
    i -> j, 0-based indices."""
    if isinstance(op, KLocalLiouvillian):
        n = op.phys.n_sites
        d = 2 ** n
        M = np.zeros((d * d, d * d), dtype=np.complex128)
        for ie in range(d):
            e = state_from_int(op.phys, ie + 1)
            for is_ in range(d):
                s = state_from_int(op.phys, is_ + 1)
                i = ie * d + is_
                for mel, cl, cr in enumerate_connections(op, s, e, skip_zero=False):
                    s2 = _apply(s, cl)
                    e2 = _apply(e, cr)
                    j = _idx(s2) + d * _idx(e2)
                    M[i, j] += mel
        return M
    hilb = op.hilb
    n = hilb.n_sites
    d = 2 ** n
    M = np.zeros((d, d), dtype=np.complex128)
    for i in range(d):
        s = state_from_int(hilb, i + 1)
        for o in operators(op):
            r = local_index(s, o.sites) - 1
            for mel, ch in o.conns[r]:
                M[i, _idx(_apply(s, ch))] += mel
    return M


def _idx(v):
    return int(sum(((int(x) + 1) // 2) << k for k, x in enumerate(v)))


# -- flat table ---------------------------------------------------------------
TERM_ROW, TERM_COL, TERM_ROWCOL, TERM_KET = 0, 1, 2, 3


@dataclass
class FlatOpTable:
    """Flat connection table (`nqh_optable`, include/nqhip.h).

    Local operator o acts on op_sites[op_site_off[o]:op_site_off[o+1]]
    (0-based sites, local bit m <-> m-th listed site).  Its 2^k rows are
    rows op_row_off[o] .. op_row_off[o+1]-1; row q owns entries
    row_ent_off[q] .. row_ent_off[q+1]-1, each a complex mel and a mask of the
    physical sites it flips.  Terms are enumerated in order; term t has kind
    0 (row side), 1 (column side), 2 (row (x) column, nested loops) or 3 (ket)
    and local operators term_op_l[t] / term_op_r[t] (-1 = identity)."""
    n_sites: int
    term_kind: np.ndarray
    term_op_l: np.ndarray
    term_op_r: np.ndarray
    op_site_off: np.ndarray
    op_sites: np.ndarray
    op_row_off: np.ndarray
    row_ent_off: np.ndarray
    ent_mel: np.ndarray        # complex128
    ent_flip: np.ndarray       # uint64

    @property
    def n_terms(self):
        return len(self.term_kind)

    @property
    def n_ops(self):
        return len(self.op_site_off) - 1

    def max_connections(self) -> int:
        """Upper bound on connections per configuration (table capacity)."""
        tot = 0
        for k, ol, orr in zip(self.term_kind, self.term_op_l, self.term_op_r):
            wl = self._max_row(ol) if ol >= 0 else 1
            wr = self._max_row(orr) if orr >= 0 else 1
            tot += wl * wr
        return int(tot)

    def _max_row(self, o):
        r0, r1 = self.op_row_off[o], self.op_row_off[o + 1]
        return int(np.max(np.diff(self.row_ent_off[r0:r1 + 1])))


def _flat_add_op(op, state):
    key = id(op)
    if key in state["index"]:
        return state["index"][key]
    o = len(state["site_off"]) - 1
    state["index"][key] = o
    state["sites"].extend(s - 1 for s in op.sites)
    state["site_off"].append(len(state["sites"]))
    for row in op.conns:
        for mel, ch in row:
            mask = 0
            for site, _val in ch:
                mask |= 1 << (site - 1)
            state["mel"].append(complex(mel))
            state["flip"].append(mask)
        state["row_ent_off"].append(len(state["mel"]))
    state["row_off"].append(len(state["row_ent_off"]) - 1)
    return o


def flatten(op) -> FlatOpTable:
    """Flatten a KLocalLiouvillian (or a ket operator) into a FlatOpTable."""
    st = {"index": {}, "sites": [], "site_off": [0], "row_off": [0],
          "row_ent_off": [0], "mel": [], "flip": []}
    kinds, ls, rs = [], [], []
    if isinstance(op, KLocalLiouvillian):
        n = op.phys.n_sites
        for t in _tensor_terms(op.HnH_l):
            kinds.append(TERM_ROW); ls.append(_flat_add_op(t.op_l, st)); rs.append(-1)
        for t in _tensor_terms(op.HnH_r):
            kinds.append(TERM_COL); ls.append(-1); rs.append(_flat_add_op(t.op_r, st))
        for t in _tensor_terms(op.LLdag):
            kinds.append(TERM_ROWCOL)
            ls.append(_flat_add_op(t.op_l, st))
            rs.append(_flat_add_op(t.op_r, st))
    else:
        n = op.hilb.n_sites
        for o in operators(op):
            kinds.append(TERM_KET); ls.append(_flat_add_op(o, st)); rs.append(-1)
    return FlatOpTable(
        n_sites=n,
        term_kind=np.array(kinds, dtype=np.int32),
        term_op_l=np.array(ls, dtype=np.int32),
        term_op_r=np.array(rs, dtype=np.int32),
        op_site_off=np.array(st["site_off"], dtype=np.int32),
        op_sites=np.array(st["sites"], dtype=np.int32),
        op_row_off=np.array(st["row_off"], dtype=np.int32),
        row_ent_off=np.array(st["row_ent_off"], dtype=np.int32),
        ent_mel=np.array(st["mel"], dtype=np.complex128),
        ent_flip=np.array(st["flip"], dtype=np.uint64),
    )

