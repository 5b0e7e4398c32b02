"""Spin-1/2 Hilbert spaces and configuration indexing.

Mirrors the reference's `HomogeneousSpin` / `SuperOpSpace` conventions:

* a site value is v in {-1, +1}, stored as float; the default fill is -1
  (src/Hilbert/HomogeneousSpin.jl:69-70);
* local index of a site: li = (v+1)/2 + 1 (HomogeneousSpin.jl:164-169);
* physical index: toint(s) = 1 + sum_i ((v_i+1)/2) * 2^(i-1), site 1 is the
  least significant bit and bit 1 <=> v = +1 (HomogeneousSpin.jl:156-162);
* doubled index: toint((s, e)) = (toint(e)-1) * 2^N + toint(s)
  (src/Hilbert/SuperHilbert.jl:50-56, inverse :23-36);
* doubled site i in 1..2N flips s_i for i <= N, e_(i-N) otherwise
  (src/Hilbert/DoubledHilbert.jl:19-27).

On the device a doubled configuration is one packed key,
key = bits(s) | bits(e) << N, so toint((s, e)) == key + 1.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "HomogeneousSpin", "SuperOpSpace", "physical", "nsites", "spacedimension",
    "state_from_int", "toint", "local_index", "key_from_states",
    "states_from_key", "keys_to_states", "states_to_keys", "flipat",
]


class HomogeneousSpin:
    """N identical spins-1/2 (HomogeneousSpin.jl:16-49, S = 1//2 only)."""

    def __init__(self, n_sites: int):
        if n_sites < 1 or n_sites > 32:
            raise ValueError("HomogeneousSpin: 1 <= N <= 32 sites supported")
        self.n_sites = int(n_sites)

    @property
    def shape(self):
        return [2] * self.n_sites

    def __eq__(self, other):
        return isinstance(other, HomogeneousSpin) and other.n_sites == self.n_sites

    def __hash__(self):
        return hash(("HomogeneousSpin", self.n_sites))

    def __repr__(self):
        return f"HomogeneousSpin({self.n_sites}, 2)"


class SuperOpSpace:
    """Doubled (row, column) space of a physical space (SuperHilbert.jl:11)."""

    def __init__(self, basis: HomogeneousSpin):
        self.basis = basis

    def __eq__(self, other):
        return isinstance(other, SuperOpSpace) and other.basis == self.basis

    def __hash__(self):
        return hash(("SuperOpSpace", self.basis))

    def __repr__(self):
        return f"SuperOpSpace({self.basis})"


def physical(h):
    return h.basis if isinstance(h, SuperOpSpace) else h


def nsites(h) -> int:
    if isinstance(h, SuperOpSpace):
        return 2 * h.basis.n_sites
    return h.n_sites


def spacedimension(h) -> int:
    if isinstance(h, SuperOpSpace):
        return (2 ** h.basis.n_sites) ** 2
    return 2 ** h.n_sites


def state_from_int(h: HomogeneousSpin, val: int) -> np.ndarray:
    """`set!(s, h, val)` for a 1-based index val (HomogeneousSpin.jl:116-125)."""
    if not 1 <= val <= 2 ** h.n_sites:
        raise IndexError("index outside the Hilbert space")
    v = val - 1
    out = np.empty(h.n_sites)
    for i in range(h.n_sites):
        v, tmp = divmod(v, 2)
        out[i] = tmp * 2 - 1
    return out


def toint(s, h: HomogeneousSpin) -> int:
    """1-based physical index of a +-1 vector (HomogeneousSpin.jl:156-162)."""
    tot = 0
    for i, v in enumerate(s):
        tot += int((v + 1) // 2) * (2 ** i)
    return tot + 1


def local_index(s, sites) -> int:
    """1-based row index of a local operator acting on `sites` (1-based)
    (HomogeneousSpin.jl:171-179)."""
    idx = 1
    for m, j in enumerate(sites):
        idx += int((s[j - 1] + 1) // 2) * (2 ** m)
    return idx


def key_from_states(s, e) -> int:
    """Packed doubled key bits(s) | bits(e) << N; equals toint((s,e)) - 1."""
    n = len(s)
    k = 0
    for i in range(n):
        if s[i] > 0:
            k |= 1 << i
        if e[i] > 0:
            k |= 1 << (i + n)
    return k


def states_from_key(key: int, n: int):
    s = np.array([1.0 if (key >> i) & 1 else -1.0 for i in range(n)])
    e = np.array([1.0 if (key >> (i + n)) & 1 else -1.0 for i in range(n)])
    return s, e


def keys_to_states(keys, n: int):
    """Vectorised unpack: returns (s, e) as N x B float64 column-major batches,
    the reference's batch layout (src/States/States.jl:16-19)."""
    keys = np.asarray(keys, dtype=np.uint64)
    bits = (keys[None, :] >> np.arange(2 * n, dtype=np.uint64)[:, None]) & np.uint64(1)
    v = bits.astype(np.float64) * 2.0 - 1.0
    return np.ascontiguousarray(v[:n]), np.ascontiguousarray(v[n:])


def states_to_keys(s, e):
    s = np.asarray(s)
    e = np.asarray(e)
    n = s.shape[0]
    w = (np.uint64(1) << np.arange(n, dtype=np.uint64))
    ks = ((s > 0).astype(np.uint64) * w[:, None]).sum(axis=0)
    ke = ((e > 0).astype(np.uint64) * w[:, None]).sum(axis=0)
    return (ks | (ke << np.uint64(n))).astype(np.uint64)


def flipat(key: int, i: int, n: int) -> int:
    """Flip doubled site i (1-based, DoubledHilbert.jl:19-27) of a packed key."""
    if not 1 <= i <= 2 * n:
        raise IndexError("doubled site out of range")
    return key ^ (1 << (i - 1))
