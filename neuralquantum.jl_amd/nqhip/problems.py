"""Lattices and the dissipative spin models used by the steady-state loop.

The construction order mirrors how the reference's drivers build them, since
operator insertion order fixes the connection order:

* dissipative Ising chain: per site i, `H += g/2 sigmax(i)` then
  `H += V/4 sigmaz(i) sigmaz(i+1)` (examples/dissipative_ising1d.jl:17-26);
  jump operators sqrt(gamma) sigmam(i) in site order.
* lattice Hamiltonians iterate `edges(lattice)` (SimpleOperators.jl:163-170);
  edges are the sorted pairs (i < j) in lexicographic order, vertices numbered
  x fastest (LightGraphs `SimpleGraph` edge order on a `HyperCube`).
"""
from __future__ import annotations

import math

from .hilbert import HomogeneousSpin
from .operators import LocalOperator, liouvillian, sigmam, sigmax, sigmay, sigmaz

__all__ = ["chain_edges", "square_edges", "dissipative_ising_chain",
           "dissipative_xyz", "ising_hamiltonian"]


def chain_edges(n: int, periodic: bool):
    """Bonds (i, i+1) of a 1D chain, 1-based; periodic adds (1, n) last."""
    edges = [(i, i + 1) for i in range(1, n)]
    if periodic and n > 2:
        edges.append((1, n))
    return sorted(edges)


def square_edges(lx: int, ly: int, periodic: bool = True):
    """Nearest-neighbour bonds of an lx x ly square lattice, vertex
    v = 1 + x + lx*y, as sorted (i < j) pairs in lexicographic order."""
    es = set()
    for y in range(ly):
        for x in range(lx):
            v = 1 + x + lx * y
            for dx, dy in ((1, 0), (0, 1)):
                nx, ny = x + dx, y + dy
                if not periodic and (nx >= lx or ny >= ly):
                    continue
                nx %= lx
                ny %= ly
                w = 1 + nx + lx * ny
                if w != v:
                    es.add((min(v, w), max(v, w)))
    return sorted(es)


def dissipative_ising_chain(n: int, V: float = 2.0, g: float = 1.0, gamma: float = 1.0,
                            periodic: bool = True):
    """H = V/4 sum sz_i sz_{i+1} + g/2 sum sx_i, jumps sqrt(gamma) sm_i.
    Returns (hilb, H, jump_ops, liouvillian)."""
    hilb = HomogeneousSpin(n)
    H = LocalOperator(hilb)
    for i in range(1, n + 1):
        H = H + (g / 2.0) * sigmax(hilb, i)
        if i < n:
            H = H + ((V / 4.0) * sigmaz(hilb, i)) * sigmaz(hilb, i + 1)
        elif periodic and n >= 2:
            # bond (N, mod(N, N) + 1) = (N, 1), as the reference's example writes it
            H = H + ((V / 4.0) * sigmaz(hilb, i)) * sigmaz(hilb, 1)
    ops = [_scaled_jump(gamma, sigmam(hilb, i)) for i in range(1, n + 1)]
    return hilb, H, ops, liouvillian(H, ops)


def ising_hamiltonian(n: int, g: float, V: float, periodic: bool = True):
    """`quantum_ising_hamiltonian` on a chain (SimpleOperators.jl:160-173):
    all field terms first, then the bonds in edge order."""
    hilb = HomogeneousSpin(n)
    H = LocalOperator(hilb)
    for i in range(1, n + 1):
        H = H + (g / 2.0) * sigmax(hilb, i)
    for (i, j) in chain_edges(n, periodic):
        H = H + ((V / 4.0) * sigmaz(hilb, i)) * sigmaz(hilb, j)
    return hilb, H


def dissipative_xyz(lx: int, ly: int, Jx: float = 0.9, Jy: float = 1.0, Jz: float = 1.1,
                    gamma: float = 1.0, periodic: bool = True):
    """H = sum_<ij> (Jx sx sx + Jy sy sy + Jz sz sz), jumps sqrt(gamma) sm_i."""
    n = lx * ly
    hilb = HomogeneousSpin(n)
    H = LocalOperator(hilb)
    for (i, j) in square_edges(lx, ly, periodic):
        H = H + (Jx * sigmax(hilb, i)) * sigmax(hilb, j)
        H = H + (Jy * sigmay(hilb, i)) * sigmay(hilb, j)
        H = H + (Jz * sigmaz(hilb, i)) * sigmaz(hilb, j)
    ops = [_scaled_jump(gamma, sigmam(hilb, i)) for i in range(1, n + 1)]
    return hilb, H, ops, liouvillian(H, ops)


def _scaled_jump(gamma: float, op):
    return op if gamma == 1.0 else math.sqrt(gamma) * op
