"""Kagome lattice geometry (host setup, runs once per run).

Mirrors ``DoubleKagome`` and its helpers from the reference:
src/Lattice.jl:1-97 and src/Hamiltonian.jl:28-153,279-283.  Sites and labels
are 1-based at this interface, exactly as in the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np


def validate_boundary_conditions(pbc, antipbc):
    """src/Lattice.jl:3-15: antiperiodic needs periodic in the same direction."""
    for d, (p, a) in enumerate(zip(pbc, antipbc)):
        if a and not p:
            which = "first" if d == 0 else "second"
            raise ValueError(
                f"Invalid boundary conditions in {which} direction: Cannot have antiperiodic "
                "boundary conditions without periodic boundary conditions")


@dataclass(frozen=True)
class DoubleKagome:
    """6-site double unit cell Kagome lattice, ``n1*n2*3`` sites (src/Lattice.jl:17-92).

    ``n1`` counts single cells along a1 and must be even; there are
    ``(n1//2)*n2`` double cells.  a1 = [4t, 0], a2 = [t, sqrt(3) t]."""

    n1: int
    n2: int
    t: float = 1.0
    PBC: tuple = (False, False)
    antiPBC: tuple = (False, False)
    a1: np.ndarray = field(init=False, repr=False, compare=False)
    a2: np.ndarray = field(init=False, repr=False, compare=False)
    r: np.ndarray = field(init=False, repr=False, compare=False)

    def __post_init__(self):
        validate_boundary_conditions(self.PBC, self.antiPBC)
        if self.n1 % 2 != 0:
            raise AssertionError("n1 must be even in DoubleKagome")
        a = 2.0 * self.t
        a1 = np.array([2.0 * a, 0.0])
        a2 = np.array([0.5 * a, 0.5 * math.sqrt(3.0) * a])
        r = np.array([[0.0, 0.0], 0.25 * a1, 0.5 * a2, 0.5 * a1, 0.75 * a1,
                      [2.5 * self.t, 0.5 * math.sqrt(3.0) * self.t]])
        object.__setattr__(self, "a1", a1)
        object.__setattr__(self, "a2", a2)
        object.__setattr__(self, "r", r)

    @classmethod
    def build(cls, t: float, n1: int, n2: int, PBC, antiPBC=(False, False)):
        """Same argument order as the reference constructor DoubleKagome(t, n1, n2, PBC; antiPBC)."""
        return cls(int(n1), int(n2), float(t), tuple(bool(x) for x in PBC),
                   tuple(bool(x) for x in antiPBC))

    @property
    def ncells(self) -> int:
        return (self.n1 // 2) * self.n2

    def cell_grid(self):
        """Integer (x, y) grid position of every double cell, cell index = x + y*(n1/2)."""
        c = np.arange(self.ncells)
        return c % (self.n1 // 2), c // (self.n1 // 2)


def ns(lat: DoubleKagome) -> int:
    """Number of sites, src/Lattice.jl:97."""
    return lat.n1 * lat.n2 * 3


def unitcell_coord(lat: DoubleKagome, s: int) -> np.ndarray:
    """Origin of the double cell holding 1-based site ``s`` (src/Hamiltonian.jl:28-38)."""
    n = ns(lat)
    if not 1 <= s <= n:
        raise AssertionError(f"s should be in the range of 1 to ns, got: {s}")
    cell = (s - 1) // 6
    h = lat.n1 // 2
    return (cell % h) * lat.a1 + (cell // h) * lat.a2


def unitcell_diff(lat: DoubleKagome, c1, c2):
    """(dx, dy) with c1 - c2 = dx*a1 + dy*a2, rounded (src/Hamiltonian.jl:56-76)."""
    d = np.asarray(c1, dtype=float) - np.asarray(c2, dtype=float)
    m = np.array([[lat.a1[0], lat.a2[0]], [lat.a1[1], lat.a2[1]]])
    x = np.linalg.solve(m, d)
    return int(np.rint(x[0])), int(np.rint(x[1]))


def get_site_coord(lat: DoubleKagome, s: int) -> np.ndarray:
    """Real-space position of 1-based site ``s`` (src/Hamiltonian.jl:279-283)."""
    return unitcell_coord(lat, s) + lat.r[(s - 1) % 6]


def site_coords(lat: DoubleKagome) -> np.ndarray:
    """(ns, 2) positions of all sites, vectorised ``get_site_coord``."""
    x, y = lat.cell_grid()
    origin = x[:, None] * lat.a1[None, :] + y[:, None] * lat.a2[None, :]
    return (origin[:, None, :] + lat.r[None, :, :]).reshape(-1, 2)


def image_shifts(lat: DoubleKagome):
    """Periodic images tried for an inter-cell pair, in the reference's order
    (src/Hamiltonian.jl:119-139): the direct displacement first, then every
    (shift1, shift2) != (0, 0) with a sign -1 per antiperiodic crossing."""
    h, n2 = lat.n1 // 2, lat.n2
    if not lat.PBC[0] and not lat.PBC[1]:
        return [(0, 0, 1.0)]
    xs = [-h, 0, h] if lat.PBC[0] else [0]
    ys = [-n2, 0, n2] if lat.PBC[1] else [0]
    out = [(0, 0, 1.0)]
    for s1 in xs:
        for s2 in ys:
            if s1 == 0 and s2 == 0:
                continue
            sign = 1.0
            if lat.antiPBC[0] and s1 != 0:
                sign = -sign
            if lat.antiPBC[1] and s2 != 0:
                sign = -sign
            out.append((s1, s2, sign))
    return out


def get_boundary_shifts(lat: DoubleKagome, s1: int, s2: int):
    """All (dx, dy, sign) displacements from s1's cell to s2's cell under the
    boundary conditions (src/Hamiltonian.jl:99-153)."""
    if s1 == s2:
        raise AssertionError(f"s1 and s2 should not be the same, got: {s1} and {s2}")
    n = ns(lat)
    for s in (s1, s2):
        if not 1 <= s <= n:
            raise AssertionError(f"site out of range: {s} in {n}")
    dx, dy = unitcell_diff(lat, unitcell_coord(lat, s2), unitcell_coord(lat, s1))
    return [(dx + a, dy + b, sg) for (a, b, sg) in image_shifts(lat)]
