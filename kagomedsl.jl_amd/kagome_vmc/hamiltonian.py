"""Spinon mean-field ansatz and J1 bond list (host setup, once per ansatz).

Mirrors src/Hamiltonian.jl:209-492 of the reference: the pi-flux hopping
tables, ``Hmat`` (hopping matrix with Peierls phases and twisted boundary
conditions), ``orbitals`` (lowest-N eigenvectors), ``get_nn`` (bond list in the
reference's column-major order) and the ``Hamiltonian`` bundle handed to the
device sampler.  The site-pair loop of the reference is evaluated here as one
vectorised pass over all pairs of different cells.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass

import numpy as np

from .lattice import DoubleKagome, image_shifts, ns, site_coords

# src/Hamiltonian.jl:219-235
pi_link_in = {
    (1, 2): 1, (1, 3): 1, (2, 3): 1, (2, 4): -1, (4, 6): 1, (4, 5): 1, (5, 6): 1,
    (2, 1): 1, (3, 1): 1, (3, 2): 1, (4, 2): -1, (6, 4): 1, (5, 4): 1, (6, 5): 1,
}
# src/Hamiltonian.jl:250-261: (label1, label2, dx, dy) -> amplitude
pi_link_inter = {
    (3, 5, -1, 1): -1, (3, 1, 0, 1): -1, (6, 2, 0, 1): -1, (6, 4, 0, 1): 1,
    (5, 1, 1, 0): 1, (1, 5, -1, 0): 1, (1, 3, 0, -1): -1, (2, 6, 0, -1): -1,
    (4, 6, 0, -1): 1, (5, 3, 1, -1): -1,
}


def link_orientation(lat: DoubleKagome, l1: int, l2: int, dx: int = 0, dy: int = 0) -> int:
    """+1 when the hop from label l1 (cell 0) to label l2 in the cell
    dx a1 + dy a2 runs counterclockwise around the triangle that holds the
    bond, -1 otherwise.  Every kagome bond belongs to exactly one triangle;
    its third vertex is the site at distance t from both ends."""
    r1 = lat.r[l1 - 1]
    r2 = dx * lat.a1 + dy * lat.a2 + lat.r[l2 - 1]
    for l in range(6):
        for cx in range(-2, 3):
            for cy in range(-2, 3):
                rk = cx * lat.a1 + cy * lat.a2 + lat.r[l]
                if (abs(np.linalg.norm(rk - r1) - lat.t) < 1e-9 and
                        abs(np.linalg.norm(rk - r2) - lat.t) < 1e-9):
                    d1, d2 = r2 - r1, rk - r1
                    return 1 if d1[0] * d2[1] - d1[1] * d2[0] > 0 else -1
    raise ValueError(f"link ({l1}, {l2}, {dx}, {dy}) is not a nearest-neighbour bond")


def flux_link_tables(lat: DoubleKagome, phi: float, link_in=None, link_inter=None, dphi: bool = False):
    """Link tables of the flux family: each amplitude of the pi-flux tables
    (src/Hamiltonian.jl:219-261) times exp(i phi o / 3), o the link's
    orientation in its triangle.  Every triangle then carries flux phi and
    every hexagon pi - 2 phi (each hexagon edge is traversed against its
    triangle), and phi = 0 is the pi-flux ansatz.  Complex amplitudes enter
    Hmat exactly as in the reference, whose tables hold numbers multiplying a
    ComplexF64 matrix (custom tables: scripts/zero_flux.jl:15-42).  With
    dphi, the tables of the derivative d/dphi (Hmat is linear in them)."""
    link_in = pi_link_in if link_in is None else link_in
    link_inter = pi_link_inter if link_inter is None else link_inter

    def amp(v, o):
        a = v * np.exp(1j * phi * o / 3.0)
        return a * (1j * o / 3.0) if dphi else a

    tin = {k: amp(v, link_orientation(lat, k[0], k[1])) for k, v in link_in.items()}
    tinter = {k: amp(v, link_orientation(lat, *k)) for k, v in link_inter.items()}
    return tin, tinter


def _peierls(B, r1, r2):
    """exp(i B/2 (x1 + x2)(y2 - y1)) (src/Hamiltonian.jl:201-202,325-326)."""
    return np.exp(1j * (B / 2.0) * (r1[..., 0] + r2[..., 0]) * (r2[..., 1] - r1[..., 1]))


def tunneling(lat: DoubleKagome, link_in=None, link_inter=None, B: float = 0.0) -> np.ndarray:
    """Upper-triangular hopping matrix T of ``Hmat`` before H = -(T + T')."""
    link_in = pi_link_in if link_in is None else link_in
    link_inter = pi_link_inter if link_inter is None else link_inter
    n = ns(lat)
    T = np.zeros((n, n), dtype=np.complex128)
    pos = site_coords(lat)
    cx, cy = lat.cell_grid()
    # in-cell: s1 < s2 in the same cell, amplitude assigned (src/Hamiltonian.jl:316-330)
    for c in range(lat.ncells):
        for a in range(6):
            for b in range(a + 1, 6):
                amp = link_in.get((a + 1, b + 1))
                if amp is None:
                    continue
                s1, s2 = 6 * c + a, 6 * c + b
                T[s1, s2] = amp * _peierls(B, pos[s1], pos[s2])
    if not link_inter:
        return T
    # inter-cell: every pair s1 < s2 in different cells, each periodic image
    # whose (label1, label2, dx, dy) has an amplitude (src/Hamiltonian.jl:333-343, 177-207)
    site = np.arange(n)
    cell, label = site // 6, site % 6
    I, J = np.triu_indices(n, 1)
    keep = cell[I] != cell[J]
    I, J = I[keep], J[keep]
    dx0 = cx[cell[J]] - cx[cell[I]]
    dy0 = cy[cell[J]] - cy[cell[I]]
    keys = np.array(list(link_inter.keys()), dtype=np.int64)
    vals = np.array(list(link_inter.values()), dtype=np.complex128)
    span = int(np.max(np.abs(keys[:, 2:]))) if len(keys) else 0
    off = 2 * max(lat.n1 // 2, lat.n2) + span + 1
    table_idx = -np.ones((6, 6, 2 * off + 1, 2 * off + 1), dtype=np.int64)
    for q, (l1, l2, dx, dy) in enumerate(keys):
        table_idx[l1 - 1, l2 - 1, dx + off, dy + off] = q
    rel = pos[J] - (cx[cell[J]][:, None] * lat.a1 + cy[cell[J]][:, None] * lat.a2)
    org1 = cx[cell[I]][:, None] * lat.a1 + cy[cell[I]][:, None] * lat.a2
    for (sh1, sh2, sign) in image_shifts(lat):
        dx, dy = dx0 + sh1, dy0 + sh2
        ok = (np.abs(dx) <= off) & (np.abs(dy) <= off)
        q = np.full(len(I), -1)
        q[ok] = table_idx[label[I][ok], label[J][ok], dx[ok] + off, dy[ok] + off]
        hit = q >= 0
        if not np.any(hit):
            continue
        r2 = org1[hit] + dx[hit][:, None] * lat.a1 + dy[hit][:, None] * lat.a2 + rel[hit]
        amp = sign * vals[q[hit]] * _peierls(B, pos[I[hit]], r2)
        np.add.at(T, (I[hit], J[hit]), amp)
    return T


def Hmat(lat: DoubleKagome, link_in=None, link_inter=None, B: float = 0.0) -> np.ndarray:
    """Single-spinon Hamiltonian H = -(T + T') (src/Hamiltonian.jl:308-354)."""
    T = tunneling(lat, link_in, link_inter, B)
    if np.any(np.tril(T, -1) != 0):
        raise RuntimeError("tunneling matrix must be upper triangular")
    return -(T + T.conj().T)


def get_nn(H: np.ndarray):
    """Bond list from the nonzeros of triu(H), column-major order: sorted by j,
    then i, with i <= j, 1-based (src/Hamiltonian.jl:447-451)."""
    jj, ii = np.nonzero(np.triu(H).T)
    return [(int(i) + 1, int(j) + 1) for i, j in zip(ii, jj)]


def orbitals(H: np.ndarray, N_up: int, N_down: int):
    """Lowest N_up / N_down eigenvectors of H (src/Hamiltonian.jl:382-393).

    Real symmetric H yields real orbitals.  Also returns the sorted spectrum."""
    Hh = H.real if np.all(H.imag == 0) else H
    w, v = np.linalg.eigh(Hh)
    p = np.argsort(w, kind="stable")
    w, v = w[p], v[:, p]
    return np.asfortranarray(v[:, :N_up]), np.asfortranarray(v[:, :N_down]), w


def shell_gap(spectrum: np.ndarray, n: int) -> float:
    """eps_{n+1} - eps_n: zero means the Fermi level is partly filled (open shell)."""
    if n <= 0 or n >= len(spectrum):
        return np.inf
    return float(spectrum[n] - spectrum[n - 1])


@dataclass
class Hamiltonian:
    """src/Hamiltonian.jl:420-427: filling, occupied orbitals, H and the J1 bonds."""

    N_up: int
    N_down: int
    U_up: np.ndarray
    U_down: np.ndarray
    H_mat: np.ndarray
    nn: list
    spectrum: np.ndarray = None

    @classmethod
    def build(cls, N_up: int, N_down: int, lat: DoubleKagome, link_in=None, link_inter=None,
              B: float = 0.0, nn=None):
        """src/Hamiltonian.jl:480-492.  ``nn`` overrides the bond list (the
        J1 list must come from the m = 0, t2 = 0 geometry when H carries
        on-site or second-neighbour terms)."""
        H = Hmat(lat, link_in, link_inter, B)
        U_up, U_down, w = orbitals(H, N_up, N_down)
        for n in {N_up, N_down}:
            if shell_gap(w, n) < 1e-8:
                warnings.warn(f"open shell at filling {n}: the Slater determinant depends on the "
                              "eigensolver's basis choice", RuntimeWarning)
        return cls(N_up, N_down, U_up, U_down, H, get_nn(H) if nn is None else list(nn), w)

    @property
    def ns(self) -> int:
        return self.U_up.shape[0]

    def bonds_array(self) -> np.ndarray:
        return np.ascontiguousarray(np.array(self.nn, dtype=np.int32).reshape(-1, 2))
