"""Host setup of the product (kagome_vmc.lattice / .hamiltonian): the
reference's known answers and agreement with the oracle's independent
restatement of Hmat / get_nn."""
import numpy as np
import pytest

from kagome_vmc import (DoubleKagome, Hamiltonian, Hmat, get_boundary_shifts, get_nn,
                        get_site_coord, ns, orbitals, shell_gap, unitcell_coord)


def test_lattice_known_answers():
    with pytest.raises(AssertionError):
        DoubleKagome.build(1.0, 3, 3, (False, False))
    assert ns(DoubleKagome.build(1.0, 4, 3, (False, False))) == 36
    with pytest.raises(ValueError):
        DoubleKagome.build(1.0, 4, 3, (False, False), (True, False))
    lat = DoubleKagome.build(1.0, 4, 3, (False, False))
    assert np.allclose(get_site_coord(lat, 3), [0.5, 0.5 * np.sqrt(3)])
    assert np.allclose(get_site_coord(lat, 13), [1.0, np.sqrt(3)])
    assert np.allclose(unitcell_coord(DoubleKagome.build(1.0, 4, 2, (False, False)), 24),
                       [5.0, np.sqrt(3)])
    lp = DoubleKagome.build(1.0, 4, 3, (True, True))
    sh = get_boundary_shifts(lp, 3, 9)
    assert (1, 0, 1.0) in sh
    with pytest.raises(AssertionError):
        get_boundary_shifts(lp, 3, 3)


def test_hmat_known_answers():
    H = Hmat(DoubleKagome.build(1.0, 4, 3, (False, False)))
    assert H[0, 1] == -1 and H[0, 2] == -1 and np.isclose(H[2, 12], 1)
    assert np.allclose(H[0, 3:], 0)
    H = Hmat(DoubleKagome.build(1.0, 4, 3, (True, True), (True, True)))
    assert H[0, 10] == 1 and H[0, 26] == -1 and H[10, 26] == 1
    H = Hmat(DoubleKagome.build(1.0, 4, 3, (True, True), (False, True)))
    assert H[0, 10] == -1 and H[0, 26] == -1


@pytest.mark.parametrize("n1,n2", [(2, 2), (4, 3), (4, 4), (6, 6)])
@pytest.mark.parametrize("pbc,apbc", [((False, False), (False, False)),
                                      ((True, False), (True, False)),
                                      ((True, True), (False, True)),
                                      ((True, True), (True, True))])
@pytest.mark.parametrize("B", [0.0, 0.1])
def test_hmat_matches_oracle(oracle_mod, n1, n2, pbc, apbc, B):
    Hp = Hmat(DoubleKagome.build(1.0, n1, n2, pbc, apbc), B=B)
    Ho = oracle_mod.OracleLattice(1.0, n1, n2, pbc, apbc).hmat(B=B)
    assert np.abs(Hp - Ho).max() < 1e-13
    assert get_nn(Hp) == oracle_mod.get_nn(Ho)


@pytest.mark.parametrize("L,apbc,closed", [(4, (False, True), True), (4, (False, False), False),
                                           (6, (False, False), True), (8, (False, True), True),
                                           (12, (False, True), True), (14, (False, False), True),
                                           (14, (False, True), False)])
def test_shell_closure(L, apbc, closed):
    """Closed shells (read off the spectrum) decide the boundary conditions:
    antiperiodic along a2 for L = 4, 8, 12, plain periodic for L = 6, 14."""
    H = Hmat(DoubleKagome.build(1.0, L, L, (True, True), apbc))
    _, _, w = orbitals(H, 3 * L * L // 2, 3 * L * L // 2)
    assert (shell_gap(w, 3 * L * L // 2) > 1e-6) == closed


def test_pi_flux_pattern():
    """Flux 0 through every triangle and pi through every hexagon: the product
    of hoppings around each 3-cycle is positive (t = -H) and around each
    elementary 6-cycle negative, for the reference's pi-flux tables."""
    H = Hmat(DoubleKagome.build(1.0, 8, 8, (True, True), (False, False)))
    T = -H.real
    n = T.shape[0]
    nbrs = [set(np.nonzero(T[i])[0]) for i in range(n)]
    tri = set()
    for i in range(n):
        for j in nbrs[i]:
            for k in nbrs[j] & nbrs[i]:
                tri.add(tuple(sorted((i, j, k))))
    assert len(tri) == 2 * n // 3
    for (i, j, k) in tri:
        assert T[i, j] * T[j, k] * T[k, i] > 0
    # hexagons: 6-cycles with no chord, found by walking neighbours
    hexes = set()
    for a in range(n):
        for b in nbrs[a]:
            for c in nbrs[b] - {a}:
                if c in nbrs[a]:
                    continue
                for d in nbrs[c] - {a, b}:
                    if d in nbrs[b]:
                        continue
                    for e in nbrs[d] - {a, b, c}:
                        if e in nbrs[c]:
                            continue
                        for f in (nbrs[e] & nbrs[a]) - {b, c, d}:
                            if f in nbrs[d] or f in nbrs[b]:
                                continue
                            cyc = (a, b, c, d, e, f)
                            hexes.add(frozenset(cyc))
                            prod = T[a, b] * T[b, c] * T[c, d] * T[d, e] * T[e, f] * T[f, a]
                            assert prod < 0
    assert len(hexes) == n // 3


def test_hamiltonian_bundle():
    lat = DoubleKagome.build(1.0, 4, 3, (True, False))
    ham = Hamiltonian.build(18, 18, lat)
    assert ham.U_up.shape == (36, 18) and ham.U_down.shape == (36, 18)
    assert np.allclose(ham.U_up.T @ ham.U_up, np.eye(18), atol=1e-12)
    assert np.allclose(ham.H_mat @ ham.U_up, ham.U_up * ham.spectrum[:18], atol=1e-12)
    assert ham.bonds_array().shape == (len(ham.nn), 2)


def _loops(H):
    """Triangles (3-cycles) and hexagons (chordless 6-cycles) of the bond
    graph of H."""
    n = H.shape[0]
    nbrs = [set(np.nonzero(H[i])[0]) - {i} for i in range(n)]
    tri = set()
    for i in range(n):
        for j in nbrs[i]:
            for k in nbrs[j] & nbrs[i]:
                tri.add(tuple(sorted((i, j, k))))
    hexes = {}
    for a in range(n):
        for b in nbrs[a]:
            for c in nbrs[b] - {a}:
                if c in nbrs[a]:
                    continue
                for d in nbrs[c] - {a, b}:
                    if d in nbrs[b]:
                        continue
                    for e in nbrs[d] - {a, b, c}:
                        if e in nbrs[c]:
                            continue
                        for f in (nbrs[e] & nbrs[a]) - {b, c, d}:
                            if f in nbrs[d] or f in nbrs[b]:
                                continue
                            hexes.setdefault(frozenset((a, b, c, d, e, f)), (a, b, c, d, e, f))
    return [list(t) for t in tri], list(hexes.values())


def _ccw(lat, pos, cyc):
    """The cycle in counterclockwise order (minimal-image steps on the torus)."""
    T1, T2 = (lat.n1 // 2) * lat.a1, lat.n2 * lat.a2
    p = np.zeros(2)
    area = 0.0
    for q in range(len(cyc)):
        d = pos[cyc[(q + 1) % len(cyc)]] - pos[cyc[q]]
        best = min((d + i * T1 + j * T2 for i in (-1, 0, 1) for j in (-1, 0, 1)), key=np.linalg.norm)
        area += p[0] * best[1] - p[1] * best[0]
        p = p + best
    return cyc if area > 0 else cyc[::-1]


@pytest.mark.parametrize("phi", [np.pi / 3, 0.9, 0.0])
def test_flux_family_pattern(phi):
    """flux_link_tables(phi): the phase of the product of t_xy = -H[x, y]
    along every counterclockwise triangle is phi and along every
    counterclockwise hexagon pi - 2 phi (mod 2 pi); |t| = 1 on every bond;
    phi = 0 gives the reference's pi-flux H exactly; the derivative tables
    equal a central difference of H in phi."""
    from kagome_vmc.hamiltonian import flux_link_tables
    from kagome_vmc.lattice import site_coords
    lat = DoubleKagome.build(1.0, 8, 8, (True, True), (False, False))
    H = Hmat(lat, *flux_link_tables(lat, phi))
    assert np.abs(H - H.conj().T).max() < 1e-15
    nz = H[H != 0]
    assert np.allclose(np.abs(nz), 1.0)
    if phi == 0.0:
        assert np.array_equal(H, Hmat(lat))
    pos = site_coords(lat)
    tri, hexes = _loops(H)
    assert len(tri) == 2 * H.shape[0] // 3 and len(hexes) == H.shape[0] // 3

    def phase(cyc):
        c = _ccw(lat, pos, cyc)
        prod = 1.0 + 0j
        for q in range(len(c)):
            prod *= -H[c[q], c[(q + 1) % len(c)]]
        return np.angle(prod)

    wrap = lambda x: (x + np.pi) % (2 * np.pi) - np.pi
    for t in tri:
        assert abs(wrap(phase(t) - phi)) < 1e-12
    for h in hexes:
        assert abs(wrap(phase(h) - (np.pi - 2 * phi))) < 1e-12
    eps = 1e-6
    dH = Hmat(lat, *flux_link_tables(lat, phi, dphi=True))
    fd = (Hmat(lat, *flux_link_tables(lat, phi + eps)) - Hmat(lat, *flux_link_tables(lat, phi - eps))) / (2 * eps)
    assert np.abs(dH - fd).max() < 1e-9


def test_flux_family_matches_oracle(oracle_mod):
    """The complex flux tables through the oracle's Hmat (complex link
    amplitudes, plus a field B) give the product's H."""
    from kagome_vmc.hamiltonian import flux_link_tables
    for (pbc, apbc, B) in (((True, True), (False, True), 0.0), ((True, True), (False, True), 0.1),
                           ((True, False), (False, False), 0.05)):
        lat = DoubleKagome.build(1.0, 4, 4, pbc, apbc)
        tin, tint = flux_link_tables(lat, np.pi / 3)
        Hp = Hmat(lat, tin, tint, B=B)
        Ho = oracle_mod.OracleLattice(1.0, 4, 4, pbc, apbc).hmat(tin, tint, B=B)
        assert np.abs(Hp - Ho).max() < 1e-13
