"""Pins the CPU oracle (oracle/kvmc_oracle.c) before it is trusted.

Each test restates a known answer held by the reference's own test suite
(test/test-Lattice.jl, test/test-Hamiltonian.jl, test/test-MonteCarlo.jl;
file:line given per test), a published known-answer vector (Random123
Philox KAT), or a mathematical fact chosen so that a plausible slip (a dropped
term, a wrong sign or index, a transposed operand) fails.
"""
import itertools

import numpy as np
import pytest


# --- Philox4x32-10 known-answer vectors (Random123 kat_vectors) ------------
@pytest.mark.parametrize("ctr,key,expect", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_kat(oracle_mod, ctr, key, expect):
    assert oracle_mod.philox4x32_10(ctr, key) == expect


def test_sweep_draw_ranges(oracle_mod):
    us, bs = zip(*[oracle_mod.sweep_draws(7, 3, s, 96) for s in range(4000)])
    assert min(us) >= 0.0 and max(us) < 1.0
    assert set(bs) == set(range(96))
    assert abs(np.mean(us) - 0.5) < 0.02


def test_random_config_is_sz0_permutation(oracle_mod):
    ku, kd = oracle_mod.random_config(11, 5, 0, 48, 24, 24)
    assert sorted(ku[ku > 0]) == list(range(1, 25))
    assert sorted(kd[kd > 0]) == list(range(1, 25))
    assert np.all((ku != 0) ^ (kd != 0))
    ku2, _ = oracle_mod.random_config(11, 5, 1, 48, 24, 24)
    assert not np.array_equal(ku, ku2)


# --- Lattice: test/test-Lattice.jl ------------------------------------------
def test_lattice_validation(oracle_mod):
    with pytest.raises(AssertionError):           # test-Lattice.jl:3
        oracle_mod.OracleLattice(1.0, 3, 3, (False, False))
    assert oracle_mod.OracleLattice(1.0, 4, 3, (False, False)).ns == 36   # :6-7
    with pytest.raises(ValueError):               # test-Hamiltonian.jl:66-79
        oracle_mod.OracleLattice(1.0, 4, 3, (False, False), (True, False))
    with pytest.raises(ValueError):
        oracle_mod.OracleLattice(1.0, 4, 3, (False, False), (False, True))


def test_nearest_neighbours(oracle_mod):
    # test/test-Lattice.jl:10-45
    H1 = oracle_mod.OracleLattice(1.0, 4, 4, (True, False)).hmat()
    assert H1.shape[0] == 48
    nn = oracle_mod.get_nn(H1)
    assert all(i < j for i, j in nn)
    assert (1, 2) in nn and (1, 3) in nn and (2, 3) in nn and (1, 4) not in nn
    H2 = oracle_mod.OracleLattice(1.0, 4, 4, (True, True), (True, True)).hmat()
    nn2 = oracle_mod.get_nn(H2)
    assert len(nn) != len(nn2)
    assert all(p in nn2 for p in nn)
    assert all(H1[i - 1, j - 1] != 0 for i, j in nn)
    # column-major order: sorted by j, then i (src/Hamiltonian.jl:447-451)
    assert nn == sorted(nn, key=lambda p: (p[1], p[0]))


# --- Hmat: test/test-Hamiltonian.jl:4-61 ------------------------------------
def test_hmat_known_entries(oracle_mod):
    H = oracle_mod.OracleLattice(1.0, 4, 3, (False, False)).hmat()
    assert np.allclose(H, H.conj().T)
    assert H[0, 0] == 0 and H[0, 1] == -1 and H[0, 2] == -1
    assert np.allclose(H[0, 3:36], 0)
    assert np.isclose(H[2, 12], 1)
    H2 = oracle_mod.OracleLattice(1.0, 4, 3, (True, False)).hmat()
    assert H2[0, 1] == -1 and H2[0, 2] == -1 and H2[0, 10] == -1 and H2[0, 12] == 0
    ax = oracle_mod.OracleLattice(1.0, 4, 3, (True, False), (True, False)).hmat()
    assert ax[0, 1] == -1 and ax[0, 2] == -1 and ax[4, 6] == -1 and ax[0, 10] == 1
    ay = oracle_mod.OracleLattice(1.0, 4, 3, (True, True), (False, True)).hmat()
    assert ay[0, 10] == -1 and ay[0, 2 + 4 * 6] == -1
    ab = oracle_mod.OracleLattice(1.0, 4, 3, (True, True), (True, True)).hmat()
    assert ab[0, 10] == 1 and ab[0, 26] == -1 and ab[10, 26] == 1
    mx = oracle_mod.OracleLattice(1.0, 4, 3, (True, True), (True, False)).hmat()
    assert mx[0, 10] == 1 and mx[0, 2 + 4 * 6] == 1


def test_peierls_phases(oracle_mod):
    # test/test-Hamiltonian.jl:722-775
    lat = oracle_mod.OracleLattice(1.0, 4, 3, (False, False))
    B = 0.1
    HB = lat.hmat(B=B)
    assert np.allclose(HB, HB.conj().T)
    r1, r3 = lat.site_coord(1), lat.site_coord(3)
    ph = (B / 2) * (r1[0] + r3[0]) * (r3[1] - r1[1])
    assert HB[0, 2] == -np.exp(1j * ph)
    assert HB[0, 1] == -1.0
    r13 = lat.site_coord(13)
    rr3 = lat.site_coord(3)
    ph = (B / 2) * (rr3[0] + r13[0]) * (r13[1] - rr3[1])
    assert HB[2, 12] == -oracle_mod.PI_LINK_INTER[(3, 1, 0, 1)] * np.exp(1j * ph)
    latp = oracle_mod.OracleLattice(1.0, 4, 3, (True, True))
    HBp = latp.hmat(B=B)
    r1, r11 = latp.site_coord(1), latp.site_coord(11)
    r2 = r11 - 2 * np.array([4.0, 0.0])
    ph = (B / 2) * (r1[0] + r2[0]) * (r2[1] - r1[1])
    assert HBp[0, 10] == -oracle_mod.PI_LINK_INTER[(1, 5, -1, 0)] * np.exp(1j * ph)


def test_custom_link_inter(oracle_mod):
    # apply_boundary_conditions! cases, test/test-Hamiltonian.jl:332-400: a
    # single inter-cell link table, no in-cell links.  H = -(T + T').
    n1, n2 = 4, 3
    cases = [((True, True), (False, False), {(1, 5, -1, 0): 1.0, (5, 1, 1, 0): 1.0}, 1, 11, 1.0),
             ((True, True), (True, False), {(1, 5, -1, 0): 1.0, (5, 1, 1, 0): 1.0}, 1, 11, -1.0),
             ((True, True), (False, True), {(3, 6, 0, -1): 1.0, (6, 3, 0, 1): 1.0}, 3,
              3 * 2 * n1 + 6, -1.0)]
    for pbc, apbc, link, s1, s2, expect in cases:
        H = oracle_mod.OracleLattice(1.0, n1, n2, pbc, apbc).hmat(link_in={}, link_inter=link)
        assert np.isclose(H[s1 - 1, s2 - 1], -expect)
    lat = oracle_mod.OracleLattice(1.0, n1, n2, (True, True), (True, True))
    H = lat.hmat(link_in={}, link_inter={(1, 6, -1, -1): 1.0})
    assert np.isclose(H[0, 35], -1.0)     # diagonal crossing: signs cancel
    H = lat.hmat(link_in={}, link_inter={(1, 5, 1, 0): 1.0})
    assert np.isclose(H[0, 6], 0.0)


def test_geometry_helpers(oracle_mod):
    # test/test-Hamiltonian.jl:467-720
    lat = oracle_mod.OracleLattice(1.0, 4, 2, (False, False))
    a1, a2 = np.array([4.0, 0.0]), np.array([1.0, np.sqrt(3.0)])
    assert np.allclose(lat.unitcell_coord(1), [0, 0]) and np.allclose(lat.unitcell_coord(6), [0, 0])
    assert np.allclose(lat.unitcell_coord(7), a1)
    assert np.allclose(lat.unitcell_coord(24), a1 + a2)
    assert np.allclose(lat.unitcell_coord(13), a2)
    lat3 = oracle_mod.OracleLattice(1.0, 4, 3, (False, False))
    for c1, c2, dx, dy in [(a1, [0, 0], 1, 0), (a2, [0, 0], 0, 1), (a1 + a2, [0, 0], 1, 1),
                           ([0, 0], a1, -1, 0), ([0, 0], a2, 0, -1), (2 * a1 + 2 * a2, [0, 0], 2, 2),
                           ([0.1, 0.1], [0, 0], 0, 0), (a1 + [0.1, 0.1], [0, 0], 1, 0),
                           (a2 - [0.1, 0.1], [0, 0], 0, 1), ([5.0, np.sqrt(3.0)], [0, 0], 1, 1)]:
        assert lat3.unitcell_diff(c1, c2) == (dx, dy)
    assert np.allclose(lat3.site_coord(1), [0, 0])
    assert np.allclose(lat3.site_coord(2), [1, 0])
    assert np.allclose(lat3.site_coord(3), [0.5, 0.5 * np.sqrt(3.0)])
    assert np.allclose(lat3.site_coord(7), [4, 0])
    assert np.allclose(lat3.site_coord(13), [1, np.sqrt(3.0)])


def test_boundary_shifts(oracle_mod):
    # test/test-Hamiltonian.jl:606-700
    n1, n2 = 4, 3
    lat = oracle_mod.OracleLattice(1.0, n1, n2, (True, True))
    assert (1, 0, 1.0) in lat.boundary_shifts(3, 9)
    sh = lat.boundary_shifts(3, 6 * (n1 // 2 - 1) + 3)
    assert (-(n1 // 2 - 1), 0, 1.0) in sh and (n1 // 2 + 1, 0, 1.0) in sh
    assert oracle_mod.OracleLattice(1.0, n1, n2, (False, False)).boundary_shifts(3, 9) == [(1, 0, 1.0)]
    shx = oracle_mod.OracleLattice(1.0, n1, n2, (True, False)).boundary_shifts(3, 9)
    assert len(shx) > 1 and all(s[1] == 0 for s in shx)
    sha = oracle_mod.OracleLattice(1.0, n1, n2, (True, False), (True, False)).boundary_shifts(
        3, 6 * (n1 // 2 - 1) + 3)
    assert any(s[2] == -1.0 for s in sha)
    shd = lat.boundary_shifts(3, 6 * n1 // 2 + 9)
    assert any(s[0] != 0 and s[1] != 0 for s in shd)


# --- getxprime / spinInteraction! / Sz: test/test-Hamiltonian.jl ------------
def test_getxprime_known_answer(oracle_mod):
    # test/test-Hamiltonian.jl:97-135: 36-site OBC, one up spin at site 1
    H = oracle_mod.OracleLattice(1.0, 4, 3, (False, False)).hmat()
    nn = oracle_mod.get_nn(H)
    ku = np.array([1] + [0] * 35)
    kd = np.array([0] + list(range(1, 36)))
    xp = oracle_mod.getxprime(nn, ku, kd)
    assert len(xp) == 3
    assert xp[(-1, -1, -1, -1)] == (len(nn) - 2) * 0.25 + 2 * (-0.25)
    assert xp[(2, 1, 1, 1)] == -0.5
    assert xp[(3, 1, 1, 2)] == -0.5


def test_spin_interaction_known_answers(oracle_mod):
    # test/test-Hamiltonian.jl:237-321, through single-bond lists
    ku, kd = np.array([1, 0, 2]), np.array([0, 1, 0])
    assert oracle_mod.getxprime([(2, 1)], ku, kd)[(2, 1, 1, 1)] == -0.5
    assert oracle_mod.getxprime([(1, 2)], ku, kd)[(2, 1, 1, 1)] == -0.5
    assert len(oracle_mod.getxprime([(1, 3)], ku, kd)) == 1      # both up: diagonal only
    assert oracle_mod.getxprime([(2, 3)], ku, kd)[(2, 2, 3, 1)] == -0.5
    assert oracle_mod.getxprime([(1, 2), (1, 2)], ku, kd)[(2, 1, 1, 1)] == -1.0
    with pytest.raises(ValueError):      # Sz ArgumentError: doubly occupied site
        oracle_mod.getxprime([(1, 2)], np.array([1, 2, 0]), np.array([0, 2, 1]))


def test_z_known_answer(oracle_mod):
    # test/test-MonteCarlo.jl:175-183
    assert oracle_mod.z_count([(1, 2), (2, 3), (1, 3)], [0, 1, 0], [1, 0, 2]) == 2


def test_tilde_u_known_answers(oracle_mod):
    # test/test-MonteCarlo.jl:54-173
    U = np.array([[1.0, 2, 3], [4, 5, 6], [7, 8, 9]])
    r = oracle_mod.tilde_u(U, [2, 3, 1])
    assert np.array_equal(r[0], U[2]) and np.array_equal(r[1], U[0]) and np.array_equal(r[2], U[1])
    with pytest.raises(ValueError):
        oracle_mod.tilde_u(np.array([[1.0, 2], [3, 4]]), [0, 0])
    U = np.array([[1.0, 2], [3, 4], [5, 6]])
    r = oracle_mod.tilde_u(U, [1, 0, 2])
    assert np.array_equal(r[0], U[0]) and np.array_equal(r[1], U[2])
    U = np.array([[1 + 1j, 2 + 2j], [3 + 3j, 4 + 4j]])
    r = oracle_mod.tilde_u(U, [2, 1])
    assert np.array_equal(r[0], U[1]) and np.array_equal(r[1], U[0])
    with pytest.raises(IndexError):
        oracle_mod.tilde_u(np.array([[1.0, 2], [3, 4]]), [3, 1])
    with pytest.raises(ValueError):
        oracle_mod.tilde_u(np.array([[1.0, 2], [3, 4]]), [1, 2, 3])


def test_update_w_closed_form(oracle_mod):
    # test/test-MonteCarlo.jl:235-252: W=[1 .2; .2 1], l=1, K=2 -> [5 -4.8; 1 0]
    W = np.asfortranarray(np.array([[1.0, 0.2], [0.2, 1.0]], dtype=np.complex128))
    oracle_mod.update_w(W, 1, 2)
    assert np.allclose(W, [[5.0, -4.8], [1.0, 0.0]], atol=1e-14)


def test_update_w_is_sherman_morrison(oracle_mod):
    # After update_W!(W, l, K): W' = U tildeU'^-1 where row l of tildeU is
    # replaced by U[K, :] (the particle labelled l hops to site K).
    rng = np.random.default_rng(1)
    U = rng.standard_normal((10, 4))
    kap = np.array([1, 0, 2, 0, 3, 0, 0, 4, 0, 0])
    W = oracle_mod.refresh_w(U, kap)
    oracle_mod.update_w(W, 2, 4)          # label 2 moves from site 3 to site 4
    kap2 = kap.copy(); kap2[2] = 0; kap2[3] = 2
    assert np.allclose(W, oracle_mod.refresh_w(U, kap2), atol=1e-12)
    assert np.allclose(W[3], np.eye(4)[1], atol=1e-12)


def test_refresh_formula(oracle_mod):
    # test/test-MonteCarlo.jl:185-232
    U = np.array([[1.0, 0.2], [0.2, 1.0]])
    W = oracle_mod.refresh_w(U, [1, 2])
    assert np.allclose(W, U @ np.linalg.inv(U), atol=1e-10)
    W = oracle_mod.refresh_w(np.array([[1.0, 0.3], [0.3, 1.0]]), [2, 1])
    tU = np.array([[0.3, 1.0], [1.0, 0.3]])
    assert np.allclose(W, np.array([[1.0, 0.3], [0.3, 1.0]]) @ np.linalg.inv(tU), atol=1e-10)
    rng = np.random.default_rng(2)
    U = rng.standard_normal((12, 6))
    kap = np.zeros(12, dtype=int)
    kap[[1, 4, 5, 7, 9, 11]] = [3, 1, 6, 2, 5, 4]
    W, ld = oracle_mod.refresh_w(U, kap, with_logdet=True)
    tU = oracle_mod.tilde_u(U, kap)
    assert np.allclose(W, U @ np.linalg.inv(tU), atol=1e-12)
    assert np.isclose(ld, np.log(abs(np.linalg.det(tU))))
    # ratio property: W[K, l] = det(tildeU with row l -> U[K]) / det(tildeU)
    for K in range(12):
        for l in range(6):
            t2 = tU.copy(); t2[l] = U[K]
            assert np.isclose(W[K, l], np.linalg.det(t2) / np.linalg.det(tU), atol=1e-10)
    # basis independence: W does not change under U -> U M
    M = rng.standard_normal((6, 6))
    assert np.allclose(oracle_mod.refresh_w(U @ M, kap), W, atol=1e-9)


def test_refresh_singular(oracle_mod):
    # test/test-MonteCarlo.jl:429-477: identical rows -> singular
    U = np.repeat(np.random.default_rng(3).random((1, 6)), 12, axis=0)
    with pytest.raises(np.linalg.LinAlgError):
        oracle_mod.refresh_w(U, [1, 2, 3, 4, 5, 6, 0, 0, 0, 0, 0, 0])


# --- exhaustive 12-site energies: independent checks ------------------------
@pytest.mark.parametrize("pbc,apbc", [((False, False), (False, False)),
                                      ((True, True), (False, False)),
                                      ((True, True), (False, True)),
                                      ((True, False), (False, False))])
def test_exhaustive_energy_is_quadratic_form(oracle_mod, pbc, apbc):
    """sum_x |psi|^2 E_L(x) / sum |psi|^2 (E_L from W, getOL) must equal
    psi^T H psi / psi^T psi with H the spin-1/2 Heisenberg matrix and psi from
    site-ordered determinants with the fermion reordering sign.  A wrong sign
    of the exchange term or a wrong W index fails this."""
    H = oracle_mod.OracleLattice(1.0, 2, 2, pbc, apbc).hmat()
    nn = oracle_mod.get_nn(H)
    Uu, Ud, _ = oracle_mod.orbitals(H, 6, 6)
    r = oracle_mod.exhaustive(nn, Uu, Ud)
    assert abs(r["E"] - r["E_quadratic_form"]) < 1e-11
    # variational principle against exact diagonalisation
    Hm, _ = oracle_mod.heisenberg_dense(12, nn, 6)
    assert r["E"] >= np.linalg.eigvalsh(Hm)[0] - 1e-12


def test_exhaustive_local_energy_matches_brute_force_ratios(oracle_mod):
    """E_L(x) from getOL equals sum_x' H_x'x psi(x')/psi(x) with psi from the
    spin-basis amplitudes (no W involved)."""
    H = oracle_mod.OracleLattice(1.0, 2, 2, (True, True), (False, True)).hmat()
    nn = oracle_mod.get_nn(H)
    Uu, Ud, _ = oracle_mod.orbitals(H, 6, 6)
    r = oracle_mod.exhaustive(nn, Uu, Ud)
    Hm, confs = oracle_mod.heisenberg_dense(12, nn, 6)
    psi = r["psi_spin"]
    el = (Hm.T @ psi) / psi
    ok = np.abs(psi) > 1e-8
    assert np.allclose(r["eloc"][ok], el[ok].real, atol=1e-9)


def test_reference_kernel_stationary_law(oracle_mod):
    """The reference transition kernel (shared uniform r, src/MonteCarlo.jl:
    546-591) accepts x -> x' through bond b iff r < (Z_mu/Z_max) min(1, a).
    Building P from the oracle's sweep decisions at thresholds, its
    stationary vector is |psi|^2 / Z_mu."""
    H = oracle_mod.OracleLattice(1.0, 2, 2, (True, False)).hmat()
    nn = oracle_mod.get_nn(H)
    Uu, Ud, _ = oracle_mod.orbitals(H, 6, 6)
    r = oracle_mod.exhaustive(nn, Uu, Ud)
    confs = r["confs"]
    index = {c: q for q, c in enumerate(confs)}
    D, nb = len(confs), len(nn)
    P = np.zeros((D, D))
    rng = np.random.default_rng(4)
    for q, (ku, kd) in enumerate(oracle_mod.all_configs(12, 6)):
        if r["weight"][q] == 0:
            continue
        zr = r["zmu"][q] / nb
        for b in range(nb):
            ch = oracle_mod.Chain.start(nn, Uu, Ud, ku, kd)
            i, j = nn[b]
            if (ku[i - 1] != 0) == (ku[j - 1] != 0):
                continue
            c2 = list(confs[q]); c2[i - 1], c2[j - 1] = c2[j - 1], c2[i - 1]
            q2 = index[tuple(c2)]
            a = r["weight"][q2] / r["weight"][q]
            thr = zr * min(1.0, a)
            # probe the oracle's decision just below and just above the threshold
            # (skipped near nodes, where the W ratio is dominated by roundoff)
            well_conditioned = a > 1e-10 and r["weight"][q] > 1e-12 * r["weight"].max()
            probes = [] if not well_conditioned else [(thr * (1 - 1e-9), True),
                                           (min(thr * (1 + 1e-9), 0.999999), False)]
            for rr, expect in probes:
                if not expect and thr >= 0.999999:
                    continue
                c = oracle_mod.Chain.start(nn, Uu, Ud, ku, kd)
                oc, _, _ = oracle_mod.run_chain(c, 1, tape_u=[rr], tape_bond=[b], policy=1, k=10**6)
                assert (oc[0] == oracle_mod.ACCEPT) == expect
            P[q2, q] += thr / nb
        P[q, q] = 1.0 - P[:, q].sum()
    w = np.where(r["zmu"] > 0, r["weight"] / np.maximum(r["zmu"], 1), 0)
    w /= w.sum()
    assert np.allclose(P @ w, w, atol=1e-12)
    assert rng is not None
