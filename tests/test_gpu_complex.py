"""Complex orbitals on the device against the complex128 oracle, through the
C ABI: flux phase phi = pi/3 (every triangle carries flux phi, every hexagon
pi - 2 phi) plus a field B = 0.1 (the reference's Peierls phases,
src/Hamiltonian.jl:201, 325), so U, W and the local energy are complex as in
the reference (src/Hamiltonian.jl:382-393, src/MonteCarlo.jl:181-182, 290).

Bars: FP64 decisions and configurations bit-exact on shared tapes
(decisions within 1e-12 of a tie excepted), W, log|det| and the complex
local energy to 1e-9; FP32 W within 2e-4, local energies within 1e-4, with
near-tie flips checked by replaying the oracle and reported.
"""
import numpy as np
import pytest

from kagome_systems import BC, first_divergence_is_near_tie, tape

pytestmark = pytest.mark.gpu

PHI, B = np.pi / 3, 0.1
_CACHE = {}


def complex_system(L, phi=PHI, field=B, fillings=None):
    key = (L, phi, field, fillings)
    if key not in _CACHE:
        from kagome_vmc import DoubleKagome, Hamiltonian
        from kagome_vmc.hamiltonian import flux_link_tables
        pbc, apbc = BC[L]
        lat = DoubleKagome.build(1.0, L, L, pbc, apbc)
        n = 3 * L * L // 2
        nu, nd = fillings or (n, n)
        tin, tint = flux_link_tables(lat, phi)
        # the J1 list of the pi-flux geometry (same nonzero pattern)
        _CACHE[key] = Hamiltonian.build(nu, nd, lat, tin, tint, B=field)
    return _CACHE[key]


def conditioned(oracle_mod, ham, nw, seed, max_cond=1e5):
    ns, n = ham.ns, ham.N_up
    ku = np.zeros((nw, ns), dtype=np.int32)
    kd = np.zeros((nw, ns), dtype=np.int32)
    w = attempt = 0
    while w < nw:
        a, b = oracle_mod.random_config(seed, 30_000 + attempt, 0, ns, n, ham.N_down)
        attempt += 1
        if max(np.linalg.cond(oracle_mod.tilde_u(ham.U_up, a)),
               np.linalg.cond(oracle_mod.tilde_u(ham.U_down, b))) < max_cond:
            ku[w], kd[w] = a, b
            w += 1
    return ku, kd


def sampler(ham, nw, **kw):
    from kagome_vmc import _native as nat
    return nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nw, **kw)


def _max_rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_orbitals_are_complex():
    for L in (4, 8):
        ham = complex_system(L)
        assert np.abs(np.imag(ham.U_up)).max() > 0.1
        assert ham.nn == complex_system(L, 0.0, 0.0).nn


@pytest.mark.parametrize("L", [4, 8])
@pytest.mark.parametrize("prec", [0, 1])
def test_complex_refresh_matches_oracle(oracle_mod, L, prec):
    """W = U tildeU^-1 from the real 2N x 2N form of the refresh equals the
    oracle's complex LU + inverse + GEMM; log|det| likewise."""
    ham = complex_system(L)
    nw = 5
    ku, kd = conditioned(oracle_mod, ham, nw, 3)
    with sampler(ham, nw, precision=prec) as ds:
        assert ds.complex
        ds.set_configs(ku, kd)
        logdet = ds.logdet()
        W = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    tol = 1e-10 if prec == 0 else 1e-6
    for w in range(nw):
        Wu, lu = oracle_mod.refresh_w(ham.U_up, ku[w], with_logdet=True)
        Wd, ld = oracle_mod.refresh_w(ham.U_down, kd[w], with_logdet=True)
        assert _max_rel(W[w][0], Wu) < tol
        assert _max_rel(W[w][1], Wd) < tol
        assert abs(logdet[w] - (lu + ld)) < 1e-9 * max(1.0, abs(lu + ld))


@pytest.mark.parametrize("L,nsweeps", [(4, 400), (8, 300)])
@pytest.mark.parametrize("policy", [0, 1])
def test_complex_fp64_tape_chain_parity(oracle_mod, L, nsweeps, policy):
    """FP64 complex W, immediate complex rank-1 updates: identical tapes give
    identical accept/reject sequences, configurations and refresh counts;
    W and the local energy (real part, as getOL) agree to 1e-9."""
    ham = complex_system(L)
    nw = 6 if L == 4 else 3
    k = 7
    ku, kd = conditioned(oracle_mod, ham, nw, 21)
    tu, tb = tape(nw, nsweeps, len(ham.nn), 3)
    with sampler(ham, nw, refresh_policy=policy, refresh_k=k) as ds:
        ds.set_configs(ku, kd)
        cnt, out = ds.sweep(nsweeps, tu, tb, record=True)
        kf = ds.get_configs()
        e, _ = ds.measure()
        W = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    nref = 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, nr = oracle_mod.run_chain(ch, nsweeps, tape_u=tu[w], tape_bond=tb[w], policy=policy, k=k)
        nref += nr
        if not np.array_equal(out[w], oc):
            t, m, _ = first_divergence_is_near_tie(oracle_mod, ham, ku[w], kd[w], tu[w], tb[w], out[w], oc,
                                                   policy=policy, k=k)
            assert m < 1e-12, f"walker {w}: decision differs at sweep {t} with margin {m:.3e}"
            continue
        np.testing.assert_array_equal(kf[0][w], ch.kup)
        np.testing.assert_array_equal(kf[1][w], ch.kdn)
        assert _max_rel(W[w][0], ch.Wup) < 1e-9
        assert _max_rel(W[w][1], ch.Wdn) < 1e-9
        e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        assert abs(e[w] - e_or) < 1e-9 * max(1.0, abs(e_or))
    assert cnt["accepted"] == int((out == 3).sum()) > 0
    assert cnt["refreshes"] == nref


@pytest.mark.parametrize("L,delay", [(4, 4), (8, 8)])
def test_complex_fp64_delayed_tape_parity(oracle_mod, L, delay):
    """FP64 complex W with `delay` pending complex rank-1 terms against the
    oracle's immediate complex updates: identical decisions (ties within
    1e-12 excepted), W and the local energy to 1e-9."""
    ham = complex_system(L)
    nw, nsweeps = 3, 300
    ku, kd = conditioned(oracle_mod, ham, nw, 23)
    tu, tb = tape(nw, nsweeps, len(ham.nn), 5)
    with sampler(ham, nw, delay=delay) as ds:
        ds.set_configs(ku, kd)
        cnt, out = ds.sweep(nsweeps, tu, tb, record=True)
        kf = ds.get_configs()
        e, _ = ds.measure()
        W = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    assert cnt["flushes"] > 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, _ = oracle_mod.run_chain(ch, nsweeps, tape_u=tu[w], tape_bond=tb[w])
        if not np.array_equal(out[w], oc):
            t, m, _ = first_divergence_is_near_tie(oracle_mod, ham, ku[w], kd[w], tu[w], tb[w], out[w], oc)
            assert m < 1e-12, f"walker {w}: decision differs at sweep {t} with margin {m:.3e}"
            continue
        np.testing.assert_array_equal(kf[0][w], ch.kup)
        np.testing.assert_array_equal(kf[1][w], ch.kdn)
        assert _max_rel(W[w][0], ch.Wup) < 1e-9
        assert _max_rel(W[w][1], ch.Wdn) < 1e-9
        e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        assert abs(e[w] - e_or) < 1e-9 * max(1.0, abs(e_or))


@pytest.mark.parametrize("L,delay", [(4, 12), (8, 12), (8, 4)])
def test_complex_fp32_tape_parity(oracle_mod, L, delay):
    """FP32 complex W (W in LDS at L=4, streamed from HBM at L=8) with
    delayed complex rank-1 terms: decisions follow the oracle except at
    near ties; W within 2e-4 and local energies (pending terms folded in)
    within 1e-4."""
    from kagome_vmc import _native as nat
    ham = complex_system(L)
    nw = 6 if L == 4 else 3
    n1, n2, k = 250, 250, 64
    ku, kd = conditioned(oracle_mod, ham, nw, 41)
    tu, tb = tape(nw, n1 + n2, len(ham.nn), 7)
    with sampler(ham, nw, precision=nat.PREC_FP32, delay=delay, refresh_policy=nat.REFRESH_EVERY_K,
                 refresh_k=k) as ds:
        ds.set_configs(ku, kd)
        _, o1 = ds.sweep(n1, tu[:, :n1], tb[:, :n1], record=True)
        e_mid, _ = ds.measure()
        c2, o2 = ds.sweep(n2, tu[:, n1:], tb[:, n1:], record=True)
        e_end, _ = ds.measure()
        kf = ds.get_configs()
        W = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    out = np.concatenate([o1, o2], axis=1)
    near = 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc1, _, _ = oracle_mod.run_chain(ch, n1, tape_u=tu[w, :n1], tape_bond=tb[w, :n1], policy=1, k=k)
        e_or_mid = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        oc2, _, _ = oracle_mod.run_chain(ch, n2, tape_u=tu[w, n1:], tape_bond=tb[w, n1:], policy=1, k=k)
        oc = np.concatenate([oc1, oc2])
        if not np.array_equal(out[w], oc):
            t, m, ok = first_divergence_is_near_tie(oracle_mod, ham, ku[w], kd[w], tu[w], tb[w], out[w], oc,
                                                    policy=1, k=k)
            assert ok, f"walker {w}: decision differs at sweep {t} with margin {m:.3e} (not a near tie)"
            near += 1
            print(f"L={L} delay={delay} walker {w}: FP32 near-tie flip at sweep {t}, margin {m:.2e}")
            continue
        assert abs(e_mid[w] - e_or_mid) < 1e-4 * max(1.0, abs(e_or_mid))
        np.testing.assert_array_equal(kf[0][w], ch.kup)
        np.testing.assert_array_equal(kf[1][w], ch.kdn)
        assert _max_rel(W[w][0], ch.Wup) < 2e-4
        assert _max_rel(W[w][1], ch.Wdn) < 2e-4
        e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        assert abs(e_end[w] - e_or) < 1e-4 * max(1.0, abs(e_or))
    assert near <= 1
    assert c2["flushes"] > 0 and int((out == 3).sum()) > 0


@pytest.mark.parametrize("L,fillings", [(10, None), (12, None), (14, None), (10, (152, 148))])
def test_complex_refresh_single_kernel_path(oracle_mod, L, fillings):
    """At L = 10, 12 and 14 the real form has 2N = 300, 432 and 588 columns,
    beyond the register-tile solve: the single-kernel refresh path factors it
    (keeping the empty-site rows of the first ns sites only; two panel rows
    per thread at L = 10 and 12, three at L = 14), also with unequal
    fillings.  W and log|det| against the oracle; then a short FP64 tape
    chain."""
    ham = complex_system(L, fillings=fillings)
    nw = 2
    ku, kd = conditioned(oracle_mod, ham, nw, 5)
    tu, tb = tape(nw, 120, len(ham.nn), 9)
    with sampler(ham, nw) as ds:
        ds.set_configs(ku, kd)
        logdet = ds.logdet()
        for w in range(nw):
            Wu, lu = oracle_mod.refresh_w(ham.U_up, ku[w], with_logdet=True)
            Wd, ld = oracle_mod.refresh_w(ham.U_down, kd[w], with_logdet=True)
            assert _max_rel(ds.get_w(w, 0), Wu) < 1e-9
            assert _max_rel(ds.get_w(w, 1), Wd) < 1e-9
            assert abs(logdet[w] - (lu + ld)) < 1e-9 * max(1.0, abs(lu + ld))
        cnt, out = ds.sweep(120, tu, tb, record=True)
        kf = ds.get_configs()
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, _ = oracle_mod.run_chain(ch, 120, tape_u=tu[w], tape_bond=tb[w])
        np.testing.assert_array_equal(out[w], oc)
        np.testing.assert_array_equal(kf[0][w], ch.kup)


def test_complex_drift_single_kernel_path(oracle_mod):
    """Drift mode of the single-kernel refresh (complex L=10, real form
    2N = 300): with drift tracking on, every policy refresh inside
    kvmc_sweep first compares the carried FP32 W with fresh rows.  The count
    equals the refresh count, the drift is positive and small, and with
    tracking off the chains and energies are unchanged (the drift pass
    writes nothing)."""
    from kagome_vmc import _native as nat
    ham = complex_system(10)
    nw = 4
    ku, kd = conditioned(oracle_mod, ham, nw, 7)
    res = {}
    for track in (True, False):
        with sampler(ham, nw, precision=nat.PREC_FP32, delay=12, refresh_policy=nat.REFRESH_EVERY_K,
                     refresh_k=16, seed=5) as ds:
            ds.set_configs(ku, kd)
            ds.track_drift(track)
            cnt, _ = ds.sweep(300)
            res[track] = (cnt, ds.drift_stats(), ds.get_configs(), ds.measure()[0])
    cnt, (dmax, dmean, n), k_on, e_on = res[True]
    assert n == cnt["refreshes"] > 0
    assert 0.0 < dmean <= dmax < 1e-3
    cnt_off, stats_off, k_off, e_off = res[False]
    assert stats_off[2] == 0
    assert np.array_equal(k_on[0], k_off[0]) and np.array_equal(k_on[1], k_off[1])
    assert np.array_equal(e_on, e_off)
    print(f"complex L=10 drift before refresh: max {dmax:.2e}, mean {dmean:.2e} over {n}")


def test_complex_refresh_largest_supported_size(oracle_mod):
    """Complex L=16 (768 sites, 384 electrons per spin): the real form has
    2N = 768 columns, the most the single-kernel refresh takes (three panel
    rows per thread for every row of tildeU).  W and log|det| of one walker
    against the oracle (a start with cond(tildeU) below 1e6, so the bar is
    1e-8), then a short FP64 tape chain."""
    from kagome_vmc import DoubleKagome, Hamiltonian
    from kagome_vmc.hamiltonian import flux_link_tables
    lat = DoubleKagome.build(1.0, 16, 16, (True, True), (False, False))
    tin, tint = flux_link_tables(lat, PHI)
    ham = Hamiltonian.build(384, 384, lat, tin, tint, B=B)
    nw, nsw = 1, 60
    ku, kd = conditioned(oracle_mod, ham, nw, 5, max_cond=1e6)
    tu, tb = tape(nw, nsw, len(ham.nn), 9)
    with sampler(ham, nw) as ds:
        ds.set_configs(ku, kd)
        logdet = ds.logdet()
        Wu, lu = oracle_mod.refresh_w(ham.U_up, ku[0], with_logdet=True)
        Wd, ld = oracle_mod.refresh_w(ham.U_down, kd[0], with_logdet=True)
        eu, ed = _max_rel(ds.get_w(0, 0), Wu), _max_rel(ds.get_w(0, 1), Wd)
        print(f"complex L=16 refresh: W max rel {max(eu, ed):.2e}")
        assert eu < 1e-8 and ed < 1e-8
        assert abs(logdet[0] - (lu + ld)) < 1e-9 * max(1.0, abs(lu + ld))
        cnt, out = ds.sweep(nsw, tu, tb, record=True)
        kf = ds.get_configs()
    ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[0], kd[0])
    oc, _, _ = oracle_mod.run_chain(ch, nsw, tape_u=tu[0], tape_bond=tb[0])
    np.testing.assert_array_equal(out[0], oc)
    np.testing.assert_array_equal(kf[0][0], ch.kup)
    np.testing.assert_array_equal(kf[1][0], ch.kdn)


def test_mixed_random_flux_groups_follow_oracle(oracle_mod):
    """The mixed-size batch with a seeded flux phase per lattice (L = 4, 6,
    8, 10, 12; complex W in FP32 with delay 12, W in LDS for L <= 6 and
    streamed above; FP64 refresh every 64 accepts, the L = 10 and 12 ones on
    the single-kernel refresh path): two walkers per group replay in the
    oracle's complex128 chain (same Philox stream) over two measurement
    intervals, with the same configurations and local energies."""
    from kagome_vmc import _native as nat
    from kagome_vmc.batch import MixedBatch, flux_cluster
    rng = np.random.default_rng(11)
    sizes = ((4, 24), (6, 12), (8, 12), (10, 6), (12, 6))
    groups = [(flux_cluster(L, float(rng.uniform(0.0, np.pi))), nw) for L, nw in sizes]
    seed, k = 19, 64
    kw = dict(precision=nat.PREC_FP32, refresh_policy=nat.REFRESH_EVERY_K, refresh_k=k, delay=12, seed=seed)
    with MixedBatch(groups, **kw) as mb:
        assert all(g.sampler.complex for g in mb.groups)
        mb.init_random(3)
        k0 = [g.sampler.get_configs() for g in mb.groups]
        r1 = mb.step(1)
        k1 = [g.sampler.get_configs() for g in mb.groups]
        r2 = mb.step(1)
        k2 = [g.sampler.get_configs() for g in mb.groups]
    diverged = 0
    for g, (ham, nw) in enumerate(groups):
        n = min(ham.N_up, ham.N_down)
        for w in (0, nw - 1):
            ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, k0[g][0][w], k0[g][1][w])
            for kk, rr in ((k1, r1), (k2, r2)):
                oracle_mod.run_chain(ch, n, seed=seed + g, walker=w, policy=1, k=k)
                if not (np.array_equal(kk[g][0][w], ch.kup) and np.array_equal(kk[g][1][w], ch.kdn)):
                    diverged += 1
                    print(f"group L={sizes[g][0]} walker {w}: FP32 chain left the oracle's (near tie)")
                    break
                e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
                assert abs(rr[g][1][0][w] - e_or) < 1e-4 * max(1.0, abs(e_or)), (g, w)
    assert diverged <= 1
