"""Parity of the HIP sampler (through the C ABI) with the CPU oracle.

Bars: bit-exact for configurations and accept/reject sequences (decisions
within 1e-12 of a tie excepted; none occur at these sizes and seeds), 1e-10
relative for W, determinant ratios and local energies in FP64.
"""
import numpy as np
import pytest

from kagome_systems import oracle_configs, system, well_conditioned_configs

pytestmark = pytest.mark.gpu

W_RTOL = 1e-10      # FP64 W / ratio / local-energy agreement (relative to max|W|)


def sampler(ham, nw, **kw):
    from kagome_vmc import _native as nat
    return nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nw, **kw)


def _max_rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("L", [2, 4, 6, 8, 12])
@pytest.mark.parametrize("prec", [0, 1])
def test_refresh_matches_oracle(oracle_mod, L, prec):
    ham = system(L)
    nw = 6 if L < 12 else 2
    ku, kd = well_conditioned_configs(oracle_mod, 11, nw, ham)
    with sampler(ham, nw, precision=prec) as ds:
        ds.set_configs(ku, kd)
        logdet = ds.logdet()
        for w in range(nw):
            Wu_o, lu = oracle_mod.refresh_w(ham.U_up, ku[w], with_logdet=True)
            Wd_o, ld = oracle_mod.refresh_w(ham.U_down, kd[w], with_logdet=True)
            tol = W_RTOL if prec == 0 else 1e-6
            assert _max_rel(ds.get_w(w, 0), Wu_o.real) < tol
            assert _max_rel(ds.get_w(w, 1), Wd_o.real) < tol
            assert abs(logdet[w] - (lu + ld)) < 1e-9 * max(1.0, abs(lu + ld))


@pytest.mark.parametrize("prec", [0, 1])
def test_refresh_many_walkers(oracle_mod, prec):
    """520 walkers refreshed in one launch pair: every walker's W and
    log-determinant equal the oracle's after set_configs, and W matches the
    oracle's refresh of the final configuration after sweeps that refresh
    most walkers at the end of the call (refresh_k = 1)."""
    ham = system(4)
    nw = 520
    ku, kd = well_conditioned_configs(oracle_mod, 13, nw, ham, max_cond=1e5)
    tol = W_RTOL if prec == 0 else 1e-6
    with sampler(ham, nw, precision=prec, refresh_policy=2, refresh_k=1, delay=1 if prec == 0 else 4) as ds:
        ds.set_configs(ku, kd)
        logdet = ds.logdet()
        for w in range(nw):
            Wu_o, lu = oracle_mod.refresh_w(ham.U_up, ku[w], with_logdet=True)
            Wd_o, ld = oracle_mod.refresh_w(ham.U_down, kd[w], with_logdet=True)
            assert _max_rel(ds.get_w(w, 0), Wu_o.real) < tol
            assert _max_rel(ds.get_w(w, 1), Wd_o.real) < tol
            assert abs(logdet[w] - (lu + ld)) < 1e-9 * max(1.0, abs(lu + ld))
        cnt, _ = ds.sweep(20)
        assert cnt["refreshes"] >= 256
        ku2, kd2 = ds.get_configs()
        for w in range(0, nw, 7):
            assert _max_rel(ds.get_w(w, 0), oracle_mod.refresh_w(ham.U_up, ku2[w]).real) < 100 * tol
            assert _max_rel(ds.get_w(w, 1), oracle_mod.refresh_w(ham.U_down, kd2[w]).real) < 100 * tol


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("n_up", [264, 240])
def test_refresh_above_256_electrons(oracle_mod, prec, n_up):
    """N_up = 264 at L=12 (432 sites, N_down = 168): beyond 256 electrons of
    one spin the refresh runs as a single kernel over [tildeU; U_empty].
    N_up = 240 (N_down = 192) is the widest two-kernel case: 15 column
    blocks, the solve kernel's 16-tile instance.  Walker 0 starts from the
    pivoted-QR configuration, walker 1 from it with 20 seeded up/down
    exchanges, each kept only while both tildeU stay well conditioned (these
    fillings are not closed shells, so U, and with it the conditioning of a
    given configuration, depends on the eigensolver)."""
    from kagome_vmc import DoubleKagome, Hamiltonian, init_conf_qr
    lat = DoubleKagome.build(1.0, 12, 12, (True, True), (False, True))
    ham = Hamiltonian.build(n_up, 432 - n_up, lat)
    ku0, kd0 = init_conf_qr(ham.U_up, ham.U_down)
    ku1, kd1 = ku0.copy(), kd0.copy()
    rng = np.random.default_rng(5)
    swaps = 0
    while swaps < 20:
        i = rng.choice(np.flatnonzero(ku1))
        j = rng.choice(np.flatnonzero(kd1))
        a, b = ku1.copy(), kd1.copy()
        a[i], a[j] = 0, a[i]
        b[j], b[i] = 0, b[j]
        if max(np.linalg.cond(oracle_mod.tilde_u(ham.U_up, a)),
               np.linalg.cond(oracle_mod.tilde_u(ham.U_down, b))) < 1e5:
            ku1, kd1 = a, b
            swaps += 1
    ku, kd = np.stack([ku0, ku1]), np.stack([kd0, kd1])
    tol = W_RTOL if prec == 0 else 1e-6
    with sampler(ham, 2, precision=prec) as ds:
        ds.set_configs(ku, kd)
        logdet = ds.logdet()
        for w in range(2):
            Wu, lu = oracle_mod.refresh_w(ham.U_up, ku[w], with_logdet=True)
            Wd, ld = oracle_mod.refresh_w(ham.U_down, kd[w], with_logdet=True)
            assert _max_rel(ds.get_w(w, 0), Wu.real) < tol
            assert _max_rel(ds.get_w(w, 1), Wd.real) < tol
            assert abs(logdet[w] - (lu + ld)) < 1e-9 * max(1.0, abs(lu + ld))
        if prec == 0:
            ds.sweep(40)
            e, _ = ds.measure()
            ku2, kd2 = ds.get_configs()
            for w in range(2):
                Wu = oracle_mod.refresh_w(ham.U_up, ku2[w])
                Wd = oracle_mod.refresh_w(ham.U_down, kd2[w])
                eo = oracle_mod.getol(ham.nn, ku2[w], kd2[w], Wu, Wd)
                assert abs(e[w] - eo) < 1e-8 * max(1.0, abs(eo))


@pytest.mark.parametrize("n_up", [25, 30])
def test_refresh_unequal_filling(oracle_mod, n_up):
    """N_up != N_down (the reference's imbalance scan, scripts/FP.jl:112-116)."""
    from kagome_vmc import DoubleKagome, Hamiltonian
    lat = DoubleKagome.build(1.0, 4, 4, (True, True), (False, True))
    ham = Hamiltonian.build(n_up, 48 - n_up, lat)
    nw = 4
    ku = np.zeros((nw, 48), dtype=np.int32)
    kd = np.zeros((nw, 48), dtype=np.int32)
    w = q = 0
    while w < nw:
        a, b = oracle_mod.random_config(3, q, 0, 48, n_up, 48 - n_up)
        q += 1
        if max(np.linalg.cond(oracle_mod.tilde_u(ham.U_up, a)),
               np.linalg.cond(oracle_mod.tilde_u(ham.U_down, b))) < 1e6:
            ku[w], kd[w] = a, b
            w += 1
    with sampler(ham, nw) as ds:
        ds.set_configs(ku, kd)
        for w in range(nw):
            assert _max_rel(ds.get_w(w, 0), oracle_mod.refresh_w(ham.U_up, ku[w]).real) < W_RTOL
            assert _max_rel(ds.get_w(w, 1), oracle_mod.refresh_w(ham.U_down, kd[w]).real) < W_RTOL
        ds.sweep(50)
        e, _ = ds.measure()
        ku2, kd2 = ds.get_configs()
        for w in range(nw):
            Wu = oracle_mod.refresh_w(ham.U_up, ku2[w])
            Wd = oracle_mod.refresh_w(ham.U_down, kd2[w])
            eo = oracle_mod.getol(ham.nn, ku2[w], kd2[w], Wu, Wd)
            assert abs(e[w] - eo) < 1e-8 * max(1.0, abs(eo))


@pytest.mark.parametrize("L", [4, 8])
def test_local_energy_matches_oracle(oracle_mod, L):
    ham = system(L)
    nw = 8
    ku, kd = well_conditioned_configs(oracle_mod, 5, nw, ham)
    with sampler(ham, nw) as ds:
        ds.set_configs(ku, kd)
        e, nap = ds.measure()
    ref = []
    napo = 0
    for w in range(nw):
        Wu = oracle_mod.refresh_w(ham.U_up, ku[w])
        Wd = oracle_mod.refresh_w(ham.U_down, kd[w])
        ref.append(oracle_mod.getol(ham.nn, ku[w], kd[w], Wu, Wd))
        napo += oracle_mod.z_count(ham.nn, ku[w], kd[w])
    ref = np.array(ref)
    assert np.all(np.abs(e - ref) <= W_RTOL * np.maximum(1.0, np.abs(ref)))
    assert nap == napo


def test_init_random_matches_oracle_draws(oracle_mod):
    ham = system(4)
    nw = 16
    with sampler(ham, nw) as ds:
        redraws = ds.init_random(seed=99)
        ku, kd = ds.get_configs()
    assert redraws >= 0
    for w in range(nw):
        for attempt in range(65):
            a, b = oracle_mod.random_config(99, w, attempt, ham.ns, ham.N_up, ham.N_down)
            if np.array_equal(a, ku[w]) and np.array_equal(b, kd[w]):
                break
        else:
            pytest.fail(f"walker {w}: device start is not an oracle draw")


def _tape(nw, nsweeps, nb, seed):
    rng = np.random.default_rng(seed)
    return rng.random((nw, nsweeps)), rng.integers(0, nb, (nw, nsweeps)).astype(np.int32)


@pytest.mark.parametrize("L,nsweeps", [(2, 400), (4, 400), (8, 200)])
@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("policy", [0, 1])
def test_tape_chain_parity(oracle_mod, L, nsweeps, kernel, policy):
    """Identical tapes (u, bond) give identical accept/reject sequences,
    configurations and refresh counts; W and E_L agree to 1e-10."""
    ham = system(L)
    nw = 8 if L < 8 else 3
    k = 7
    ku, kd = well_conditioned_configs(oracle_mod, 21, nw, ham)
    tu, tb = _tape(nw, nsweeps, len(ham.nn), 3)
    with sampler(ham, nw, kernel=kernel, refresh_policy=policy, refresh_k=k) as ds:
        ds.set_configs(ku, kd)
        cnt, out = ds.sweep(nsweeps, tu, tb, record=True)
        kfu, kfd = ds.get_configs()
        e_dev, _ = ds.measure()
        Wdev = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    total_refresh = 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, nref = oracle_mod.run_chain(ch, nsweeps, tape_u=tu[w], tape_bond=tb[w],
                                           kernel=kernel, policy=policy, k=k)
        total_refresh += nref
        np.testing.assert_array_equal(out[w], oc)
        np.testing.assert_array_equal(kfu[w], ch.kup)
        np.testing.assert_array_equal(kfd[w], ch.kdn)
        assert _max_rel(Wdev[w][0], ch.Wup.real) < 1e-9
        assert _max_rel(Wdev[w][1], ch.Wdn.real) < 1e-9
        e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        assert abs(e_dev[w] - e_or) < 1e-9 * max(1.0, abs(e_or))
    assert cnt["refreshes"] == total_refresh
    assert cnt["sweeps"] == nw * nsweeps
    assert cnt["accepted"] == int((out == 3).sum())


def test_philox_chain_parity(oracle_mod):
    """Device-RNG chains replay exactly in the oracle (same Philox stream)."""
    ham = system(4)
    nw, nsweeps, seed = 6, 500, 1234
    ku, kd = well_conditioned_configs(oracle_mod, 8, nw, ham)
    with sampler(ham, nw, seed=seed) as ds:
        ds.set_configs(ku, kd)
        ds.set_sweep_count(0)
        _, out1 = ds.sweep(200, record=True)
        _, out2 = ds.sweep(nsweeps - 200, record=True)
        kfu, _ = ds.get_configs()
        assert np.all(ds.sweep_counts() == nsweeps)
    out = np.concatenate([out1, out2], axis=1)
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, _ = oracle_mod.run_chain(ch, nsweeps, seed=seed, walker=w)
        np.testing.assert_array_equal(out[w], oc)
        np.testing.assert_array_equal(kfu[w], ch.kup)


@pytest.mark.parametrize("delay", [1, 4, 8, 12, 16])
def test_fp32_delayed_update_tracks_fp64(oracle_mod, delay):
    """FP32 W with `delay` pending rank-1 terms: W after the same tape agrees
    with the FP64 oracle chain to FP32 accuracy, decisions coincide except
    near ties at FP32 resolution."""
    ham = system(4)
    nw, nsweeps = 8, 300
    ku, kd = well_conditioned_configs(oracle_mod, 31, nw, ham)
    tu, tb = _tape(nw, nsweeps, len(ham.nn), 9)
    with sampler(ham, nw, precision=1, delay=delay, refresh_policy=1, refresh_k=64) as ds:
        ds.set_configs(ku, kd)
        cnt, out = ds.sweep(nsweeps, tu, tb, record=True)
        e32, _ = ds.measure()
        kfu, kfd = ds.get_configs()
        W32 = [ds.get_w(w, 0) for w in range(nw)]
        # get_w flushed every walker: E_loc from the flushed W equals E_loc
        # with the pending terms folded in, to FP32 rounding
        e32_flushed, _ = ds.measure()
    np.testing.assert_allclose(e32_flushed, e32, rtol=2e-5, atol=2e-5)
    same = 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, _ = oracle_mod.run_chain(ch, nsweeps, tape_u=tu[w], tape_bond=tb[w], policy=1, k=64)
        if np.array_equal(out[w], oc):
            same += 1
            assert _max_rel(W32[w], ch.Wup.real) < 2e-4
            e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
            assert abs(e32[w] - e_or) < 1e-4 * max(1.0, abs(e_or))
    assert same >= nw - 1
    if delay > 1:
        assert cnt["flushes"] <= cnt["accepted"] // delay + nw


@pytest.mark.parametrize("n_up", [20, 30])
def test_fp32_delayed_update_unequal_filling(oracle_mod, n_up):
    """FP32 W at delay 12 with N_up != N_down at L=4: the two spins' W have
    different row counts (28 and 20 empty-site rows for n_up = 20), and the
    pending columns of both use the taller one's stride.  W and E_L track
    the FP64 oracle chain on the same tape, and E_L with the pending terms
    folded in equals E_L after a flush."""
    from kagome_vmc import DoubleKagome, Hamiltonian
    lat = DoubleKagome.build(1.0, 4, 4, (True, True), (False, True))
    ham = Hamiltonian.build(n_up, 48 - n_up, lat)
    nw, nsweeps = 6, 300
    ku = np.zeros((nw, 48), dtype=np.int32)
    kd = np.zeros((nw, 48), dtype=np.int32)
    w = q = 0
    while w < nw:
        a, b = oracle_mod.random_config(5, q, 0, 48, n_up, 48 - n_up)
        q += 1
        if max(np.linalg.cond(oracle_mod.tilde_u(ham.U_up, a)),
               np.linalg.cond(oracle_mod.tilde_u(ham.U_down, b))) < 1e5:
            ku[w], kd[w] = a, b
            w += 1
    tu, tb = _tape(nw, nsweeps, len(ham.nn), 23)
    with sampler(ham, nw, precision=1, delay=12, refresh_policy=1, refresh_k=10**6) as ds:
        ds.set_configs(ku, kd)
        cnt, out = ds.sweep(nsweeps, tu, tb, record=True)
        e32, _ = ds.measure()
        Wu = [ds.get_w(w, 0) for w in range(nw)]
        Wd = [ds.get_w(w, 1) for w in range(nw)]
        e32_flushed, _ = ds.measure()
    np.testing.assert_allclose(e32_flushed, e32, rtol=2e-5, atol=2e-5)
    assert cnt["accepted"] > 12 * nw // 2   # several flushes and live pending terms
    same = 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, _ = oracle_mod.run_chain(ch, nsweeps, tape_u=tu[w], tape_bond=tb[w], policy=1, k=10**6)
        if np.array_equal(out[w], oc):
            same += 1
            assert _max_rel(Wu[w], ch.Wup.real) < 2e-4
            assert _max_rel(Wd[w], ch.Wdn.real) < 2e-4
            e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
            assert abs(e32[w] - e_or) < 1e-4 * max(1.0, abs(e_or))
    assert same >= nw - 1


def test_refresh_cadence_reference_policy(oracle_mod):
    """Reference cadence: a refresh after every sweep with counter % N == 0
    that passed both early returns (src/MonteCarlo.jl:593-604)."""
    ham = system(4)
    nw, nsweeps = 4, 240
    ku, kd = well_conditioned_configs(oracle_mod, 41, nw, ham)
    tu, tb = _tape(nw, nsweeps, len(ham.nn), 17)
    with sampler(ham, nw) as ds:
        ds.set_configs(ku, kd)
        cnt, out = ds.sweep(nsweeps, tu, tb, record=True)
    n_occ = ham.N_up
    expect = sum(1 for w in range(nw) for s in range(nsweeps)
                 if s % n_occ == 0 and out[w, s] >= 2)
    assert cnt["refreshes"] == expect


def test_singular_start_reports_walker():
    """A rank-1 U makes every tildeU singular (test/test-MonteCarlo.jl:429-477)."""
    from kagome_vmc import _native as nat
    ham = system(2)
    U = np.repeat(np.random.default_rng(0).random((1, 6)), 12, axis=0)
    ds = nat.DeviceSampler(ham.bonds_array(), U, ham.U_down, 3)
    ku = np.zeros((3, 12), dtype=np.int32)
    kd = np.zeros((3, 12), dtype=np.int32)
    for w in range(3):
        ku[w, :6] = np.arange(1, 7)
        kd[w, 6:] = np.arange(1, 7)
    with pytest.raises(nat.SingularError):
        ds.set_configs(ku, kd)
    assert ds.failed_walker() == 0
    with pytest.raises(nat.KvmcError):
        ds.sweep(1)           # W was never initialised
    ds.close()


def test_invalid_configuration_rejected():
    from kagome_vmc import _native as nat
    ham = system(2)
    with sampler(ham, 1) as ds:
        ku = np.zeros((1, 12), dtype=np.int32)
        kd = np.zeros((1, 12), dtype=np.int32)
        ku[0, :6] = np.arange(1, 7)
        kd[0, 5:11] = np.arange(1, 7)       # site 6 doubly occupied, site 12 empty
        with pytest.raises(nat.KvmcError, match="doubly occupied"):
            ds.set_configs(ku, kd)


def test_checkpoint_roundtrip(tmp_path, oracle_mod):
    from kagome_vmc import MC, MCContext, init, read_checkpoint, run, write_checkpoint
    params = dict(n1=4, n2=4, PBC=(True, True), antiPBC=(False, True), N_up=24, N_down=24,
                  nwalkers=8, seed=5)
    mc = MC(params)
    ctx = MCContext(seed=5, thermalization=10)
    init(mc, ctx, params)
    run(mc, ctx, 100)
    path = str(tmp_path / "ckpt.npz")
    write_checkpoint(mc, ctx, path)
    e1, _ = mc.sampler.measure()
    mc2 = MC(params)
    ctx2 = MCContext(seed=5, thermalization=10)
    read_checkpoint(mc2, ctx2, path)
    assert ctx2.sweeps == ctx.sweeps
    np.testing.assert_array_equal(mc2.kappa[0], mc.kappa[0])
    e2, _ = mc2.sampler.measure()
    assert np.allclose(e1, e2, rtol=1e-9, atol=1e-9)
    # both continue identically (same Philox stream, same counters)
    run(mc, ctx, 50)
    run(mc2, ctx2, 50)
    np.testing.assert_array_equal(mc2.kappa[0], mc.kappa[0])
    mc.close(); mc2.close()


def test_w_drift_is_small_after_many_updates(oracle_mod):
    ham = system(8)
    with sampler(ham, 8, precision=1, delay=8, refresh_policy=1, refresh_k=10**6, seed=3) as ds:
        ds.init_random(3)
        ds.sweep(2000)
        d32 = ds.w_drift()
    with sampler(ham, 8, precision=0, refresh_policy=1, refresh_k=10**6, seed=3) as ds:
        ds.init_random(3)
        ds.sweep(2000)
        d64 = ds.w_drift()
    assert d64 < 1e-8
    assert d32 < 1e-2


def test_stagger_refresh_moves_refreshes_not_decisions(oracle_mod):
    """kvmc_stagger_refresh only shifts when walkers refresh: the chains'
    configurations after a run equal those of an unstaggered run (FP64 W,
    where a refresh changes W only by rounding)."""
    ham = system(4)
    nw = 64
    ku, kd = well_conditioned_configs(oracle_mod, 21, nw, ham)
    out = []
    for stagger in (False, True):
        with sampler(ham, nw, refresh_policy=1, refresh_k=16, seed=5) as ds:
            ds.set_configs(ku, kd)
            if stagger:
                ds.stagger_refresh(9)
            cnt, _ = ds.sweep(60)
            out.append((ds.get_configs(), cnt["refreshes"], cnt["accepted"]))
    (c0, r0, a0), (c1, r1, a1) = out
    assert np.array_equal(c0[0], c1[0]) and np.array_equal(c0[1], c1[1])
    assert a0 == a1
    assert r1 > r0      # staggered walkers reach their first refresh sooner


@pytest.mark.parametrize("L,prec,policy,k", [(4, 1, 1, 8), (4, 0, 0, 0), (8, 1, 1, 16), (8, 0, 1, 4)])
def test_graph_replay_matches_launch_loop(oracle_mod, L, prec, policy, k):
    """kvmc_enable_graph: every replay (with eager measurements and a change
    of nsweeps between replays) leaves the same configurations, energies and
    counters as the ordinary launch loop."""
    ham = system(L)
    nw = 96
    ku, kd = well_conditioned_configs(oracle_mod, 31, nw, ham)
    runs = []
    for graph in (False, True):
        with sampler(ham, nw, precision=prec, refresh_policy=policy, refresh_k=k, delay=4 if prec else 1,
                     seed=13) as ds:
            ds.set_configs(ku, kd)
            if graph:
                ds.enable_graph(True)
            rec = []
            for n in (ham.N_up, ham.N_up, 7, ham.N_up):
                cnt, _ = ds.sweep(n)
                e, _ = ds.measure()
                rec.append((cnt, e, ds.get_configs()))
            runs.append(rec)
    for (c0, e0, k0), (c1, e1, k1) in zip(*runs):
        for key in ("sweeps", "moves", "accepted", "refreshes", "flushes"):
            assert c0[key] == c1[key], key
        assert np.array_equal(e0, e1)
        assert np.array_equal(k0[0], k1[0]) and np.array_equal(k0[1], k1[1])


def test_refresh_at_end_policy(oracle_mod):
    """AT_END refreshes a walker after the call in which it reaches k
    accepts: decisions equal those of EVERY_K (FP64 W), one sweep launch per
    call, and no walker is refreshed more than once per call."""
    ham = system(4)
    nw = 64
    ku, kd = well_conditioned_configs(oracle_mod, 41, nw, ham)
    out = {}
    for policy in (1, 2):
        with sampler(ham, nw, refresh_policy=policy, refresh_k=8, seed=17) as ds:
            ds.set_configs(ku, kd)
            rec = []
            for _ in range(3):
                cnt, _ = ds.sweep(48)
                rec.append(cnt)
            out[policy] = (rec, ds.get_configs(), ds.measure()[0])
    (r1, c1, e1), (r2, c2, e2) = out[1], out[2]
    assert np.array_equal(c1[0], c2[0]) and np.array_equal(c1[1], c2[1])
    assert np.allclose(e1, e2, rtol=1e-10, atol=1e-10)
    for a, b in zip(r1, r2):
        assert a["accepted"] == b["accepted"]
        assert b["launches"] == 1 and 0 < b["refreshes"] <= nw
        assert a["refreshes"] >= b["refreshes"]
