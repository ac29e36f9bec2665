"""Parity of the FP32 sampler with W streamed from HBM (the benchmarked
kernel instances) against the FP64 oracle, through the C ABI.

The FP32 form keeps W, the ratios and up to `delay` pending rank-1 terms in
float; the oracle keeps complex128 W with immediate updates
(src/MonteCarlo.jl:279-292, 538-607).  On the same (u, bond) tapes the two
chains take the same decisions except where a ratio lies within FP32
resolution of its Metropolis threshold; such a flip is accepted only after
replaying the oracle to the first differing sweep and checking that the
decision there was a near tie (relative margin below 1e-3), and it is
reported.  Bars for walkers whose decisions agree:
  - configurations bit-exact,
  - W within 2e-4 of the oracle's W relative to max|W| (both spins),
  - local energies (FP64 accumulation over FP32 W with live pending terms)
    within 1e-4 relative,
  - refresh counts equal to the oracle's under the same policy.
"""
import numpy as np
import pytest

from kagome_systems import first_divergence_is_near_tie, stagger_since, system, tape, \
    well_conditioned_configs

pytestmark = pytest.mark.gpu

W_TOL32 = 2e-4      # FP32 W against the FP64 oracle, relative to max|W|
E_TOL32 = 1e-4      # FP32 local energy against the oracle, relative


def sampler(ham, nw, **kw):
    from kagome_vmc import _native as nat
    return nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nw, **kw)


def _max_rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _pending_after(outcomes, delay, k, since0=0):
    """Pending rank-1 terms left after an EVERY_K run with these outcomes:
    a flush every `delay` accepts, cleared by each refresh (k accepts)."""
    c, acc = 0, since0
    for oc in outcomes:
        if oc == 3:
            acc += 1
            c += 1
            if c == delay:
                c = 0
            if acc >= k:
                acc, c = 0, 0
    return c


@pytest.mark.parametrize("L,delay", [(8, 12), (8, 16), (12, 12), (12, 16)])
def test_fp32_streamed_w_tape_parity(oracle_mod, L, delay):
    """k_sweep<float, 4, 12|16, streamed W> and the workgroup local-energy
    kernel (pending terms folded in) against the oracle chain on a shared
    tape, in two calls with a measurement between them."""
    from kagome_vmc import _native as nat
    ham = system(L)
    nw = 3 if L == 8 else 2
    n1, n2, k = 300, 300, 64
    ku, kd = well_conditioned_configs(oracle_mod, 61, nw, ham, max_cond=1e5)
    tu, tb = tape(nw, n1 + n2, len(ham.nn), 29)
    with sampler(ham, nw, precision=nat.PREC_FP32, delay=delay, refresh_policy=nat.REFRESH_EVERY_K,
                 refresh_k=k) as ds:
        ds.set_configs(ku, kd)
        c1, o1 = ds.sweep(n1, tu[:, :n1], tb[:, :n1], record=True)
        e_mid, _ = ds.measure()
        k_mid = ds.get_configs()
        c2, o2 = ds.sweep(n2, tu[:, n1:], tb[:, n1:], record=True)
        e_end, _ = ds.measure()
        k_end = ds.get_configs()
        W = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    out = np.concatenate([o1, o2], axis=1)
    # the middle measurement ran with live pending terms for some walker,
    # and every walker flushed at least once before it
    pend = [_pending_after(o1[w], delay, k) for w in range(nw)]
    assert max(pend) > 0, pend
    assert np.all((o1 == 3).sum(1) >= delay)
    assert c1["flushes"] + c2["flushes"] >= nw
    near_ties = 0
    n_refresh = 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc1, _, r1 = oracle_mod.run_chain(ch, n1, tape_u=tu[w, :n1], tape_bond=tb[w, :n1], policy=1, k=k)
        e_or_mid = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        ku_mid, kd_mid = ch.kup.copy(), ch.kdn.copy()
        oc2, _, r2 = oracle_mod.run_chain(ch, n2, tape_u=tu[w, n1:], tape_bond=tb[w, n1:], policy=1, k=k)
        oc = np.concatenate([oc1, oc2])
        if not np.array_equal(out[w], oc):
            t, m, ok = first_divergence_is_near_tie(oracle_mod, ham, ku[w], kd[w], tu[w], tb[w], out[w], oc,
                                                    policy=1, k=k)
            assert ok, f"walker {w}: decision differs at sweep {t} with margin {m:.3e} (not a near tie)"
            near_ties += 1
            print(f"L={L} delay={delay} walker {w}: FP32 near-tie flip at sweep {t}, margin {m:.2e}")
            continue
        n_refresh += r1 + r2
        np.testing.assert_array_equal(k_mid[0][w], ku_mid)
        np.testing.assert_array_equal(k_mid[1][w], kd_mid)
        assert abs(e_mid[w] - e_or_mid) < E_TOL32 * max(1.0, abs(e_or_mid)), (w, e_mid[w], e_or_mid)
        np.testing.assert_array_equal(k_end[0][w], ch.kup)
        np.testing.assert_array_equal(k_end[1][w], ch.kdn)
        e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        assert abs(e_end[w] - e_or) < E_TOL32 * max(1.0, abs(e_or)), (w, e_end[w], e_or)
        assert _max_rel(W[w][0], ch.Wup.real) < W_TOL32
        assert _max_rel(W[w][1], ch.Wdn.real) < W_TOL32
    assert near_ties <= 1
    if near_ties == 0:
        assert c1["refreshes"] + c2["refreshes"] == n_refresh
    print(f"L={L} delay={delay}: {nw - near_ties}/{nw} walkers identical, {near_ties} near-tie flips, "
          f"{n_refresh} refreshes, pending at the middle measurement {pend}")


@pytest.mark.parametrize("k", [16, 64, 216])
@pytest.mark.parametrize("policy", ["every_k", "at_end"])
def test_fp32_refresh_policies_L12(oracle_mod, policy, k):
    """L=12, FP32, delay 12, W from HBM, refresh every k accepted moves
    (EVERY_K: mid-call, as soon as a walker reaches k; AT_END: after the
    call in which it reaches k), with refresh phases staggered as in the
    benchmark.  Decisions and configurations follow the oracle's EVERY_K
    chain (a refresh changes W only by rounding); the refresh count follows
    the policy; the carried W and local energies agree with an FP64 refresh
    of the final configurations."""
    from kagome_vmc import _native as nat
    ham = system(12)
    nw, calls, S, seed_st = 4, 3, 216, 77
    pol = nat.REFRESH_EVERY_K if policy == "every_k" else nat.REFRESH_AT_END
    ku, kd = well_conditioned_configs(oracle_mod, 71, nw, ham, max_cond=1e5)
    tu, tb = tape(nw, calls * S, len(ham.nn), 31)
    since0 = stagger_since(oracle_mod, seed_st, nw, k)
    with sampler(ham, nw, precision=nat.PREC_FP32, delay=12, refresh_policy=pol, refresh_k=k) as ds:
        ds.set_configs(ku, kd)
        ds.stagger_refresh(seed_st)
        outs, cnts = [], []
        for c in range(calls):
            cnt, o = ds.sweep(S, tu[:, c * S:(c + 1) * S], tb[:, c * S:(c + 1) * S], record=True)
            outs.append(o)
            cnts.append(cnt)
        drift = ds.w_drift()
        e_dev, _ = ds.measure()
        kf = ds.get_configs()
        W = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    out = np.concatenate(outs, axis=1)
    refreshes = sum(c["refreshes"] for c in cnts)
    # expected refreshes under the policy
    if policy == "at_end":
        expect = 0
        for w in range(nw):
            acc = since0[w]
            for o in outs:
                acc += int((o[w] == 3).sum())
                if acc >= k:
                    expect += 1
                    acc = 0
        assert all(c["launches"] == 1 for c in cnts)
    near_ties = 0
    n_or_refresh = 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        ch.acc_since = since0[w]
        oc, _, nr = oracle_mod.run_chain(ch, calls * S, tape_u=tu[w], tape_bond=tb[w], policy=1, k=k)
        n_or_refresh += nr
        if not np.array_equal(out[w], oc):
            t, m, ok = first_divergence_is_near_tie(oracle_mod, ham, ku[w], kd[w], tu[w], tb[w], out[w], oc,
                                                    policy=1, k=k, acc_since=since0[w])
            assert ok, f"walker {w}: decision differs at sweep {t} with margin {m:.3e} (not a near tie)"
            near_ties += 1
            print(f"{policy} k={k} walker {w}: FP32 near-tie flip at sweep {t}, margin {m:.2e}")
            continue
        np.testing.assert_array_equal(kf[0][w], ch.kup)
        np.testing.assert_array_equal(kf[1][w], ch.kdn)
        Wu = oracle_mod.refresh_w(ham.U_up, ch.kup)
        Wd = oracle_mod.refresh_w(ham.U_down, ch.kdn)
        assert _max_rel(W[w][0], Wu.real) < W_TOL32
        assert _max_rel(W[w][1], Wd.real) < W_TOL32
        e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, Wu, Wd)
        assert abs(e_dev[w] - e_or) < E_TOL32 * max(1.0, abs(e_or)), (w, e_dev[w], e_or)
    assert near_ties <= 1
    assert drift < W_TOL32
    if near_ties == 0:
        assert refreshes == (n_or_refresh if policy == "every_k" else expect)
    assert refreshes > 0
    print(f"{policy} k={k}: refreshes {refreshes}, accepts {int((out == 3).sum())}, W drift {drift:.2e}, "
          f"near-tie flips {near_ties}")


def test_fp64_streamed_w_tape_parity_L12(oracle_mod):
    """FP64, immediate updates, W streamed from HBM at L=12, reference
    refresh cadence: bit-exact decisions and configurations, W and local
    energies to 1e-9."""
    ham = system(12)
    nw, n = 2, 433
    ku, kd = well_conditioned_configs(oracle_mod, 81, nw, ham)
    tu, tb = tape(nw, n, len(ham.nn), 37)
    # u = 0 never pre-rejects: the cadence sweeps (counter % 216 == 0) reach
    # the ratio whenever their bond is antiparallel, so refreshes happen
    tu[:, ::ham.N_up] = 0.0
    with sampler(ham, nw) as ds:
        ds.set_configs(ku, kd)
        cnt, out = ds.sweep(n, tu, tb, record=True)
        e, _ = ds.measure()
        kf = ds.get_configs()
        W = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    nref = 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, nr = oracle_mod.run_chain(ch, n, tape_u=tu[w], tape_bond=tb[w])
        nref += nr
        np.testing.assert_array_equal(out[w], oc)
        np.testing.assert_array_equal(kf[0][w], ch.kup)
        np.testing.assert_array_equal(kf[1][w], ch.kdn)
        assert _max_rel(W[w][0], ch.Wup.real) < 1e-9
        assert _max_rel(W[w][1], ch.Wdn.real) < 1e-9
        e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        assert abs(e[w] - e_or) < 1e-9 * max(1.0, abs(e_or))
    assert cnt["refreshes"] == nref and nref > 0


@pytest.mark.parametrize("L,delay", [(4, 4), (8, 8)])
def test_fp64_delayed_updates_tape_parity(oracle_mod, L, delay):
    """FP64 W with `delay` pending rank-1 terms (W resident in LDS at L=4,
    streamed at L=8) against the oracle's immediate updates: the terms are
    the same Sherman-Morrison factors, so decisions are identical except
    within 1e-12 of a tie, and W and local energies agree to 1e-9."""
    ham = system(L)
    nw, n = 3, 300
    ku, kd = well_conditioned_configs(oracle_mod, 83, nw, ham)
    tu, tb = tape(nw, n, len(ham.nn), 41)
    with sampler(ham, nw, delay=delay) as ds:
        ds.set_configs(ku, kd)
        cnt, out = ds.sweep(n, tu, tb, record=True)
        e, _ = ds.measure()
        kf = ds.get_configs()
        W = [(ds.get_w(w, 0), ds.get_w(w, 1)) for w in range(nw)]
    assert cnt["flushes"] > 0
    for w in range(nw):
        ch = oracle_mod.Chain.start(ham.nn, ham.U_up, ham.U_down, ku[w], kd[w])
        oc, _, _ = oracle_mod.run_chain(ch, n, tape_u=tu[w], tape_bond=tb[w])
        if not np.array_equal(out[w], oc):
            t, m, _ = first_divergence_is_near_tie(oracle_mod, ham, ku[w], kd[w], tu[w], tb[w], out[w], oc)
            assert m < 1e-12, f"walker {w}: decision differs at sweep {t} with margin {m:.3e}"
            continue
        np.testing.assert_array_equal(kf[0][w], ch.kup)
        np.testing.assert_array_equal(kf[1][w], ch.kdn)
        assert _max_rel(W[w][0], ch.Wup.real) < 1e-9
        assert _max_rel(W[w][1], ch.Wdn.real) < 1e-9
        e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
        assert abs(e[w] - e_or) < 1e-9 * max(1.0, abs(e_or))


def test_mixed_groups_follow_oracle_chains(oracle_mod):
    """The mixed-size batch (L = 4, 6, 8, 10, 12 groups, FP32 W with delay 12,
    FP64 refresh every 64 accepts; W in LDS for L <= 6, streamed above):
    for two walkers of every group, the device-RNG chain replays in the
    oracle (same Philox stream, seed + group) over two measurement intervals,
    with the same configurations and local energies after each."""
    from kagome_vmc import _native as nat
    from kagome_vmc.batch import MixedBatch, pi_flux_cluster
    sizes = ((4, 32), (6, 16), (8, 16), (10, 8), (12, 8))
    groups = [(pi_flux_cluster(L), nw) for L, nw in sizes]
    seed, k = 41, 64
    kw = dict(precision=nat.PREC_FP32, refresh_policy=nat.REFRESH_EVERY_K, refresh_k=k, delay=12, seed=seed)
    with MixedBatch(groups, **kw) as mb:
        mb.init_random(5)
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
            ok = True
            for step, (kk, rr) in enumerate(((k1, r1), (k2, r2))):
                oracle_mod.run_chain(ch, n, seed=seed + g, walker=w, policy=1, k=k)
                if not (np.array_equal(kk[g][0][w], ch.kup) and np.array_equal(kk[g][1][w], ch.kdn)):
                    ok = False
                    break
                e_or = oracle_mod.getol(ham.nn, ch.kup, ch.kdn, ch.Wup, ch.Wdn)
                e_dev = rr[g][1][0][w]
                assert abs(e_dev - e_or) < E_TOL32 * max(1.0, abs(e_or)), (g, w, step, e_dev, e_or)
            if not ok:
                diverged += 1
                print(f"group L={sizes[g][0]} walker {w}: FP32 chain left the oracle's (near tie)")
    assert diverged <= 1
    for g, (ham, nw) in enumerate(groups):
        assert r1[g][0]["sweeps"] == nw * min(ham.N_up, ham.N_down)


def test_drift_before_refresh_is_recorded(oracle_mod):
    """kvmc_track_drift: every policy refresh inside kvmc_sweep first measures
    the carried W against a fresh FP64 refresh.  The count equals the
    refresh count, the drift is positive (FP32 W drifts) and small, and
    with tracking off the chains are unchanged (the drift pass writes
    nothing).  An FP64 handle drifts by rounding only."""
    from kagome_vmc import _native as nat
    ham = system(8)
    nw = 32
    ku, kd = well_conditioned_configs(oracle_mod, 91, nw, ham, max_cond=1e5)
    res = {}
    for prec, track in ((nat.PREC_FP32, True), (nat.PREC_FP32, False), (nat.PREC_FP64, True)):
        with sampler(ham, nw, precision=prec, delay=12 if prec else 1, refresh_policy=nat.REFRESH_EVERY_K,
                     refresh_k=16, seed=9) as ds:
            ds.set_configs(ku, kd)
            ds.track_drift(track)
            cnt, _ = ds.sweep(400)
            res[(prec, track)] = (cnt, ds.drift_stats(), ds.get_configs(), ds.measure()[0])
    cnt, (dmax, dmean, n), k32, e32 = res[(nat.PREC_FP32, True)]
    assert n == cnt["refreshes"] > 0
    assert 0.0 < dmean <= dmax < 1e-3
    cnt_off, stats_off, k_off, e_off = res[(nat.PREC_FP32, False)]
    assert stats_off[2] == 0
    assert np.array_equal(k32[0], k_off[0]) and np.array_equal(k32[1], k_off[1])
    assert np.array_equal(e32, e_off)
    cnt64, (d64, _, n64), _, _ = res[(nat.PREC_FP64, True)]
    assert n64 == cnt64["refreshes"] > 0 and d64 < 1e-10
    print(f"FP32 drift before refresh: max {dmax:.2e}, mean {dmean:.2e} over {n}; FP64 max {d64:.2e}")
