"""Mixed cluster sizes on one GPU: each group of a MixedBatch runs the same
chain it runs alone (groups are independent handles on their own streams)."""
import numpy as np
import pytest

from kagome_vmc import _native as nat
from kagome_vmc.batch import MixedBatch, pi_flux_cluster

pytestmark = pytest.mark.gpu


def test_groups_match_their_solo_runs():
    """The middle group carries its own launch shape (the autotuner's kind of
    override); its solo twin is built with the same threads and delay."""
    groups = [(pi_flux_cluster(2), 64, None), (pi_flux_cluster(4), 32, {"threads": 64, "delay": 8}),
              (pi_flux_cluster(6), 16, None)]
    kw = dict(precision=nat.PREC_FP32, refresh_policy=nat.REFRESH_EVERY_K, refresh_k=16, delay=4,
              seed=3)
    with MixedBatch(groups, **kw) as mb:
        mb.init_random(11)
        res = mb.step(3)
        confs = [g.sampler.get_configs() for g in mb.groups]
    for g, (ham, nw, shape) in enumerate(groups):
        solo_kw = dict(kw, seed=kw["seed"] + g, **(shape or {}))
        with nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nw, **solo_kw) as ds:
            ds.init_random(11)
            e_solo = []
            for _ in range(3):
                ds.sweep(ham.N_up)
                e_solo.append(ds.measure()[0])
            ku, kd = ds.get_configs()
        assert np.array_equal(ku, confs[g][0]) and np.array_equal(kd, confs[g][1])
        for a, b in zip(res[g][1], e_solo):
            assert np.array_equal(a, b)
        assert res[g][0]["sweeps"] == 3 * nw * ham.N_up


@pytest.mark.parametrize("prec", [nat.PREC_FP32, nat.PREC_FP64])
def test_autotune_on_device(tmp_path, prec):
    from kagome_vmc.autotune import autotune, candidates
    ham = pi_flux_cluster(4)
    r = autotune(ham, 128, prec, 32, steps=1, cache_path=str(tmp_path / "t.json"))
    assert (r["threads"], r["delay"]) in candidates(ham.ns, prec)
    assert r["walker_sweeps_per_s"] > 0 and len(r["trials"]) == len(candidates(ham.ns, prec))


def test_group_graph_matches_per_handle_sweeps():
    """kvmc_group_sweep (one graph launch over handles of L = 4, 8, 12, real
    and complex orbitals, FP32 and FP64) leaves every handle exactly where
    its own kvmc_sweep calls leave a twin handle: configurations, local
    energies and counters, over calls that change the sweep counts."""
    from kagome_vmc.batch import flux_cluster
    specs = [(pi_flux_cluster(4), 64, dict(precision=nat.PREC_FP32, delay=12)),
             (flux_cluster(8, 0.7), 32, dict(precision=nat.PREC_FP32, delay=12)),
             (pi_flux_cluster(12), 16, dict(precision=nat.PREC_FP64)),
             (pi_flux_cluster(12), 16, dict(precision=nat.PREC_FP32, delay=12))]
    calls = [[24, 96, 216, 216], [24, 96, 216, 216], [10, 0, 216, 50]]

    def make(g, ham, nw, kw):
        ds = nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nw, refresh_policy=nat.REFRESH_EVERY_K,
                               refresh_k=16, seed=40 + g, **kw)
        ds.init_random(7 + g)
        return ds

    solo = [make(g, *sp) for g, sp in enumerate(specs)]
    grouped = [make(g, *sp) for g, sp in enumerate(specs)]
    try:
        with nat.SamplerGroup(grouped) as grp:
            for ns in calls:
                cg = grp.sweep(ns)
                for g, (ds1, ds2) in enumerate(zip(solo, grouped)):
                    c1, _ = ds1.sweep(ns[g])
                    for key in ("sweeps", "moves", "accepted", "refreshes", "flushes"):
                        assert c1[key] == cg[g][key], (g, key)
                    k1, k2 = ds1.get_configs(), ds2.get_configs()
                    assert np.array_equal(k1[0], k2[0]) and np.array_equal(k1[1], k2[1])
                    assert np.array_equal(ds1.measure()[0], ds2.measure()[0])
                assert sum(c["refreshes"] for c in cg) > 0
    finally:
        for ds in solo + grouped:
            ds.close()
