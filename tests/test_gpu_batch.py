"""Mixed cluster sizes on one GPU: each group of a MixedBatch runs the same
chain it runs alone (groups are independent handles on their own streams)."""
import numpy as np
import pytest

from kagome_vmc import _native as nat
from kagome_vmc.batch import MixedBatch, pi_flux_cluster

pytestmark = pytest.mark.gpu


def test_groups_match_their_solo_runs():
    groups = [(pi_flux_cluster(2), 64), (pi_flux_cluster(4), 32), (pi_flux_cluster(6), 16)]
    kw = dict(precision=nat.PREC_FP32, refresh_policy=nat.REFRESH_EVERY_K, refresh_k=16, delay=4,
              seed=3)
    with MixedBatch(groups, **kw) as mb:
        mb.init_random(11)
        res = mb.step(3)
        confs = [g.sampler.get_configs() for g in mb.groups]
    for g, (ham, nw) in enumerate(groups):
        solo_kw = dict(kw, seed=kw["seed"] + g)
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


def test_autotune_on_device(tmp_path):
    from kagome_vmc.autotune import autotune, candidates
    ham = pi_flux_cluster(4)
    r = autotune(ham, 128, nat.PREC_FP32, 32, steps=1, cache_path=str(tmp_path / "t.json"))
    assert (r["threads"], r["delay"]) in candidates(ham.ns, nat.PREC_FP32)
    assert r["walker_sweeps_per_s"] > 0 and len(r["trials"]) == len(candidates(ham.ns, nat.PREC_FP32))
