"""Autotuner candidate set and cache, without the GPU (a stand-in timer)."""
import json

from kagome_vmc import _native as nat
from kagome_vmc.autotune import autotune, candidates, workload_key
from kagome_systems import system


def test_candidates_respect_the_column_floor():
    # L=12 FP32: 108 float4 rows per column -> at least 128 threads
    c = candidates(432, nat.PREC_FP32)
    assert sorted({t for t, _ in c}) == [128, 256] and sorted({d for _, d in c}) == [4, 8, 12]
    # FP64 has no delayed updates; 216 double2 rows -> 256 threads
    assert candidates(432, nat.PREC_FP64) == [(256, 1), (256, 4), (256, 8)]
    # small cluster: one wavefront is a candidate
    assert (64, 8) in candidates(48, nat.PREC_FP32)


def test_autotune_picks_fastest_and_caches(tmp_path):
    ham = system(4)
    calls = []

    def timer(t, d):
        calls.append((t, d))
        return 1e6 * (1 + (t == 64) + 0.1 * d)

    path = str(tmp_path / "tune.json")
    r = autotune(ham, 256, nat.PREC_FP32, 64, cache_path=path, timer=timer)
    assert (r["threads"], r["delay"]) == (64, 12)
    assert len(calls) == len(candidates(ham.ns, nat.PREC_FP32))
    with open(path) as f:
        assert workload_key(ham, 256, nat.PREC_FP32, 64) in json.load(f)
    # cached: no new timing
    r2 = autotune(ham, 256, nat.PREC_FP32, 64, cache_path=path, timer=timer)
    assert r2["threads"] == 64 and len(calls) == len(candidates(ham.ns, nat.PREC_FP32))
    # forced re-tune
    autotune(ham, 256, nat.PREC_FP32, 64, cache_path=path, timer=timer, refresh=True)
    assert len(calls) == 2 * len(candidates(ham.ns, nat.PREC_FP32))
