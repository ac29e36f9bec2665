"""Launch-shape autotuner for the sweep kernel.

Two knobs set how the sweep kernel maps onto the GPU:
  * threads -- workgroup size (one workgroup per walker): more threads
               stream W faster in a flush, fewer let more walkers share a
               CU and hide the latency of the proposals;
  * delay   -- accepted moves kept as pending rank-1 terms before
               W is streamed (fewer HBM passes, more LDS and register use).
The best pair depends on the cluster size, the precision and the walker
count, so it is measured: each candidate runs a short timed stretch of the
real workload on the device and the fastest wins.  Results are cached in a
JSON file keyed by the workload.
"""
from __future__ import annotations

import json
import os
import time

from . import _native as nat

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "kagome_vmc", "autotune.json")


def candidates(ns: int, precision: int):
    """(threads, delay) pairs worth timing.  threads below the number of
    W row vectors (ns / 4 in FP32, ns / 2 in FP64) are raised to it by the
    library, so only distinct effective sizes are listed."""
    vec = 4 if precision == nat.PREC_FP32 else 2
    nvec = ns // vec if ns % vec == 0 else ns
    floor = (nvec + 63) // 64 * 64
    threads = sorted({max(t, floor) for t in (64, 128, 256) if max(t, floor) <= 512})
    delays = (4, 8, 12) if precision == nat.PREC_FP32 else (1, 4, 8)
    return [(t, d) for t in threads for d in delays]


def workload_key(ham, nwalkers: int, precision: int, refresh_k: int) -> str:
    return f"gfx950:ns={ham.ns}:up={ham.N_up}:dn={ham.N_down}:nw={nwalkers}:prec={precision}:k={refresh_k}"


def _time_candidate(ham, nwalkers, precision, refresh_k, threads, delay, seed, steps):
    ds = nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nwalkers, precision=precision,
                           refresh_policy=nat.REFRESH_EVERY_K, refresh_k=refresh_k, delay=delay,
                           threads=threads, seed=seed)
    try:
        ds.init_random(seed)
        ds.stagger_refresh(seed)
        nocc = min(ham.N_up, ham.N_down)
        ds.sweep(nocc)                       # warm-up (first launches, caches)
        t0 = time.perf_counter()
        cnt, _ = ds.sweep(nocc * steps)
        wall = time.perf_counter() - t0
        return cnt["sweeps"] / wall
    finally:
        ds.close()


def autotune(ham, nwalkers: int, precision: int = nat.PREC_FP32, refresh_k: int = 216, *, seed: int = 0,
             steps: int = 2, cache_path: str | None = DEFAULT_CACHE, timer=None, refresh: bool = False):
    """Returns {"threads", "delay", "walker_sweeps_per_s", "trials"} for the
    fastest candidate.  `timer(threads, delay) -> rate` replaces the device
    measurement (tests); `refresh=True` ignores a cached entry."""
    key = workload_key(ham, nwalkers, precision, refresh_k)
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            cache = json.load(f)
        if key in cache and not refresh:
            return cache[key]
    if timer is None:
        def timer(t, d):
            return _time_candidate(ham, nwalkers, precision, refresh_k, t, d, seed, steps)
    trials = []
    for t, d in candidates(ham.ns, precision):
        trials.append({"threads": t, "delay": d, "walker_sweeps_per_s": float(timer(t, d))})
    best = max(trials, key=lambda r: r["walker_sweeps_per_s"])
    result = {"threads": best["threads"], "delay": best["delay"],
              "walker_sweeps_per_s": best["walker_sweeps_per_s"], "trials": trials}
    if cache_path:
        os.makedirs(os.path.dirname(cache_path) or ".", exist_ok=True)
        cache[key] = result
        tmp = cache_path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, cache_path)
    return result
