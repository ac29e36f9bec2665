"""Throughput benchmark of the Kagome VMC sampler on one MI355X.

Workload (default): the 12x12 pi-flux Kagome cluster (432 sites, 216 up /
216 down, PBC with antiperiodic a2 so the shell is closed), 8192 walkers
from seeded uniform Sz=0 starts, synthetic (no data files).  One "step" is
N = 216 sweeps of every walker (one reference sweep = one spin-exchange
proposal) followed by one local-energy measurement of every walker, i.e. one
measurement interval of the reference cadence.

Two forms run from identical seeds:
  fp32  W in FP32 with `delay` pending rank-1 terms per flush, FP64 local
        energy, FP64 LU refresh every k accepted moves  -> `value`
  fp64  W in FP64, immediate rank-1 update (the reference algorithm),
        FP64 refresh every 216 accepted moves
`value` is the FP32 form's walker-sweeps/s; the FP64 form, the ratio and
|dE/site| between the forms are reported beside it.

Prints one JSON line.  `--gpus N` with N > 1 exits: this build runs on one
GPU.
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "kagomedsl.jl_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# closed-shell boundary conditions per cluster size
BC = {4: ((True, True), (False, True)), 6: ((True, True), (False, False)),
      8: ((True, True), (False, True)), 10: ((True, True), (False, False)),
      12: ((True, True), (False, True))}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def sweep_bytes(cnt, s, ns, nup, ndn, nw):
    """Algorithmic HBM bytes of the sweep kernel for one kvmc_sweep call.

    flush:    read + write both W once          2 s ns (nup + ndn)
    accepted: column l + row K of both W        s (2 ns + nup + ndn)
    move:     two W gathers for the ratio       2 s
    launch:   kappa load + store per walker     4 * 4 ns"""
    return (cnt["flushes"] * 2 * s * ns * (nup + ndn) + cnt["accepted"] * s * (2 * ns + nup + ndn)
            + cnt["moves"] * 2 * s + cnt["launches"] * nw * 16 * ns)


def refresh_flops(ns, n):
    """FP64 flops of one spin's refresh in this kernel: elimination of the
    ns x n matrix [tildeU; U_empty] with pivots from the top n rows, then the
    unit-lower solve of the ns - n empty rows."""
    lu = sum((ns - k - 1) + 2 * (ns - k - 1) * (n - k - 1) for k in range(n))
    solve = (ns - n) * n * (n - 1)
    return lu + solve


def run_form(nat, ham, args, precision, delay, k, label):
    ns, n = ham.ns, ham.N_up
    nw = args.walkers
    S = args.sweeps_per_step or n
    ds = nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nw, precision=precision,
                           refresh_policy=nat.REFRESH_EVERY_K, refresh_k=k, delay=delay,
                           seed=args.seed)
    t0 = time.perf_counter()
    redraws = ds.init_random(args.seed)
    log(f"[{label}] init {nw} walkers L={args.L}: {time.perf_counter() - t0:.2f}s "
        f"(redraws {redraws})")
    for w in range(args.warmup):
        t0 = time.perf_counter()
        ds.sweep(S)
        ds.measure()
        log(f"[{label}] warmup step {w}: {time.perf_counter() - t0:.3f}s")
    sync()
    ds.profile(True)
    tot = dict(sweeps=0, moves=0, accepted=0, refreshes=0, flushes=0, launches=0)
    nap = 0
    energies = []
    sync()
    t0 = time.perf_counter()
    for st in range(args.steps):
        cnt, _ = ds.sweep(S)
        e, na = ds.measure()
        for key in tot:
            tot[key] += cnt[key]
        nap += na
        energies.append(e)
    sync()
    wall = time.perf_counter() - t0
    prof = ds.profile()
    drift = ds.w_drift() if args.drift else None
    ds.close()
    s = 8 if precision == nat.PREC_FP64 else 4
    b_sweep = sweep_bytes(tot, s, ns, n, ham.N_down, nw)
    b_eloc = s * 2 * nap + 8 * ns * nw * args.steps
    flops_ref = tot["refreshes"] * (refresh_flops(ns, n) + refresh_flops(ns, ham.N_down))
    res = {
        "walker_sweeps_per_s": nw * S * args.steps / wall,
        "ms_per_step": 1e3 * wall / args.steps,
        "eloc_evals_per_s_kernel": nw * args.steps / (prof["measure_ms"] * 1e-3),
        "accept_rate": tot["accepted"] / tot["sweeps"],
        "counters": tot,
        "kernel_ms": {k2: prof[k2] for k2 in ("sweep_ms", "refresh_ms", "measure_ms")},
        "launches": {k2: prof[k2] for k2 in ("sweep_launches", "refresh_launches", "measure_launches")},
        "sweep_bytes": b_sweep,
        "sweep_gbs": b_sweep / (prof["sweep_ms"] * 1e-3) / 1e9,
        "rank1_equiv_gbs": tot["accepted"] * 2 * s * ns * (n + ham.N_down) / (prof["sweep_ms"] * 1e-3) / 1e9,
        "eloc_gbs": b_eloc / (prof["measure_ms"] * 1e-3) / 1e9,
        "refresh_tflops": (flops_ref / (prof["refresh_ms"] * 1e-3) / 1e12) if prof["refresh_ms"] > 0 else None,
        "energy_per_site": float(np.mean(energies) / ns),
        "w_drift": drift,
    }
    log(f"[{label}] {res['walker_sweeps_per_s']:.3e} walker-sweeps/s, acc {res['accept_rate']:.3f}, "
        f"sweep {res['sweep_gbs']:.0f} GB/s, kernel ms {res['kernel_ms']}")
    return res, np.array(energies)


def cpu_baseline(ham, seconds, threads, seed):
    """The oracle (C restatement of the reference algorithm: complex128 W,
    immediate rank-1 updates, LU+inverse+GEMM refresh every N sweeps), one
    chain per host thread, timed for a bounded wall budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    oracle.build()
    counts = [0] * threads
    t_end = [0.0]

    def work(t):
        ku, kd = oracle.random_config(seed, 100000 + t, 0, ham.ns, ham.N_up, ham.N_down)
        ch = oracle.Chain.start(ham.nn, ham.U_up, ham.U_down, ku, kd)
        chunk = ham.N_up
        while time.perf_counter() < t_end[0]:
            oracle.run_chain(ch, chunk, seed=seed, walker=100000 + t)
            counts[t] += chunk

    t0 = time.perf_counter()
    t_end[0] = t0 + seconds
    ths = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()
    wall = time.perf_counter() - t0
    return sum(counts) / wall, sum(counts), wall


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--L", type=int, default=12)
    ap.add_argument("--walkers", type=int, default=8192)
    ap.add_argument("--sweeps-per-step", type=int, default=0)
    ap.add_argument("--delay", type=int, default=8)
    ap.add_argument("--k", type=int, default=64, help="FP32 form: accepted moves per refresh")
    ap.add_argument("--k64", type=int, default=216, help="FP64 form: accepted moves per refresh")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--drift", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); walkers are "
              "independent chains, so N GPUs would each run this same single-GPU job.",
              file=sys.stderr)
        sys.exit(2)

    from kagome_vmc import DoubleKagome, Hamiltonian
    from kagome_vmc import _native as nat
    nat.load()
    L = args.L
    pbc, apbc = BC[L]
    lat = DoubleKagome.build(1.0, L, L, pbc, apbc)
    n = 3 * L * L // 2
    ham = Hamiltonian.build(n, n, lat)
    log(f"L={L} ns={ham.ns} nb={len(ham.nn)} walkers={args.walkers}")

    r32, e32 = run_form(nat, ham, args, nat.PREC_FP32, args.delay, args.k, "fp32")
    r64 = e64 = None
    if not args.skip_fp64:
        r64, e64 = run_form(nat, ham, args, nat.PREC_FP64, 1, args.k64, "fp64")

    cpu = None
    if args.cpu_seconds > 0:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        rate, nsw, wall = cpu_baseline(ham, args.cpu_seconds, threads, args.seed)
        cpu = {"value": rate, "unit": "walker-sweeps/s", "cores": threads, "kind": "port",
               "sample": f"oracle C restatement (complex128 W, reference kernel and cadence), "
                         f"L={L}, {threads} chains x {nsw // threads} sweeps in {wall:.1f}s"}
        log(f"[cpu] {rate:.3e} walker-sweeps/s on {threads} threads")

    value = r32["walker_sweeps_per_s"]
    line = {
        "metric": "walker-sweeps/s at L=12 (FP32-update vs FP64) at |dE/site|<=1e-5; E_loc evals/s",
        "value": value,
        "unit": "walker-sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"pi-flux Kagome L={L} ({ham.ns} sites, {n}/{n}), {args.walkers} walkers, "
                        f"step = {args.sweeps_per_step or n} sweeps + 1 E_loc; FP32 W, delay {args.delay}, "
                        f"FP64 refresh every {args.k} accepts",
            "L": L, "walkers": args.walkers, "delay": args.delay, "refresh_k": args.k,
        },
        "roofline": {
            "bound": "hbm", "achieved": r32["sweep_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": r32["sweep_gbs"] / HBM_PEAK_GBS, "traffic": None,
            "kernel": "k_sweep (FP32, delayed rank-1 updates)",
        },
        "cpu_baseline": cpu,
        "fp32": r32,
        "fp64": r64,
    }
    if r64 is not None:
        line["fp64_walker_sweeps_per_s"] = r64["walker_sweeps_per_s"]
        line["speedup_fp32_vs_fp64"] = value / r64["walker_sweeps_per_s"]
        line["dE_per_site_fp32_vs_fp64"] = abs(float(np.mean(e32) - np.mean(e64))) / ham.ns
    line["eloc_evals_per_s"] = r32["eloc_evals_per_s_kernel"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
