"""Throughput benchmark of the Kagome VMC sampler on one MI355X.

Workload (default): the 12x12 pi-flux Kagome cluster (432 sites, 216 up /
216 down, PBC with antiperiodic a2 so the shell is closed), 8192 walkers
from seeded uniform Sz=0 starts, synthetic (no data files).  One "step" is
N = 216 sweeps of every walker (one reference sweep = one spin-exchange
proposal) followed by one local-energy measurement of every walker, i.e. one
measurement interval of the reference cadence.

Forms (each its own handle):
  fp32       W in FP32 with `delay` pending rank-1 terms per flush, FP64 local
             energy, FP64 LU refresh every k accepted moves (EVERY_K: a walker
             is refreshed as soon as it reaches k)            -> `value`
  fp64       the plain FP64 sampler: W in FP64, immediate rank-1 updates,
             FP64 refresh every N accepted moves (EVERY_K, k = N)
  fp64_ref   FP64 with the reference's refresh cadence (src/MonteCarlo.jl:
             593-604: every min(N_up, N_dn) sweeps that pass the early returns)
  fp32 k=... the FP32 form at the other refresh intervals

Accuracy: fp32 and fp64 start from the same configurations (drawn once),
use the same Philox key, the same refresh interval and the same refresh
phases, so their chains take identical decisions except at near ties.  After
`--therm` thermalisation steps and the warm-up steps, every timed and probe
step pairs each walker's FP32 local energy with its FP64 one; the line
reports the mean |dE/site| with its standard error over walkers, the
fraction of walkers whose configurations left the FP64 chain's, and the
relative W drift just before every refresh of the FP32 form (probe steps).

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

# HBM bytes (FETCH_SIZE + WRITE_SIZE, calibrated against a plain copy) per
# walker refreshed by the kernel pair k_refresh_blocked + k_refresh_solve, from
# rocprofv3 --pmc passes over this script's default FP32 run (summaries in
# profiles/).
REFRESH_TRAFFIC_BYTES_PER_WALKER = {12: 6_157_947}
# the same for the FP32 sweep kernel (k_sweep_ib), per walker-sweep (delay 12,
# L=12, refresh every 216 accepts)
SWEEP_TRAFFIC_BYTES_PER_WALKER_SWEEP = {12: 10_394}

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
    """Algorithmic HBM bytes of the sweep kernel for one kvmc_sweep call.  W
    holds the rows of the empty sites only (ns - N per spin; the occupied
    rows are unit rows), plus a site -> row map per spin.

    flush:    read + write both W once          2 s ((ns - nup) nup + (ns - ndn) ndn)
    accepted: column l + row K of both W        s ((ns - nup) + (ns - ndn) + nup + ndn) = 2 s ns
    move:     two W gathers for the ratio       2 s
    launch:   kappa and row maps, load + store  2 * 2 * 2 * 4 ns"""
    rows = (ns - nup) * nup + (ns - ndn) * ndn
    return (cnt["flushes"] * 2 * s * rows + cnt["accepted"] * 2 * s * ns
            + cnt["moves"] * 2 * s + cnt["launches"] * nw * 32 * ns)


def refresh_flops(ns, n):
    """Algorithmic FP64 flops of one spin's refresh W = U tildeU^-1: pivoted
    LU of the n x n tildeU (2/3 n^3) and the two triangular solves for the ns
    rows of W (2 ns n^2).  The kernel itself skips the occupied rows of W
    (they are unit rows), so it executes fewer flops than this count."""
    return 2.0 * n**3 / 3.0 + 2.0 * ns * n * n


def refresh_min_bytes(ns, n, s):
    """Algorithmic HBM bytes of one spin's refresh: read U (FP64, ns x n) and
    kappa, write the ns - n empty-site rows of W (s bytes per element) and
    the site -> row map."""
    return 8 * ns * n + 4 * ns + s * (ns - n) * n + 4 * ns


def make_sampler(nat, ham, args, precision, delay, k, policy_name):
    """One form's handle: k <= 0 selects the reference refresh cadence."""
    policy = (nat.REFRESH_REFERENCE if k <= 0 else
              {"every_k": nat.REFRESH_EVERY_K, "at_end": nat.REFRESH_AT_END}[policy_name])
    return nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, args.walkers, precision=precision,
                             refresh_policy=policy, refresh_k=max(k, 0), delay=delay,
                             threads=args.threads, seed=args.seed)


def prepare(ds, args, label, configs=None):
    """Seeded uniform starts (or the given configurations), then refresh
    phases spread as in a long run, so that any timed window holds the
    steady-state refresh rate of an every-k policy."""
    t0 = time.perf_counter()
    if configs is None:
        redraws = ds.init_random(args.seed)
    else:
        ds.set_configs(*configs)
        redraws = 0
    ds.stagger_refresh(args.seed)
    if args.graph:
        ds.enable_graph(True)
    log(f"[{label}] start {args.walkers} walkers L={args.L}: {time.perf_counter() - t0:.2f}s "
        f"(redraws {redraws})")
    return redraws


def untimed_steps(ds, S, nsteps, label, what):
    for w in range(nsteps):
        t0 = time.perf_counter()
        ds.sweep(S)
        ds.measure()
        if w == nsteps - 1 or w < 2:
            log(f"[{label}] {what} step {w}: {time.perf_counter() - t0:.3f}s")


def timed_steps(nat, ham, ds, args, precision, label):
    """K timed steps (S sweeps + one local-energy pass of every walker),
    bracketed by device synchronisation; kernel times from the handle's own
    HIP events.  Returns the summary and the per-step local energies."""
    ns, n = ham.ns, ham.N_up
    nw = args.walkers
    S = args.sweeps_per_step or n
    sync()
    # graph replay needs profiling off: kernel times are then not split
    ds.profile(not args.graph)
    tot = dict(sweeps=0, moves=0, accepted=0, refreshes=0, flushes=0, launches=0)
    nap = 0
    energies = []
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        cnt, _ = ds.sweep(S)
        e, na = ds.measure()
        for key in tot:
            tot[key] += cnt[key]
        nap += na
        energies.append(e)
    sync()
    wall = time.perf_counter() - t0
    prof = ds.profile()
    ds.profile(False)
    s = 8 if precision == nat.PREC_FP64 else 4
    b_sweep = sweep_bytes(tot, s, ns, n, ham.N_down, nw)
    b_eloc = s * 2 * nap + 8 * ns * nw * args.steps
    flops_walker = refresh_flops(ns, n) + refresh_flops(ns, ham.N_down)
    flops_ref = tot["refreshes"] * flops_walker

    def rate(num, ms):
        return num / (ms * 1e-3) if ms > 0 else None

    res = {
        "walker_sweeps_per_s": nw * S * args.steps / wall,
        "ms_per_step": 1e3 * wall / args.steps,
        "graph_replay": bool(args.graph),
        "eloc_evals_per_s_kernel": rate(nw * args.steps, prof["measure_ms"]),
        "accept_rate": tot["accepted"] / tot["sweeps"],
        "counters": tot,
        "kernel_ms": {k2: prof[k2] for k2 in ("sweep_ms", "refresh_ms", "measure_ms")},
        "launches": {k2: prof[k2] for k2 in ("sweep_launches", "refresh_launches", "measure_launches")},
        "sweep_bytes": b_sweep,
        "sweep_gbs": (rate(b_sweep, prof["sweep_ms"]) or 0) / 1e9,
        "rank1_equiv_gbs": (rate(tot["accepted"] * 2 * s * ns * (n + ham.N_down), prof["sweep_ms"]) or 0) / 1e9,
        "eloc_gbs": (rate(b_eloc, prof["measure_ms"]) or 0) / 1e9,
        "refresh_tflops": (rate(flops_ref, prof["refresh_ms"]) or 0) / 1e12 or None,
        "refresh_flops_per_walker": flops_walker,
        "refresh_walkers_per_launch": (tot["refreshes"] / prof["refresh_launches"]) if prof["refresh_launches"] else 0,
        "refresh_ms_per_launch": (prof["refresh_ms"] / prof["refresh_launches"]) if prof["refresh_launches"] else 0,
        "sweep_ms_per_launch": prof["sweep_ms"] / max(prof["sweep_launches"], 1),
        "sweep_bytes_per_launch": b_sweep / max(prof["sweep_launches"], 1),
        "walker_sweeps_per_launch": nw * S * args.steps / max(prof["sweep_launches"], 1),
        "accepts_per_launch": tot["accepted"] / max(prof["sweep_launches"], 1),
        "energy_per_site": float(np.mean(energies) / ns),
    }
    log(f"[{label}] {res['walker_sweeps_per_s']:.3e} walker-sweeps/s, acc {res['accept_rate']:.3f}, "
        f"sweep {res['sweep_gbs']:.0f} GB/s, kernel ms {res['kernel_ms']}, refreshes {tot['refreshes']}")
    return res, np.array(energies)


def run_form(nat, ham, args, precision, delay, k, policy_name, label, warmup=None):
    """A form timed on its own (own starts): thermalisation-free warm-up,
    then the timed steps."""
    ds = make_sampler(nat, ham, args, precision, delay, k, policy_name)
    prepare(ds, args, label)
    untimed_steps(ds, args.sweeps_per_step or ham.N_up, args.warmup if warmup is None else warmup,
                  label, "warmup")
    res, e = timed_steps(nat, ham, ds, args, precision, label)
    ds.close()
    return res, e


def paired_accuracy(nat, ham, args, ds32, ds64, e32, e64):
    """Paired FP32-vs-FP64 local energies over the timed steps (e32, e64:
    steps x walkers) and `probe_steps` further steps, in which the FP32
    handle also records its relative W drift just before every refresh and
    the carried FP32 W is checked against a fresh FP64 refresh of the same
    configurations (a third handle)."""
    ns = ham.ns
    S = args.sweeps_per_step or ham.N_up
    d_steps = [(a - b) / ns for a, b in zip(e32, e64)]
    diverged = np.zeros(args.walkers, dtype=bool)
    carried_mean = carried_max = 0.0
    ds32.track_drift(True)
    ref = nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, args.walkers, precision=nat.PREC_FP64)
    for st in range(args.probe_steps):
        ds32.sweep(S)
        a, _ = ds32.measure()
        ds64.sweep(S)
        b, _ = ds64.measure()
        d_steps.append((a - b) / ns)
        k32, k64 = ds32.get_configs(), ds64.get_configs()
        diverged |= np.any(k32[0] != k64[0], axis=1) | np.any(k32[1] != k64[1], axis=1)
        ref.set_configs(*k32)
        e_fresh, _ = ref.measure()
        carried_mean = max(carried_mean, abs(float(np.mean(a) - np.mean(e_fresh))) / ns)
        carried_max = max(carried_max, float(np.abs(a - e_fresh).max()) / ns)
        log(f"[accuracy] probe step {st}: mean dE/site {np.mean(d_steps[-1]):+.2e}, "
            f"diverged walkers {int(diverged.sum())}")
    ref.close()
    dmax, dmean, dcount = ds32.drift_stats()
    ds32.track_drift(False)
    D = np.stack(d_steps)                    # steps x walkers, per site
    per_walker = D.mean(axis=0)
    mean = float(per_walker.mean())
    stderr = float(per_walker.std(ddof=1) / np.sqrt(len(per_walker)))
    same = ~diverged
    out = {
        "dE_per_site_mean": mean,
        "dE_per_site_stderr": stderr,
        "abs_dE_per_site": abs(mean),
        # the bound as stated (|mean| <= 1e-5 with a standard error below it),
        # and the mean's upper two-sigma limit beside it
        "within_1e-5": bool(abs(mean) <= 1e-5 and stderr < 1e-5),
        "abs_dE_per_site_upper_2sigma": abs(mean) + 2 * stderr,
        "steps_paired": int(D.shape[0]),
        "walkers": int(D.shape[1]),
        "diverged_fraction": float(diverged.mean()),
        "identical_walkers_dE_per_site_mean": float(per_walker[same].mean()) if same.any() else None,
        "identical_walkers_max_abs_dE_per_site": float(np.abs(D[:, same]).max()) if same.any() else None,
        "w_drift_before_refresh_max": dmax,
        "w_drift_before_refresh_mean": dmean,
        "w_drift_refreshes": int(dcount),
        "carried_vs_refreshed_dE_per_site": carried_mean,
        "carried_vs_refreshed_max_walker_dE_per_site": carried_max,
    }
    log(f"[accuracy] dE/site {mean:+.2e} +- {stderr:.1e} over {D.shape[1]} walkers x {D.shape[0]} steps; "
        f"diverged {out['diverged_fraction']:.4f}; W drift before refresh max {dmax:.2e} "
        f"mean {dmean:.2e} over {dcount} refreshes")
    return out


def host_cpu():
    """Logical CPUs of the host, the ones this process may use, and the model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return os.cpu_count() or 1, usable, model


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


def copy_calibration(mib):
    """A plain device copy of `mib` MiB (torch's copy kernel), run a few times
    so that rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE can be read against a
    known byte count."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "launches": 3}), flush=True)


def roofline_refresh(r, ns, traffic_per_walker):
    wpl = r["refresh_walkers_per_launch"]
    if not wpl or not r["refresh_ms_per_launch"]:
        return None
    achieved = r["refresh_flops_per_walker"] * wpl / (r["refresh_ms_per_launch"] * 1e-3) / 1e12
    return {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS,
            "traffic": traffic_per_walker * wpl if traffic_per_walker else None,
            "kernel": "k_refresh_blocked + k_refresh_solve (FP64 pivoted LU of tildeU, then MFMA "
                      "triangular solves for the empty-site rows of W = U tildeU^-1)",
            "units_per_launch": wpl, "launch_ms": r["refresh_ms_per_launch"]}


def sweep_kernel_name(ham, s, delay):
    """The sweep kernel the library runs for this shape: FP32 W streamed from
    HBM with at most 256 empty-site rows per spin and 1 < delay <= 16 uses
    the implicit-pending-row kernel, everything else k_sweep."""
    rows = [((ham.ns - n) + 3) // 4 * 4 for n in (ham.N_up, ham.N_down)]
    resident = s * sum(r * n for r, n in zip(rows, (ham.N_up, ham.N_down))) <= 64 * 1024
    if s == 4 and 1 < delay <= 16 and max(rows) <= 256 and not resident:
        return "k_sweep_ib (exchange step + delayed Sherman-Morrison updates, pending rows implicit)"
    return "k_sweep (exchange step + delayed Sherman-Morrison updates)"


def roofline_sweep(r, traffic_per_walker_sweep, ham, s=4, delay=12):
    """HBM roofline of the sweep kernel.  `achieved` counts the bytes the
    algorithm as built must move (flushes of the empty-site rows of W once
    per `delay` accepts, the column and row of each accept, two gathers per
    move; sweep_bytes).  Beside it, the per-accept unit (a full read and
    write of both ns x N W per accepted move, 2 s ns^2, plus 2 s + 16 per
    proposal): the work an immediate rank-1 update of the full W would have
    to stream, delivered in the same time."""
    if not r["sweep_ms_per_launch"]:
        return None
    t = r["sweep_ms_per_launch"] * 1e-3
    achieved = r["sweep_bytes_per_launch"] / t / 1e9
    wsl = r["walker_sweeps_per_launch"]
    per_accept = (r["accepts_per_launch"] * 2 * s * ham.ns * ham.ns + wsl * (2 * s + 16)) / t / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic_per_walker_sweep * wsl if traffic_per_walker_sweep else None,
            "units_per_launch": wsl,
            "kernel": sweep_kernel_name(ham, s, delay),
            "launch_ms": r["sweep_ms_per_launch"],
            "achieved_per_accept_unit": per_accept,
            "frac_per_accept_unit": per_accept / HBM_PEAK_GBS}


MIXED_GROUPS = ((4, 4096), (6, 2048), (8, 2048), (10, 1024), (12, 1024))


def run_mixed(args):
    """All groups of MIXED_GROUPS (L, walkers) at once, one handle and host
    thread per group, FP32 W with delayed updates and FP64 refresh every `k`
    accepts; each group's workgroup size and delay come from the autotuner.
    Each lattice takes a flux phase phi drawn uniformly from [0, pi] with the
    seed (complex orbitals, complex W).  One step = n_occupied sweeps + one
    local-energy pass in every group.  Also times each group alone, for the
    overlap gain and for each group's kernel times and HBM roofline."""
    from kagome_vmc import _native as nat
    from kagome_vmc.autotune import autotune
    from kagome_vmc.batch import MixedBatch, flux_cluster
    nat.load()
    rng = np.random.default_rng(args.seed)
    phis = [float(rng.uniform(0.0, np.pi)) for _ in MIXED_GROUPS]
    hams = [flux_cluster(L, phi) for (L, _), phi in zip(MIXED_GROUPS, phis)]
    # launch shape per group (workgroup size and delay) from the autotuner,
    # each candidate timed on the group's own workload; --no-autotune keeps
    # --delay and the library's default workgroup size for every group
    shapes = []
    for ham, (_, nw) in zip(hams, MIXED_GROUPS):
        if args.no_autotune:
            shapes.append({"threads": 0, "delay": args.delay})
        else:
            r = autotune(ham, nw, nat.PREC_FP32, args.k, seed=args.seed, steps=2, cache_path=None)
            shapes.append({"threads": r["threads"], "delay": r["delay"]})
        log(f"[mixed] L={int(round((ham.ns / 3) ** 0.5))}: threads {shapes[-1]['threads']}, "
            f"delay {shapes[-1]['delay']}")
    groups = [(ham, nw, shp) for ham, (_, nw), shp in zip(hams, MIXED_GROUPS, shapes)]
    kw = dict(precision=nat.PREC_FP32, refresh_policy=nat.REFRESH_EVERY_K, refresh_k=args.k, seed=args.seed)

    def timed(batch, profile=False):
        batch.init_random(args.seed)
        for g in batch.groups:
            g.sampler.stagger_refresh(args.seed)
        batch.step(args.warmup)
        sync()
        if profile:
            for g in batch.groups:
                g.sampler.profile(True)
        t0 = time.perf_counter()
        res = batch.step(args.steps)
        sync()
        wall = time.perf_counter() - t0
        prof = [g.sampler.profile() for g in batch.groups] if profile else None
        return wall, res, prof

    # all groups in one graph launch per step (kvmc_group_sweep); with
    # --threads-mixed one host thread per group instead
    with MixedBatch(groups, graph=not args.threads_mixed, **kw) as mb:
        wall, res, _ = timed(mb)
    with MixedBatch(groups, **kw) as mb:
        wall_threads, res_threads, _ = timed(mb)
    per = []
    alone_total = 0.0
    for (ham, nw, shp), (cnt, _), phi in zip(groups, res, phis):
        with MixedBatch([(ham, nw, shp)], **kw) as one:
            t1, r1, prof = timed(one, profile=True)
        alone_total += t1
        c1 = r1[0][0]
        s = 8   # complex float W elements
        b = sweep_bytes(c1, s, ham.ns, ham.N_up, ham.N_down, nw)
        sweep_gbs = b / (prof[0]["sweep_ms"] * 1e-3) / 1e9 if prof[0]["sweep_ms"] > 0 else None
        per.append({"L": int(round((ham.ns / 3) ** 0.5)), "walkers": nw, "phi": phi,
                    "threads": shp["threads"], "delay": shp["delay"],
                    "walker_sweeps": cnt["sweeps"],
                    "alone_walker_sweeps_per_s": c1["sweeps"] / t1, "refreshes": cnt["refreshes"],
                    # both W of a walker within 64 KiB stay in LDS (empty-site rows only)
                    "resident_w_in_lds": ((ham.ns - ham.N_up + 3) // 4 * 4) * (ham.N_up + ham.N_down) * s
                                         <= 64 * 1024,
                    "alone_kernel_ms": {k2: prof[0][k2] for k2 in ("sweep_ms", "refresh_ms", "measure_ms")},
                    "alone_sweep_gbs": sweep_gbs,
                    "alone_sweep_hbm_frac": (sweep_gbs / HBM_PEAK_GBS) if sweep_gbs else None})
        log(f"[mixed] L={per[-1]['L']} phi={phi:.3f}: alone {per[-1]['alone_walker_sweeps_per_s']:.3e} "
            f"walker-sweeps/s, sweep kernel {sweep_gbs or 0:.0f} GB/s")
    total = sum(c["sweeps"] for c, _ in res)
    line = {
        "metric": "walker-sweeps/s, mixed cluster sizes L in {4,6,8,10,12} on one GPU",
        "value": total / wall, "unit": "walker-sweeps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "mixed: " + ", ".join(f"L={L} x {nw}" for L, nw in MIXED_GROUPS)
                               + f"; seeded flux phase phi in [0, pi] per lattice (complex W); FP32 W with "
                               + ("the autotuner's workgroup size and delay per group" if not args.no_autotune
                                  else f"delay {args.delay}")
                               + f", FP64 refresh every {args.k} accepts",
                   "groups": per},
        "sequential_walker_sweeps_per_s": total / alone_total,
        "launch": "one graph launch per step for all groups" if not args.threads_mixed else "one host thread per group",
        "threads_walker_sweeps_per_s": sum(c["sweeps"] for c, _ in res_threads) / wall_threads,
    }
    print(json.dumps(line), flush=True)


def run_grad(args):
    """Gradient throughput: walkers from |psi|^2 (unbiased kernel, FP64 W) on
    the ansatz family at theta = (phi, ratio, m, t2), drawn uniformly from
    [0, pi] x [0.8, 1.2] x [0, 0.2] x [-0.1, 0.1] with the seed (complex
    orbitals unless phi = 0), or theta0 = (0, 1, 0, 0) with --theta theta0.
    One step = N sweeps + one local-energy and log-derivative pass; reports
    measure_grad evaluations/s, the sampled <O_k>, <E_L conj(O_k)>, the SR
    matrix S and the gradient with its standard error over walkers."""
    from kagome_vmc import _native as nat
    from kagome_vmc.gradient import THETA0, GradientSampler, ansatz_family, averages
    from kagome_vmc.hamiltonian import shell_gap
    from kagome_vmc.lattice import DoubleKagome
    nat.load()
    L = args.L
    pbc, apbc = BC[L]
    fam = ansatz_family(DoubleKagome.build(1.0, L, L, pbc, apbc))
    n = 3 * L * L // 2
    if args.theta == "seed":
        rng = np.random.default_rng(args.seed)
        theta = np.array([rng.uniform(0.0, np.pi), rng.uniform(0.8, 1.2), rng.uniform(0.0, 0.2),
                          rng.uniform(-0.1, 0.1)])
    else:
        theta = THETA0.copy()
    gap = shell_gap(np.linalg.eigvalsh(fam.hamiltonian(theta)), n)
    gs = GradientSampler(fam, theta, n, n, args.walkers, seed=args.seed, precision=nat.PREC_FP64,
                         delay=args.grad_delay)
    s = gs.sampler
    s.init_random(args.seed)
    for _ in range(args.warmup):     # warm-up steps do the timed steps' work (first launches included)
        s.sweep(n)
        s.measure_grad()
    sync()
    t_grad = 0.0
    E, O = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.sweep(n)
        t1 = time.perf_counter()
        e, o = s.measure_grad()
        t_grad += time.perf_counter() - t1
        E.append(e)
        O.append(o)
    wall = time.perf_counter() - t0
    is_cplx = bool(gs.sampler.complex)
    gs.close()
    E, O = np.stack(E), np.stack(O)
    # complex log-derivatives (the flux direction): dE = 2 Re(<O* E_L> - E <O*>)
    w_e = E.real.mean(0)
    Oc = np.conj(O)
    w_g = 2.0 * ((E[..., None] * Oc).mean(0) - E.mean(0)[:, None] * Oc.mean(0)).real
    avg = averages(E.ravel(), O.reshape(-1, O.shape[-1]))
    cpair = lambda z: [[float(np.real(v)), float(np.imag(v))] for v in np.ravel(z)]
    line = {
        "metric": f"gradient evaluations/s (E_L and 4 log-derivatives per walker), L={L}",
        "value": args.walkers * args.steps / t_grad, "unit": "walker-evaluations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Kagome L={L} ({3 * L * L} sites), {args.walkers} walkers from |psi|^2 at "
                               f"theta = (phi, ratio, m, t2) = ({', '.join(f'{v:.6g}' for v in theta)}) "
                               f"({'drawn with the seed' if args.theta == 'seed' else 'theta0'}); "
                               f"step = {n} sweeps + measure_grad",
                   "theta": theta.tolist(), "complex_orbitals": is_cplx,
                   "fermi_shell_gap": gap},
        "energy_per_site": float(w_e.mean() / (3 * L * L)),
        "O_mean_re_im": cpair(avg["O"]),
        "E_conjO_mean_re_im": cpair(avg["EO"]),
        "S": avg["S"].tolist(),
        "gradient_per_site": (w_g.mean(0) / (3 * L * L)).tolist(),
        "gradient_stderr_per_site": (w_g.std(0, ddof=1) / np.sqrt(args.walkers) / (3 * L * L)).tolist(),
        "walker_sweeps_per_s": args.walkers * n * args.steps / wall,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--L", type=int, default=12)
    ap.add_argument("--walkers", type=int, default=8192)
    ap.add_argument("--sweeps-per-step", type=int, default=0)
    ap.add_argument("--delay", type=int, default=12)
    ap.add_argument("--threads", type=int, default=0, help="sweep-kernel workgroup size (0: automatic)")
    ap.add_argument("--k", type=int, default=216, help="FP32 form: accepted moves per refresh")
    ap.add_argument("--k64", type=int, default=0,
                    help="plain FP64 form: accepted moves per refresh (0: N, one per electron)")
    ap.add_argument("--policy", default="every_k", choices=("every_k", "at_end"),
                    help="FP32 headline: refresh mid-call (every_k) or after the call (at_end)")
    ap.add_argument("--k-extra", default="16,64,at_end:16,at_end:216",
                    help="further FP32 forms beside the headline: k (every-k policy) or at_end:k")
    ap.add_argument("--therm", type=int, default=20,
                    help="thermalisation steps of the paired fp32/fp64 forms before the warm-up")
    ap.add_argument("--skip-fp64-ref", action="store_true",
                    help="skip the FP64 form with the reference refresh cadence")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--probe-steps", type=int, default=14,
                    help="untimed steps after the timed region, each checking the carried W")
    ap.add_argument("--graph", action="store_true",
                    help="replay captured sweep sequences (no per-kernel times)")
    ap.add_argument("--grad-delay", type=int, default=1,
                    help="gradient form: pending FP64 rank-1 updates per flush (1: immediate)")
    ap.add_argument("--theta", default="seed", choices=("seed", "theta0"),
                    help="gradient form: parameters drawn with the seed, or theta0 = (0, 1, 0, 0)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="mixed batch: --delay and the default workgroup size for every group")
    ap.add_argument("--grad", action="store_true",
                    help="only run the gradient workload (unbiased kernel, FP64, log-derivatives)")
    ap.add_argument("--mixed", action="store_true",
                    help="only run the mixed-cluster-size batch (L = 4..12 groups at once)")
    ap.add_argument("--threads-mixed", action="store_true",
                    help="mixed batch: one host thread per group instead of one graph launch")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); walkers are "
              "independent chains, so N GPUs would each run this same single-GPU job.",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.mixed:
        run_mixed(args)
        return
    if args.grad:
        run_grad(args)
        return

    from kagome_vmc import DoubleKagome, Hamiltonian
    from kagome_vmc import _native as nat
    nat.load()
    L = args.L
    pbc, apbc = BC[L]
    lat = DoubleKagome.build(1.0, L, L, pbc, apbc)
    n = 3 * L * L // 2
    ham = Hamiltonian.build(n, n, lat)
    log(f"L={L} ns={ham.ns} nb={len(ham.nn)} walkers={args.walkers}")

    S = args.sweeps_per_step or n
    k64 = args.k64 or n
    # headline pair: FP32 and plain FP64 from the same configurations, Philox
    # key, refresh interval and refresh phases
    ds32 = make_sampler(nat, ham, args, nat.PREC_FP32, args.delay, args.k, args.policy)
    prepare(ds32, args, "fp32")
    ds64 = None
    if not args.skip_fp64:
        ds64 = make_sampler(nat, ham, args, nat.PREC_FP64, 1, k64, "every_k")
        prepare(ds64, args, "fp64", configs=ds32.get_configs())
    for ds, label in ((ds32, "fp32"), (ds64, "fp64")):
        if ds is not None:
            untimed_steps(ds, S, args.therm + args.warmup, label, "thermalisation/warm-up")
    r32, e32 = timed_steps(nat, ham, ds32, args, nat.PREC_FP32, "fp32")
    r64 = e64 = acc = None
    if ds64 is not None:
        r64, e64 = timed_steps(nat, ham, ds64, args, nat.PREC_FP64, "fp64")
        acc = paired_accuracy(nat, ham, args, ds32, ds64, e32, e64)
        ds64.close()
    ds32.close()
    r64ref = None
    if not args.skip_fp64 and not args.skip_fp64_ref:
        r64ref, _ = run_form(nat, ham, args, nat.PREC_FP64, 1, 0, "every_k", "fp64 reference cadence")
    extra = {}
    for item in [x.strip() for x in args.k_extra.split(",") if x.strip()]:
        pol, _, kk = item.rpartition(":")
        pol = pol or "every_k"
        k = int(kk)
        rk, _ = run_form(nat, ham, args, nat.PREC_FP32, args.delay, k, pol, f"fp32 {pol} k={k}")
        extra[f"{pol}:{k}"] = {key: rk[key] for key in (
            "walker_sweeps_per_s", "ms_per_step", "kernel_ms", "refresh_tflops")}
        extra[f"{pol}:{k}"]["refreshes"] = rk["counters"]["refreshes"]
        if r64 is not None:
            extra[f"{pol}:{k}"]["speedup_vs_fp64"] = rk["walker_sweeps_per_s"] / r64["walker_sweeps_per_s"]

    cpu = None
    if args.cpu_seconds > 0:
        nproc, usable, model = host_cpu()
        # the GPU box allots 16 host threads to each GPU
        threads = args.cpu_threads or min(usable, 16)
        rate, nsw, wall = cpu_baseline(ham, args.cpu_seconds, threads, args.seed)
        cpu = {"value": rate, "unit": "walker-sweeps/s", "cores": threads, "kind": "port",
               "nproc": nproc, "usable_cpus": usable, "cpu_model": model,
               "per_core": rate / threads,
               "sample": f"oracle C restatement (complex128 W, reference kernel and cadence), "
                         f"L={L}, {threads} independent chains (one per thread) x {nsw // threads} sweeps "
                         f"in {wall:.1f}s"}
        log(f"[cpu] {rate:.3e} walker-sweeps/s on {threads} threads ({nproc} logical CPUs, {model})")

    # the dominant kernel of the headline form carries the roofline
    traffic_ok = args.delay == 12 and args.policy == "every_k" and args.k == 216
    rr = roofline_refresh(r32, ham.ns, REFRESH_TRAFFIC_BYTES_PER_WALKER.get(L) if traffic_ok else None)
    rs = roofline_sweep(r32, SWEEP_TRAFFIC_BYTES_PER_WALKER_SWEEP.get(L) if traffic_ok else None, ham,
                        delay=args.delay)
    dominant = rr if rr and r32["kernel_ms"]["refresh_ms"] > r32["kernel_ms"]["sweep_ms"] else rs
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
        "dtype_detail": "W and determinant ratios in f32 (delayed rank-1 updates); local energy, "
                        "refresh LU and solve in f64",
        "data": "synthetic",
        "config": {
            "workload": f"pi-flux Kagome L={L} ({ham.ns} sites, {n}/{n}), {args.walkers} walkers, "
                        f"step = {S} sweeps + 1 E_loc; FP32 W, delay {args.delay}, "
                        f"FP64 refresh every {args.k} accepts ({args.policy}); {args.therm} thermalisation "
                        f"steps before the warm-up",
            "L": L, "walkers": args.walkers, "delay": args.delay, "refresh_k": args.k,
            "refresh_policy": args.policy,
        },
        "roofline": dominant,
        "roofline_other": rs if dominant is rr else rr,
        "cpu_baseline": cpu,
        "accuracy_fp32_vs_fp64": acc,
        "fp32": r32,
        "fp64": r64,
        "fp64_reference_cadence": r64ref,
        "fp32_other": extra,
    }
    if r64 is not None:
        line["fp64_walker_sweeps_per_s"] = r64["walker_sweeps_per_s"]
        line["speedup_fp32_vs_fp64"] = value / r64["walker_sweeps_per_s"]
    if r64ref is not None:
        line["speedup_fp32_vs_fp64_reference_cadence"] = value / r64ref["walker_sweeps_per_s"]
    line["eloc_evals_per_s"] = r32["eloc_evals_per_s_kernel"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
