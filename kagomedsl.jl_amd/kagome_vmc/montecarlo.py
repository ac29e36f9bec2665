"""Host mirror of the reference's Carlo.jl plugin for KagomeDSL.MC.

The reference implements Carlo's ``AbstractMC`` hooks on a one-chain ``MC``
(src/MonteCarlo.jl:17-28, 165-185, 435-441, 538-607, 628-634, 675-687,
715-755).  Here ``MC`` holds a batch of walkers whose state lives on the GPU
(libkagome_vmc.so); every hook advances or measures all walkers at once.

Carlo's run loop (``sweep!``, then ``ctx.sweeps += 1``, then ``measure!`` once
thermalised) is reproduced by :func:`run`, which hands the device whole
stretches of sweeps between two measurement points.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat
from .hamiltonian import Hamiltonian, pi_link_in, pi_link_inter
from .lattice import DoubleKagome

_PREC = {"fp64": nat.PREC_FP64, "fp32": nat.PREC_FP32}
_KERNEL = {"reference": nat.KERNEL_REFERENCE, "unbiased": nat.KERNEL_UNBIASED}


@dataclass
class MCContext:
    """Carlo's MCContext: seed, sweep counter, thermalisation length and the
    measurement store (samples are kept only once thermalised)."""

    seed: int = 0
    thermalization: int = 0
    binsize: int = 1
    sweeps: int = 0
    observables: dict = field(default_factory=dict)

    @classmethod
    def from_params(cls, params):
        return cls(seed=int(params.get("seed", 0)),
                   thermalization=int(params.get("thermalization", 0)),
                   binsize=int(params.get("binsize", 1)))

    def is_thermalized(self) -> bool:
        return self.sweeps > self.thermalization

    def measure(self, name, value):
        if self.is_thermalized():
            self.observables.setdefault(name, []).append(np.asarray(value))


class MC:
    """Batched counterpart of ``MC(params)`` (src/MonteCarlo.jl:165-185).

    Reference keys: ``n1, n2, PBC, N_up, N_down`` (required), ``antiPBC, B,
    link_in, link_inter`` (optional).  Device keys: ``nwalkers`` (1),
    ``precision`` ("fp64" | "fp32"), ``kernel`` ("reference" | "unbiased"),
    ``refresh`` ("reference" or an int k: refresh every k accepted moves),
    ``delay`` (pending rank-1 updates per flush; 1 = immediate), ``threads``
    (either may be "auto": chosen by the autotuner), ``seed`` (device RNG
    key; defaults to the context seed)."""

    def __init__(self, params: dict):
        n1, n2 = int(params["n1"]), int(params["n2"])
        PBC = tuple(params["PBC"])
        antiPBC = tuple(params.get("antiPBC", (False, False)))
        self.lat = DoubleKagome.build(1.0, n1, n2, PBC, antiPBC)
        self.N_up, self.N_down = int(params["N_up"]), int(params["N_down"])
        self.Ham = Hamiltonian.build(self.N_up, self.N_down, self.lat,
                                     params.get("link_in", pi_link_in),
                                     params.get("link_inter", pi_link_inter),
                                     float(params.get("B", 0.0)), nn=params.get("nn"))
        self.ns = n1 * n2 * 3
        self.nwalkers = int(params.get("nwalkers", 1))
        refresh = params.get("refresh", "reference")
        policy = nat.REFRESH_REFERENCE if refresh == "reference" else nat.REFRESH_EVERY_K
        k = 0 if refresh == "reference" else int(refresh)
        prec = _PREC[params.get("precision", "fp64")]
        threads, delay = params.get("threads", 0), params.get("delay", 1)
        if threads == "auto" or delay == "auto":
            from .autotune import autotune
            tuned = autotune(self.Ham, self.nwalkers, prec, k or min(self.N_up, self.N_down),
                             seed=int(params.get("seed", 0)))
            threads = tuned["threads"] if threads == "auto" else threads
            delay = tuned["delay"] if delay == "auto" else delay
        self.sampler = nat.DeviceSampler(
            self.Ham.bonds_array(), self.Ham.U_up, self.Ham.U_down, self.nwalkers,
            precision=prec, kernel=_KERNEL[params.get("kernel", "reference")],
            refresh_policy=policy, refresh_k=k, delay=int(delay), threads=int(threads),
            seed=int(params.get("rng_seed", params.get("seed", 0))))
        self.counters = {"sweeps": 0, "moves": 0, "accepted": 0, "refreshes": 0}

    @property
    def n_occupied(self) -> int:
        return min(self.N_up, self.N_down)

    @property
    def kappa(self):
        """(kappa_up, kappa_down), each nwalkers x ns, 1-based labels."""
        return self.sampler.get_configs()

    def close(self):
        self.sampler.close()


def init_conf_qr(U_up: np.ndarray, U_down: np.ndarray):
    """init_conf_qr! (src/MonteCarlo.jl:326-357): up electrons on the first
    N_up pivot columns of a column-pivoted QR of U_up^T; down electrons on the
    first N_down pivots of the same factorisation of U_down^T restricted to
    the remaining sites (in increasing site order, as setdiff(1:ns, ...)).
    Labels follow the pivot order.  Returns (kappa_up, kappa_down), 1-based."""
    import scipy.linalg as sla
    U_up = np.real_if_close(np.asarray(U_up))
    U_down = np.real_if_close(np.asarray(U_down))
    ns, n_up = U_up.shape
    n_dn = U_down.shape[1]
    kup = np.zeros(ns, dtype=np.int32)
    kdn = np.zeros(ns, dtype=np.int32)
    _, _, p_up = sla.qr(U_up.T, mode="economic", pivoting=True)
    sites_up = p_up[:n_up]
    kup[sites_up] = np.arange(1, n_up + 1)
    if n_dn > 0:
        avail = np.setdiff1d(np.arange(ns), sites_up)
        _, _, p_dn = sla.qr(U_down[avail, :].T, mode="economic", pivoting=True)
        kdn[avail[p_dn[:n_dn]]] = np.arange(1, n_dn + 1)
    return kup, kdn


def find_initial_configuration(sampler, U_up, U_down):
    """find_initial_configuration! (src/MonteCarlo.jl:382-410): every walker
    starts from the QR configuration; a singular tildeU raises the
    reference's error text."""
    kup, kdn = init_conf_qr(U_up, U_down)
    nw = sampler.nwalkers
    try:
        sampler.set_configs(np.tile(kup, (nw, 1)), np.tile(kdn, (nw, 1)))
    except nat.SingularError as e:
        raise RuntimeError("QR-based configuration is singular. The Hamiltonian may be "
                           "rank-deficient.") from e
    return kup, kdn


def init(mc: MC, ctx: MCContext, params: dict | None = None, kappa=None):
    """Carlo.init! (src/MonteCarlo.jl:435-441).  With ``params["init"] ==
    "qr"`` every walker starts from the reference's deterministic pivoted-QR
    configuration; by default each walker starts from a seeded uniform Sz=0
    configuration (redrawn while tildeU is singular); ``kappa = (kappa_up,
    kappa_down)`` sets the configurations directly."""
    params = params or {}
    if kappa is not None:
        mc.sampler.set_configs(*kappa)
    elif params.get("init") == "qr":
        find_initial_configuration(mc.sampler, mc.Ham.U_up, mc.Ham.U_down)
    else:
        seed = int(params.get("init_seed", ctx.seed))
        mc.sampler.init_random(seed)
    mc.sampler.set_sweep_count(ctx.sweeps)


def sweep(mc: MC, ctx: MCContext, nsweeps: int = 1):
    """``nsweeps`` Carlo.sweep! calls on every walker (src/MonteCarlo.jl:538-607)
    without measuring; advances ctx.sweeps."""
    cnt, _ = mc.sampler.sweep(nsweeps)
    for k in mc.counters:
        mc.counters[k] += cnt[k]
    ctx.sweeps += nsweeps
    if ctx.is_thermalized():
        ctx.measure("acc", cnt["accepted"] / max(1, cnt["sweeps"]))
    return cnt


def measure(mc: MC, ctx: MCContext):
    """Carlo.measure! (src/MonteCarlo.jl:628-634): local energy of every walker
    when ctx.sweeps is a multiple of n_occupied."""
    if ctx.sweeps % mc.n_occupied == 0:
        eloc, _ = mc.sampler.measure()
        ctx.measure("OL", eloc)
        return eloc
    return None


def run(mc: MC, ctx: MCContext, nsteps: int):
    """Carlo's step loop for ``nsteps`` steps: sweep!, sweeps += 1, measure! when
    thermalised.  Sweeps between two measurement points go to the device in
    one call."""
    done = 0
    while done < nsteps:
        nocc = mc.n_occupied
        target = (ctx.sweeps // nocc + 1) * nocc
        if target <= ctx.thermalization:
            target = (ctx.thermalization // nocc + 1) * nocc
        chunk = min(nsteps - done, target - ctx.sweeps)
        sweep(mc, ctx, chunk)
        done += chunk
        if ctx.is_thermalized():
            measure(mc, ctx)
    return ctx


def binning_error(x: np.ndarray, binsize: int) -> float:
    n = len(x) // max(1, binsize)
    if n < 2:
        return float("nan")
    b = x[: n * binsize].reshape(n, binsize).mean(axis=1)
    return float(b.std(ddof=1) / math.sqrt(n))


def register_evaluables(params: dict):
    """Carlo.register_evaluables (src/MonteCarlo.jl:675-687): energy = <OL>/ns.
    Returns a function of the context giving (energy per site, error).  The
    error bins the per-measurement walker averages with the context's binsize."""
    ns = int(params["n1"]) * int(params["n2"]) * 3

    def energy(ctx: MCContext):
        ol = np.array([np.mean(x) for x in ctx.observables.get("OL", [])])
        if len(ol) == 0:
            return float("nan"), float("nan")
        return float(ol.mean() / ns), binning_error(ol / ns, ctx.binsize)

    return {"energy": energy}


def write_checkpoint(mc: MC, ctx: MCContext, path: str):
    """Carlo.write_checkpoint (src/MonteCarlo.jl:715-719): kappa_up/kappa_down of
    every walker plus the sweep counter (npz container; no HDF5 here)."""
    ku, kd = mc.sampler.get_configs()
    np.savez(path, kappa_up=ku.astype(np.int64), kappa_down=kd.astype(np.int64),
             sweeps=np.int64(ctx.sweeps), seed=np.int64(ctx.seed))


def read_checkpoint(mc: MC, ctx: MCContext, path: str):
    """Carlo.read_checkpoint! (src/MonteCarlo.jl:751-755).  Unlike the
    reference, W is rebuilt from the restored kappa at once."""
    with np.load(path, allow_pickle=False) as z:
        mc.sampler.set_configs(z["kappa_up"].astype(np.int32), z["kappa_down"].astype(np.int32))
        ctx.sweeps = int(z["sweeps"])
    mc.sampler.set_sweep_count(ctx.sweeps)
