"""Walker groups of different cluster sizes on one GPU.

Each group is one device handle (its own HIP stream and buffers).  The
groups' sweeps either run as one graph launch per step (graph=True: every
group's sweep sequence is a parallel branch of one captured hipGraph,
kvmc_group_sweep) or are driven by one host thread per group; either way
the groups' kernels overlap on the GPU.  Groups are independent Markov
chains, so a group's results do not depend on which other groups run beside
it, or on how they are launched.
"""
from __future__ import annotations

import threading
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from .hamiltonian import Hamiltonian, Hmat, flux_link_tables, get_nn, shell_gap
from .lattice import DoubleKagome

# antiPBC choices tried in order for a closed shell at half filling
_BC_CANDIDATES = ((False, True), (False, False), (True, False), (True, True))


def flux_cluster(L: int, phi: float = 0.0, n_up: int | None = None) -> Hamiltonian:
    """The L x L cluster (3 L^2 sites, periodic in both directions) of the
    flux family at phase phi (phi = 0: the pi-flux ansatz; otherwise complex
    orbitals), at half filling, with the first antiperiodic choice that
    closes the shell (the Slater determinant is then unique).  The J1 bonds
    are those of the pi-flux geometry."""
    n = 3 * L * L // 2 if n_up is None else n_up
    for apbc in _BC_CANDIDATES:
        lat = DoubleKagome.build(1.0, L, L, (True, True), apbc)
        tin, tint = flux_link_tables(lat, phi)
        w = np.linalg.eigvalsh(Hmat(lat, tin, tint))
        if shell_gap(w, n) > 1e-8 and shell_gap(w, 3 * L * L - n) > 1e-8:
            return Hamiltonian.build(n, 3 * L * L - n, lat, tin, tint, nn=get_nn(Hmat(lat).real))
    raise ValueError(f"no boundary condition closes the shell at L={L}, phi={phi}, N_up={n}")


def pi_flux_cluster(L: int, n_up: int | None = None) -> Hamiltonian:
    """The L x L pi-flux cluster at half filling (flux_cluster at phi = 0)."""
    return flux_cluster(L, 0.0, n_up)


@dataclass
class Group:
    ham: Hamiltonian
    sampler: nat.DeviceSampler

    @property
    def n_occupied(self) -> int:
        return min(self.ham.N_up, self.ham.N_down)


class MixedBatch:
    """groups: iterable of (Hamiltonian, nwalkers) or (Hamiltonian, nwalkers,
    overrides).  Sampler keywords (precision, refresh_policy, refresh_k,
    delay, threads, kernel, seed) apply to all groups; a group's overrides
    dict (e.g. the autotuner's threads and delay) replaces them for that
    group; group g draws with seed + g.  graph=True sweeps all groups with
    one graph launch per step."""

    def __init__(self, groups, graph: bool = False, **sampler_kw):
        seed = int(sampler_kw.pop("seed", 0))
        self.groups = []
        for g, item in enumerate(groups):
            ham, nw = item[0], item[1]
            kw = dict(sampler_kw)
            if len(item) > 2 and item[2]:
                kw.update(item[2])
            ds = nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, int(nw),
                                   seed=seed + g, **kw)
            self.groups.append(Group(ham, ds))
        self._graph = nat.SamplerGroup([g.sampler for g in self.groups]) if graph else None

    def _parallel(self, fn):
        out = [None] * len(self.groups)
        err = []

        def run(i):
            try:
                out[i] = fn(self.groups[i])
            except Exception as e:  # re-raised in the caller's thread
                err.append(e)

        ths = [threading.Thread(target=run, args=(i,)) for i in range(len(self.groups))]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        if err:
            raise err[0]
        return out

    def init_random(self, seed: int):
        return self._parallel(lambda g: g.sampler.init_random(seed))

    def step(self, nsteps: int = 1):
        """nsteps measurement intervals on every group: n_occupied sweeps of
        each walker, then its local energies.  Returns per group the summed
        counters and the list of per-walker energy arrays."""
        if self._graph is not None:
            keys = ("sweeps", "moves", "accepted", "refreshes", "flushes", "launches")
            tots = [dict.fromkeys(keys, 0) for _ in self.groups]
            energies = [[] for _ in self.groups]
            for _ in range(nsteps):
                cnts = self._graph.sweep([g.n_occupied for g in self.groups])
                for tot, cnt in zip(tots, cnts):
                    for k in keys:
                        tot[k] += cnt[k]
                for e, g in zip(energies, self.groups):
                    e.append(g.sampler.measure()[0])
            return list(zip(tots, energies))

        def run(g):
            tot = dict(sweeps=0, moves=0, accepted=0, refreshes=0, flushes=0, launches=0)
            energies = []
            for _ in range(nsteps):
                cnt, _ = g.sampler.sweep(g.n_occupied)
                for k in tot:
                    tot[k] += cnt[k]
                energies.append(g.sampler.measure()[0])
            return tot, energies
        return self._parallel(run)

    def close(self):
        if self._graph is not None:
            self._graph.close()
            self._graph = None
        for g in self.groups:
            g.sampler.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
