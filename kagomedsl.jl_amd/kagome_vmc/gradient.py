"""Variational gradient of the projected-determinant energy.

The reference has no differentiation; this module adds it on top of the same
ansatz construction.  Four parameters theta = (phi, ratio, m, t2) around the
pi-flux state (the J1 bond list always comes from the pi-flux geometry, so
on-site or second-neighbour terms in H never enter the Heisenberg sum):

    H(theta) = H_up(phi) + ratio H_down(phi) + m D_stag + t2 G_nnn
      phi    flux phase: every link of the reference's pi-flux tables
             carries exp(i phi o / 3), o = +-1 its orientation in its
             triangle, so each triangle holds flux phi and each hexagon
             pi - 2 phi (flux_link_tables); H is complex for phi not in
             {0, pi},
      ratio  hopping ratio between the bonds outside the in-cell triangles
             (1-2-3 and 4-5-6: the 2-4 link and every inter-cell link,
             H_down) and those inside (H_up),
      m      staggered on-site mass (+ on labels 1-3, - on labels 4-6),
      t2     second-neighbour hopping on pairs at distance sqrt(3) t joined
             by a unique two-bond path, signed as that path's product of
             pi-flux amplitudes.
theta0 = (0, 1, 0, 0) is the reference pi-flux ansatz.  A field B adds the
reference's Peierls phases (src/Hamiltonian.jl:201, 325) to every hopping.

Orbital derivatives come from first-order perturbation theory on the closed
shell, dU[:, m] = sum_{n virtual} u_n (u_n^H dH u_m) / (eps_m - eps_n); the
device turns them into O_k(x) = d log det tildeU / d theta_k for every walker
(kvmc_measure_grad), complex when psi is.  With walkers drawn from |psi|^2
(the unbiased kernel) and the complex local energy E_L,

    dE/dtheta_k = 2 Re( <conj(O_k) E_L> - E <conj(O_k)> ),
    S_kl        = Re( <conj(O_k) O_l> - <conj(O_k)><O_l> ).

`EnergyFunction` exposes E(theta) as a torch.autograd.Function whose backward
returns that gradient (the counterpart of a ChainRulesCore rrule).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from .hamiltonian import Hmat, flux_link_tables, get_nn, pi_link_in, pi_link_inter, shell_gap
from .lattice import DoubleKagome, site_coords

NPARAMS = 4
PARAM_NAMES = ("phi", "ratio", "m", "t2")
THETA0 = np.array([0.0, 1.0, 0.0, 0.0])

_UP_LINKS = {(1, 2), (1, 3), (2, 3), (4, 5), (4, 6), (5, 6)}


@dataclass
class AnsatzFamily:
    lat: DoubleKagome
    D: np.ndarray      # staggered mass (real diagonal)
    G_nnn: np.ndarray  # second-neighbour hopping (real symmetric)
    nn: list           # J1 bonds of the pi-flux geometry
    B: float = 0.0     # magnetic field of the Peierls phases

    def _halves(self, phi, dphi=False):
        """H_up(phi) (in-cell triangles) and H_down(phi) (the other bonds),
        or their phi-derivatives."""
        tin, tinter = flux_link_tables(self.lat, phi, dphi=dphi)
        up_in = {k: v for k, v in tin.items() if (min(k), max(k)) in _UP_LINKS}
        down_in = {k: v for k, v in tin.items() if (min(k), max(k)) not in _UP_LINKS}
        return Hmat(self.lat, up_in, {}, B=self.B), Hmat(self.lat, down_in, tinter, B=self.B)

    def hamiltonian(self, theta) -> np.ndarray:
        phi, ratio, m, t2 = (float(x) for x in theta)
        Hu, Hd = self._halves(phi)
        H = Hu + ratio * Hd + m * self.D + t2 * self.G_nnn
        return H.real.copy() if np.all(H.imag == 0) else H

    def generators(self, theta) -> list:
        """dH/dtheta_k at theta."""
        phi, ratio = float(theta[0]), float(theta[1])
        dHu, dHd = self._halves(phi, dphi=True)
        Hu, Hd = self._halves(phi)
        real = lambda X: X.real.copy() if np.all(X.imag == 0) else X
        return [real(dHu + ratio * dHd), real(Hd), self.D, self.G_nnn]


def _torus_distance(lat: DoubleKagome, pos: np.ndarray) -> np.ndarray:
    """Minimal-image distance between every pair of sites on the torus
    spanned by (n1/2) a1 and n2 a2 (open directions are never wrapped)."""
    T1 = (lat.n1 // 2) * lat.a1
    T2 = lat.n2 * lat.a2
    d = pos[None, :, :] - pos[:, None, :]
    best = np.linalg.norm(d, axis=-1)
    for i in ((-1, 0, 1) if lat.PBC[0] else (0,)):
        for j in ((-1, 0, 1) if lat.PBC[1] else (0,)):
            best = np.minimum(best, np.linalg.norm(d + i * T1 + j * T2, axis=-1))
    return best


def ansatz_family(lat: DoubleKagome, B: float = 0.0) -> AnsatzFamily:
    H_pi = Hmat(lat).real
    n = H_pi.shape[0]
    stag = np.diag([1.0 if (s % 6) < 3 else -1.0 for s in range(n)])
    nn_mask = H_pi != 0
    pos = site_coords(lat)
    two = H_pi @ H_pi
    dist = _torus_distance(lat, pos)
    nnn = (~nn_mask) & (np.abs(two) == 1.0) & (np.abs(dist - math.sqrt(3.0) * lat.t) < 1e-9)
    g_nnn = np.where(nnn, np.sign(two), 0.0)
    return AnsatzFamily(lat, stag, g_nnn, get_nn(H_pi), float(B))


def orbital_derivatives(H: np.ndarray, G: list, n: int):
    """Occupied orbitals U (ns x n) and first-order dU_k for each dH = G[k]
    (parallel-transport gauge, U^H dU_k = 0).  Raises if the shell is open
    (the derivative would not exist).  Complex when H or any G[k] is."""
    w, v = np.linalg.eigh(H)
    if shell_gap(w, n) < 1e-8:
        raise ValueError("open shell: orbital derivatives undefined")
    U, V = v[:, :n], v[:, n:]
    denom = w[:n][None, :] - w[n:][:, None]          # (virt, occ): eps_m - eps_n
    dU = []
    for Gk in G:
        C = (V.conj().T @ Gk @ U) / denom
        dU.append(np.asfortranarray(V @ C))
    return np.asfortranarray(U), dU, w


def averages(eloc, O, weights=None):
    """Weighted <E>, <O_k>, <conj(O_k) E_L>, S and the gradient
    2 Re(<conj(O) E_L> - E <conj(O)>) (real formulas for real inputs)."""
    eloc = np.asarray(eloc)
    O = np.asarray(O)
    w = np.ones(eloc.shape[0]) if weights is None else np.asarray(weights, dtype=float)
    w = w / w.sum()
    E = float(np.real(w @ eloc))
    Om = w @ O
    EO = (w * eloc) @ O.conj()
    S = np.real((O.conj() * w[:, None]).T @ O - np.outer(Om.conj(), Om))
    grad = 2.0 * np.real(EO - E * Om.conj())
    return {"E": E, "O": Om, "EO": EO, "S": S, "grad": grad}


def _is_complex(*arrays) -> bool:
    return any(np.iscomplexobj(a) and np.any(np.imag(a) != 0) for a in arrays)


class GradientSampler:
    """Walkers drawn from |psi(theta)|^2 (unbiased kernel) with the orbital
    derivatives loaded, so each measurement yields E_L and O_k per walker.
    Complex orbitals, or complex derivatives (the phi direction), run on a
    complex handle."""

    def __init__(self, family: AnsatzFamily, theta, n_up: int, n_down: int, nwalkers: int,
                 seed: int = 0, precision: int = nat.PREC_FP64, delay: int = 1):
        self.family = family
        self.theta = np.asarray(theta, dtype=float)
        H = family.hamiltonian(self.theta)
        G = family.generators(self.theta)
        self.U_up, self.dU_up, _ = orbital_derivatives(H, G, n_up)
        self.U_dn, self.dU_dn, _ = orbital_derivatives(H, G, n_down)
        cplx = _is_complex(self.U_up, self.U_dn, *self.dU_up, *self.dU_dn)
        self.sampler = nat.DeviceSampler(np.array(family.nn, dtype=np.int32), self.U_up, self.U_dn,
                                         nwalkers, precision=precision, delay=delay,
                                         kernel=nat.KERNEL_UNBIASED, seed=seed,
                                         orbitals=nat.ORB_COMPLEX if cplx else nat.ORB_REAL)
        self.sampler.set_orbital_derivs(self.dU_up, self.dU_dn)

    def run(self, seed: int, therm: int, nmeas: int, every: int):
        """Random starts, `therm` sweeps, then `nmeas` measurements `every`
        sweeps apart.  Returns the averages and the raw samples, E_L of shape
        (nmeas, nwalkers) and O of shape (nmeas, nwalkers, nparams)."""
        s = self.sampler
        s.init_random(seed)
        s.sweep(therm)
        E, O = [], []
        for _ in range(nmeas):
            s.sweep(every)
            e, o = s.measure_grad()
            E.append(e)
            O.append(o)
        E, O = np.stack(E), np.stack(O)
        res = averages(E.ravel(), O.reshape(-1, O.shape[-1]))
        res["samples"] = (E, O)
        return res

    def close(self):
        self.sampler.close()


def exhaustive_average(family: AnsatzFamily, theta, n_up: int, n_down: int, det_floor=1e-10):
    """Exact |psi|^2 averages over every configuration with n_up up spins,
    evaluated on the device (one walker per configuration); configurations
    whose determinant vanishes (|det| below det_floor times the largest) are
    dropped, as their weight is zero."""
    import itertools
    H = family.hamiltonian(theta)
    G = family.generators(theta)
    U_up, dU_up, _ = orbital_derivatives(H, G, n_up)
    U_dn, dU_dn, _ = orbital_derivatives(H, G, n_down)
    ns = H.shape[0]
    confs = []
    for ups in itertools.combinations(range(ns), n_up):
        ku = np.zeros(ns, dtype=np.int32)
        kd = np.zeros(ns, dtype=np.int32)
        upset = set(ups)
        a = b = 0
        for s in range(ns):
            if s in upset:
                a += 1; ku[s] = a
            else:
                b += 1; kd[s] = b
        det = abs(np.linalg.det(U_up[list(ups), :]) *
                  np.linalg.det(U_dn[[s for s in range(ns) if s not in upset], :]))
        confs.append((ku, kd, det))
    dmax = max(c[2] for c in confs)
    keep = [c for c in confs if c[2] > det_floor * dmax]
    ku = np.stack([c[0] for c in keep])
    kd = np.stack([c[1] for c in keep])
    cplx = _is_complex(U_up, U_dn, *dU_up, *dU_dn)
    with nat.DeviceSampler(np.array(family.nn, dtype=np.int32), U_up, U_dn, len(keep),
                           kernel=nat.KERNEL_UNBIASED,
                           orbitals=nat.ORB_COMPLEX if cplx else nat.ORB_REAL) as ds:
        ds.set_configs(ku, kd)
        ds.set_orbital_derivs(dU_up, dU_dn)
        eloc, O = ds.measure_grad()
        logdet = ds.logdet()
    w = np.exp(2.0 * (logdet - logdet.max()))
    return averages(eloc, O, w)


def _torch_function():
    import torch

    class EnergyFunction(torch.autograd.Function):
        """E(theta) with the VMC gradient as its backward.

        forward(theta, evaluate) calls evaluate(theta_numpy) -> dict with
        "E" and "grad" (exhaustive_average or a GradientSampler run) and
        returns E; backward multiplies the stored gradient by grad_output."""

        @staticmethod
        def forward(ctx, theta, evaluate):
            res = evaluate(theta.detach().cpu().double().numpy())
            ctx.save_for_backward(torch.as_tensor(res["grad"], dtype=theta.dtype,
                                                  device=theta.device))
            return torch.as_tensor(res["E"], dtype=theta.dtype, device=theta.device)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out * g, None

    return EnergyFunction


def energy(theta, evaluate):
    """Differentiable energy: torch tensor theta -> scalar E (see EnergyFunction)."""
    return _torch_function().apply(theta, evaluate)
