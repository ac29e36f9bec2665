"""CPU oracle for the KagomeDSL VMC hot path -- TEST INFRASTRUCTURE ONLY.

This module is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it; the product package
(``kagomedsl.jl_amd/``) never does, and nothing it ships calls into it.

It wraps ``libkvmc_oracle.so`` (``kvmc_oracle.c``, a C restatement of the
reference algorithm, complex128 W as in the reference) and adds the numpy
pieces the reference delegates to LAPACK:

* ``orbitals``      -- src/Hamiltonian.jl:382-393 (eigen + sortperm)
* ``exhaustive``    -- brute-force energies over every Sz=0 configuration,
  under |psi|^2 and under the reference kernel's stationary law |psi|^2/Z_mu
* ``heisenberg_dense`` -- the spin-1/2 J1 Heisenberg matrix in the Sz basis,
  used to pin the variational energy as a quadratic form psi^T H psi.

Index conventions follow the reference: 1-based sites and labels, kappa[s] = 0
for an empty site, column-major (Fortran-order) matrices.
"""
from __future__ import annotations

import ctypes as C
import itertools
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

# pi-flux hopping tables: src/Hamiltonian.jl:219-235 and :250-261
PI_LINK_IN = {
    (1, 2): 1, (1, 3): 1, (2, 3): 1, (2, 4): -1, (4, 6): 1, (4, 5): 1, (5, 6): 1,
    (2, 1): 1, (3, 1): 1, (3, 2): 1, (4, 2): -1, (6, 4): 1, (5, 4): 1, (6, 5): 1,
}
PI_LINK_INTER = {
    (3, 5, -1, 1): -1, (3, 1, 0, 1): -1, (6, 2, 0, 1): -1, (6, 4, 0, 1): 1,
    (5, 1, 1, 0): 1, (1, 5, -1, 0): 1, (1, 3, 0, -1): -1, (2, 6, 0, -1): -1,
    (4, 6, 0, -1): 1, (5, 3, 1, -1): -1,
}

KERNEL_REFERENCE, KERNEL_UNBIASED = 0, 1
REFRESH_REFERENCE, REFRESH_EVERY_K = 0, 1
PRE_REJECT, NO_MOVE, REJECT, ACCEPT = 0, 1, 2, 3


class _Lattice(C.Structure):
    _fields_ = [
        ("n1", C.c_int), ("n2", C.c_int), ("pbc", C.c_int * 2), ("apbc", C.c_int * 2),
        ("t", C.c_double), ("a1", C.c_double * 2), ("a2", C.c_double * 2),
        ("r", (C.c_double * 2) * 6),
    ]


def build(force: bool = False) -> str:
    """Compiles libkvmc_oracle.so with the Makefile beside this file."""
    import subprocess

    so = os.path.join(_HERE, "libkvmc_oracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "libkvmc_oracle.so")
        if not os.path.exists(so):
            build()
        L = C.CDLL(so)
        P = C.POINTER
        L.kvo_lattice_init.argtypes = [P(_Lattice), C.c_double, C.c_int, C.c_int, C.c_int,
                                       C.c_int, C.c_int, C.c_int]
        L.kvo_unitcell_coord.argtypes = [P(_Lattice), C.c_int, P(C.c_double)]
        L.kvo_unitcell_diff.argtypes = [P(_Lattice), P(C.c_double), P(C.c_double), P(C.c_int),
                                        P(C.c_int)]
        L.kvo_site_coord.argtypes = [P(_Lattice), C.c_int, P(C.c_double)]
        L.kvo_boundary_shifts.argtypes = [P(_Lattice), C.c_int, C.c_int, P(C.c_int), P(C.c_int),
                                          P(C.c_double)]
        L.kvo_hmat.argtypes = [P(_Lattice), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_void_p, C.c_double, C.c_void_p]
        L.kvo_get_nn.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
        L.kvo_tilde_u.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.kvo_refresh_w.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                    P(C.c_double)]
        L.kvo_update_w.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int]
        L.kvo_z.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.kvo_getol.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                C.c_void_p, C.c_void_p, C.c_void_p, P(C.c_int)]
        L.kvo_getol.restype = C.c_double
        L.kvo_getol_c.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, P(C.c_double), P(C.c_double), P(C.c_int)]
        L.kvo_getxprime.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                    P(C.c_double), C.c_void_p, C.c_void_p]
        L.kvo_philox4x32_10.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.kvo_sweep_draws.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int,
                                      P(C.c_double), P(C.c_int)]
        L.kvo_random_config.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_int, C.c_int,
                                        C.c_int, C.c_void_p, C.c_void_p]
        L.kvo_run_chain.argtypes = [
            C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
            C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int,
            P(C.c_int), C.c_int64, C.c_int64, C.c_void_p, P(C.c_int64), C.c_void_p,
            P(C.c_int64)]
        L.kvo_run_chain.restype = C.c_int64
        _LIB = L
    return _LIB


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# --------------------------------------------------------------------------
# Lattice and ansatz
# --------------------------------------------------------------------------
class OracleLattice:
    """DoubleKagome(t, n1, n2, PBC; antiPBC): src/Lattice.jl:68-97."""

    def __init__(self, t, n1, n2, pbc, antipbc=(False, False)):
        self._s = _Lattice()
        rc = lib().kvo_lattice_init(C.byref(self._s), float(t), int(n1), int(n2), int(pbc[0]),
                                    int(pbc[1]), int(antipbc[0]), int(antipbc[1]))
        if rc == 1:
            raise ValueError("antiperiodic boundary conditions need periodic ones")
        if rc == 2:
            raise AssertionError("n1 must be even in DoubleKagome")
        self.n1, self.n2 = int(n1), int(n2)
        self.ns = 3 * self.n1 * self.n2

    def unitcell_coord(self, s):
        out = (C.c_double * 2)()
        lib().kvo_unitcell_coord(C.byref(self._s), int(s), out)
        return np.array(out[:])

    def unitcell_diff(self, c1, c2):
        a = (C.c_double * 2)(*map(float, c1))
        b = (C.c_double * 2)(*map(float, c2))
        dx, dy = C.c_int(), C.c_int()
        lib().kvo_unitcell_diff(C.byref(self._s), a, b, C.byref(dx), C.byref(dy))
        return dx.value, dy.value

    def site_coord(self, s):
        out = (C.c_double * 2)()
        lib().kvo_site_coord(C.byref(self._s), int(s), out)
        return np.array(out[:])

    def boundary_shifts(self, s1, s2):
        dx = (C.c_int * 9)()
        dy = (C.c_int * 9)()
        sg = (C.c_double * 9)()
        n = lib().kvo_boundary_shifts(C.byref(self._s), int(s1), int(s2), dx, dy, sg)
        return [(dx[k], dy[k], sg[k]) for k in range(n)]

    def hmat(self, link_in=None, link_inter=None, B=0.0):
        """Hmat(lat; link_in, link_inter, B): src/Hamiltonian.jl:308-354."""
        link_in = PI_LINK_IN if link_in is None else link_in
        link_inter = PI_LINK_INTER if link_inter is None else link_inter
        ki = list(link_in.items())
        in_l1 = np.array([k[0] for k, _ in ki], dtype=np.int32)
        in_l2 = np.array([k[1] for k, _ in ki], dtype=np.int32)
        in_v = np.array([v for _, v in ki], dtype=np.complex128)
        kt = list(link_inter.items())
        it = [np.array([k[q] for k, _ in kt], dtype=np.int32) for q in range(4)]
        it_v = np.array([v for _, v in kt], dtype=np.complex128)
        H = np.zeros((self.ns, self.ns), dtype=np.complex128, order="F")
        rc = lib().kvo_hmat(C.byref(self._s), len(ki), _ptr(in_l1), _ptr(in_l2), _ptr(in_v),
                            len(kt), _ptr(it[0]), _ptr(it[1]), _ptr(it[2]), _ptr(it[3]),
                            _ptr(it_v), float(B), _ptr(H))
        if rc:
            raise RuntimeError("tunneling matrix must be upper triangular")
        return H


def get_nn(H: np.ndarray):
    """get_nn: src/Hamiltonian.jl:447-451 -> list of 1-based (i, j), i <= j."""
    Hf = np.asfortranarray(H.astype(np.complex128))
    ns = Hf.shape[0]
    out = np.zeros(ns * ns * 2, dtype=np.int32)
    nb = lib().kvo_get_nn(ns, _ptr(Hf), _ptr(out))
    return [(int(out[2 * b]), int(out[2 * b + 1])) for b in range(nb)]


def orbitals(H: np.ndarray, n_up: int, n_dn: int):
    """orbitals: src/Hamiltonian.jl:382-393 (eigen(Hermitian) + sortperm)."""
    w, v = np.linalg.eigh(H)
    p = np.argsort(w, kind="stable")
    w, v = w[p], v[:, p]
    return np.asfortranarray(v[:, :n_up]), np.asfortranarray(v[:, :n_dn]), w


def bonds_array(nn) -> np.ndarray:
    return np.ascontiguousarray(np.array(nn, dtype=np.int32).reshape(-1, 2))


# --------------------------------------------------------------------------
# Sampler pieces
# --------------------------------------------------------------------------
def tilde_u(U, kappa):
    U = np.asfortranarray(U, dtype=np.complex128)
    k = np.ascontiguousarray(kappa, dtype=np.int32)
    ns, m = U.shape
    if len(k) != ns:
        raise ValueError("DimensionMismatch")
    out = np.zeros((m, m), dtype=np.complex128, order="F")
    rc = lib().kvo_tilde_u(ns, m, _ptr(U), _ptr(k), _ptr(out))
    if rc == 1:
        raise ValueError("kappa is not valid")
    if rc == 2:
        raise IndexError("label out of bounds")
    return out


def refresh_w(U, kappa, with_logdet=False):
    """reevaluateW! for one spin: W = U * (tilde_U \\ I)  (src/MonteCarlo.jl:55-66)."""
    U = np.asfortranarray(U, dtype=np.complex128)
    k = np.ascontiguousarray(kappa, dtype=np.int32)
    ns, m = U.shape
    W = np.zeros((ns, m), dtype=np.complex128, order="F")
    ld = C.c_double(0.0)
    rc = lib().kvo_refresh_w(ns, m, _ptr(U), _ptr(k), _ptr(W), C.byref(ld))
    if rc == 3:
        raise np.linalg.LinAlgError("SingularException")
    if rc:
        raise ValueError("invalid kappa")
    return (W, ld.value) if with_logdet else W


def update_w(W, l, K):
    """update_W!(W, l, K): src/MonteCarlo.jl:279-292 (in place, 1-based)."""
    assert W.flags.f_contiguous and W.dtype == np.complex128
    ns, m = W.shape
    lib().kvo_update_w(ns, m, _ptr(W), int(l), int(K))
    return W


def z_count(bonds, kup, kdn):
    b = bonds_array(bonds)
    return lib().kvo_z(len(b), _ptr(b), _ptr(np.ascontiguousarray(kup, dtype=np.int32)),
                       _ptr(np.ascontiguousarray(kdn, dtype=np.int32)))


def getol(bonds, kup, kdn, Wup, Wdn):
    b = bonds_array(bonds)
    Wu = np.asfortranarray(Wup, dtype=np.complex128)
    Wd = np.asfortranarray(Wdn, dtype=np.complex128)
    ku = np.ascontiguousarray(kup, dtype=np.int32)
    kd = np.ascontiguousarray(kdn, dtype=np.int32)
    err = C.c_int(0)
    e = lib().kvo_getol(Wu.shape[0], Wu.shape[1], Wd.shape[1], len(b), _ptr(b), _ptr(ku),
                        _ptr(kd), _ptr(Wu), _ptr(Wd), C.byref(err))
    if err.value:
        raise ValueError("site empty or doubly occupied")
    return e


def getol_complex(bonds, kup, kdn, Wup, Wdn):
    """The complex local energy (real part: getol)."""
    b = bonds_array(bonds)
    Wu = np.asfortranarray(Wup, dtype=np.complex128)
    Wd = np.asfortranarray(Wdn, dtype=np.complex128)
    ku = np.ascontiguousarray(kup, dtype=np.int32)
    kd = np.ascontiguousarray(kdn, dtype=np.int32)
    re, im, err = C.c_double(0.0), C.c_double(0.0), C.c_int(0)
    lib().kvo_getol_c(Wu.shape[0], len(b), _ptr(b), _ptr(ku), _ptr(kd), _ptr(Wu), _ptr(Wd),
                      C.byref(re), C.byref(im), C.byref(err))
    if err.value:
        raise ValueError("site empty or doubly occupied")
    return complex(re.value, im.value)


def getxprime(bonds, kup, kdn):
    b = bonds_array(bonds)
    ku = np.ascontiguousarray(kup, dtype=np.int32)
    kd = np.ascontiguousarray(kdn, dtype=np.int32)
    keys = np.zeros(8 * max(len(b), 1), dtype=np.int32)
    coef = np.zeros(2 * max(len(b), 1), dtype=np.float64)
    diag = C.c_double(0.0)
    n = lib().kvo_getxprime(len(b), _ptr(b), _ptr(ku), _ptr(kd), C.byref(diag), _ptr(keys),
                            _ptr(coef))
    if n < 0:
        raise ValueError("site empty or doubly occupied")
    out = {(-1, -1, -1, -1): diag.value}
    for q in range(n):
        out[tuple(int(x) for x in keys[4 * q:4 * q + 4])] = float(coef[q])
    return out


def philox4x32_10(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().kvo_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return [int(x) for x in o]


def sweep_draws(seed, walker, sweep, nb):
    u = C.c_double()
    b = C.c_int()
    lib().kvo_sweep_draws(seed, walker, sweep, nb, C.byref(u), C.byref(b))
    return u.value, b.value


def random_config(seed, walker, attempt, ns, n_up, n_dn):
    ku = np.zeros(ns, dtype=np.int32)
    kd = np.zeros(ns, dtype=np.int32)
    lib().kvo_random_config(seed, walker, attempt, ns, n_up, n_dn, _ptr(ku), _ptr(kd))
    return ku, kd


@dataclass
class Chain:
    """One walker's state for ``run_chain`` (reference MC fields, src/MonteCarlo.jl:17-28)."""

    bonds: np.ndarray
    Uup: np.ndarray
    Udn: np.ndarray
    kup: np.ndarray
    kdn: np.ndarray
    Wup: np.ndarray
    Wdn: np.ndarray
    sweeps: int = 0
    acc_since: int = 0

    @classmethod
    def start(cls, bonds, Uup, Udn, kup, kdn):
        Uup = np.asfortranarray(Uup, dtype=np.complex128)
        Udn = np.asfortranarray(Udn, dtype=np.complex128)
        kup = np.ascontiguousarray(kup, dtype=np.int32).copy()
        kdn = np.ascontiguousarray(kdn, dtype=np.int32).copy()
        return cls(bonds_array(bonds), Uup, Udn, kup, kdn, refresh_w(Uup, kup),
                   refresh_w(Udn, kdn))


def run_chain(ch: Chain, nsweeps, *, tape_u=None, tape_bond=None, seed=0, walker=0,
              kernel=KERNEL_REFERENCE, policy=REFRESH_REFERENCE, k=0, measure_every=0,
              thermalization=0):
    """Runs ``nsweeps`` Carlo.sweep! calls (src/MonteCarlo.jl:538-607) on one chain.

    Returns (outcomes[int8 per sweep], eloc list, n_refresh)."""
    ns, nup = ch.Uup.shape
    ndn = ch.Udn.shape[1]
    outcomes = np.zeros(nsweeps, dtype=np.int8)
    n_meas = nsweeps // measure_every + 2 if measure_every > 0 else 1
    eloc = np.zeros(n_meas, dtype=np.float64)
    ne = C.c_int64(0)
    nr = C.c_int64(0)
    acc_since = C.c_int(ch.acc_since)
    if tape_u is not None:
        tu = np.ascontiguousarray(tape_u, dtype=np.float64)
        tb = np.ascontiguousarray(tape_bond, dtype=np.int32)
        tup, tbp = _ptr(tu), _ptr(tb)
    else:
        tup, tbp = None, None
    acc = lib().kvo_run_chain(ns, nup, ndn, len(ch.bonds), _ptr(ch.bonds), _ptr(ch.Uup),
                              _ptr(ch.Udn), _ptr(ch.kup), _ptr(ch.kdn), _ptr(ch.Wup),
                              _ptr(ch.Wdn), ch.sweeps, nsweeps, tup, tbp, seed, walker, kernel,
                              policy, k, C.byref(acc_since), measure_every, thermalization,
                              _ptr(eloc), C.byref(ne), _ptr(outcomes), C.byref(nr))
    if acc < 0:
        raise np.linalg.LinAlgError("SingularException in refresh")
    ch.sweeps += nsweeps
    ch.acc_since = acc_since.value
    return outcomes, eloc[: ne.value].copy(), nr.value


# --------------------------------------------------------------------------
# Exhaustive enumeration (small clusters)
# --------------------------------------------------------------------------
def all_configs(ns, n_up):
    """Every Sz configuration with n_up up spins, as label vectors (labels in
    increasing site order)."""
    for ups in itertools.combinations(range(ns), n_up):
        kup = np.zeros(ns, dtype=np.int32)
        kdn = np.zeros(ns, dtype=np.int32)
        upset = set(ups)
        a = b = 0
        for s in range(ns):
            if s in upset:
                a += 1
                kup[s] = a
            else:
                b += 1
                kdn[s] = b
        yield kup, kdn


def heisenberg_dense(ns, bonds, n_up):
    """J1=1 Heisenberg matrix on the Sz sector with n_up up spins.  Basis: the
    configurations of ``all_configs`` in order.  Bonds with i == j (diagonal
    entries of H, present when H has on-site terms) count as in the reference:
    each adds Sz_i^2 = 1/4 and no flip."""
    confs = [tuple(int(x != 0) for x in kup) for kup, _ in all_configs(ns, n_up)]
    index = {c: q for q, c in enumerate(confs)}
    D = len(confs)
    Hm = np.zeros((D, D))
    for q, c in enumerate(confs):
        for (i, j) in bonds:
            i0, j0 = i - 1, j - 1
            si = 0.5 if c[i0] else -0.5
            sj = 0.5 if c[j0] else -0.5
            Hm[q, q] += si * sj
            if i0 != j0 and c[i0] != c[j0]:
                c2 = list(c)
                c2[i0], c2[j0] = c[j0], c[i0]
                Hm[index[tuple(c2)], q] += 0.5
    return Hm, confs


def spin_amplitudes(Uup, Udn, confs):
    """psi(sigma) = s(sigma) det U_up[ups] det U_dn[downs] with the site-ordered
    fermion sign s(sigma) = (-1)^{#pairs i<j with sigma_i = dn, sigma_j = up}."""
    out = np.zeros(len(confs), dtype=np.complex128)
    for q, c in enumerate(confs):
        ups = [s for s, x in enumerate(c) if x]
        dns = [s for s, x in enumerate(c) if not x]
        npairs = 0
        seen_dn = 0
        for x in c:
            if x:
                npairs += seen_dn
            else:
                seen_dn += 1
        sgn = -1.0 if npairs % 2 else 1.0
        out[q] = sgn * np.linalg.det(Uup[ups, :]) * np.linalg.det(Udn[dns, :])
    return out


def exhaustive(bonds, Uup, Udn):
    """Brute-force averages over every configuration with N_up up spins.

    Returns a dict with E (per cluster, |psi|^2 weight), E_biased (weight
    |psi|^2 / Z_mu: the stationary law of the reference kernel), the quadratic
    form psi^T H psi / psi^T psi, and the per-config arrays."""
    ns, nup = Uup.shape
    ndn = Udn.shape[1]
    assert nup + ndn == ns
    b = bonds_array(bonds)
    E, Ec, w, Zmu, logdet = [], [], [], [], []
    confs = []
    for kup, kdn in all_configs(ns, nup):
        confs.append(tuple(int(x != 0) for x in kup))
        try:
            Wu, lu = refresh_w(Uup, kup, with_logdet=True)
            Wd, ld = refresh_w(Udn, kdn, with_logdet=True)
        except np.linalg.LinAlgError:
            E.append(0.0); Ec.append(0.0); w.append(0.0); Zmu.append(z_count(b, kup, kdn))
            logdet.append(-np.inf)
            continue
        E.append(getol(b, kup, kdn, Wu, Wd))
        Ec.append(getol_complex(b, kup, kdn, Wu, Wd))
        w.append(np.exp(2.0 * (lu + ld)))
        Zmu.append(z_count(b, kup, kdn))
        logdet.append(lu + ld)
    E = np.array(E); w = np.array(w); Zmu = np.array(Zmu, dtype=np.float64)
    wb = np.where(Zmu > 0, w / np.maximum(Zmu, 1), 0.0)
    Hm, hconfs = heisenberg_dense(ns, [tuple(x) for x in b], nup)
    assert hconfs == confs
    psi = spin_amplitudes(np.asarray(Uup), np.asarray(Udn), confs)
    quad = float(np.real(np.vdot(psi, Hm @ psi) / np.vdot(psi, psi)))
    return {
        "E": float(np.sum(w * E) / np.sum(w)),
        "E_biased": float(np.sum(wb * E) / np.sum(wb)),
        "E_quadratic_form": quad,
        "eloc": E, "eloc_complex": np.array(Ec), "weight": w, "zmu": Zmu, "logdet": np.array(logdet),
        "confs": confs,
        "psi_spin": psi,
    }


# --------------------------------------------------------------------------
# Differentiation (no counterpart in the reference, which has none)
# --------------------------------------------------------------------------
def log_derivs(U, dU_list, kappa):
    """O_k = d log det tildeU / d theta_k = tr(tildeU^-1 d tildeU_k), with
    tildeU = U[R, :] and d tildeU_k = dU_k[R, :] (rows in label order), by a
    dense solve per parameter.  Pinned against finite differences of
    log|det tildeU(theta)|."""
    Ut = tilde_u(U, kappa)
    out = np.zeros(len(dU_list), dtype=np.complex128)
    for k, dU in enumerate(dU_list):
        dUt = tilde_u(np.asarray(dU, dtype=np.complex128), kappa)
        out[k] = np.trace(np.linalg.solve(Ut, dUt))
    if not (np.iscomplexobj(U) or any(np.iscomplexobj(d) for d in dU_list)):
        return out.real
    return out


def exhaustive_grad(bonds, Uup, Udn, dUup, dUdn):
    """|psi|^2-weighted averages over every configuration of the local energy
    E_L (complex for a complex psi) and the log-derivatives O_k: <E>,
    <O_k>, <conj(O_k) E_L>, S_kl = Re(<conj(O_k) O_l> - <conj(O_k)><O_l>) and
    grad_k = dE/dtheta_k = 2 Re(<conj(O_k) E_L> - E <conj(O_k)>), from
    dE = 2 Re <d psi|(H - E)|psi> / <psi|psi>.  For real psi these are the
    real formulas 2(<E_L O_k> - <E><O_k>)."""
    r = exhaustive(bonds, Uup, Udn)
    ns, nup = Uup.shape
    P = len(dUup)
    O = np.zeros((len(r["eloc"]), P), dtype=np.complex128)
    for q, (kup, kdn) in enumerate(all_configs(ns, nup)):
        if r["weight"][q] > 0:
            O[q] = log_derivs(Uup, dUup, kup) + log_derivs(Udn, dUdn, kdn)
    w = r["weight"] / r["weight"].sum()
    El = r["eloc_complex"]
    E = float(np.real(w @ El))
    Om = w @ O
    EO = (w * El) @ O.conj()
    S = np.real((O.conj() * w[:, None]).T @ O - np.outer(Om.conj(), Om))
    grad = 2.0 * np.real(EO - E * Om.conj())
    if np.all(np.imag(O) == 0) and np.all(np.imag(El) == 0):
        O, Om, EO = O.real, Om.real, EO.real
    r.update({"O_per_config": O, "O": Om, "EO": EO, "S": S, "grad": grad})
    return r
