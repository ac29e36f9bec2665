"""ctypes binding of libkagome_vmc.so (the C ABI in include/kagome_vmc.h).

The library is the only compute path: there is no CPU fallback.  Loading
fails loudly when the shared object has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkagome_vmc.so")

KVMC_OK, KVMC_ERR_INVALID, KVMC_ERR_SINGULAR, KVMC_ERR_HIP, KVMC_ERR_NOMEM, KVMC_ERR_NOT_READY = range(6)
PREC_FP64, PREC_FP32 = 0, 1
ORB_REAL, ORB_COMPLEX = 0, 1
KERNEL_REFERENCE, KERNEL_UNBIASED = 0, 1
REFRESH_REFERENCE, REFRESH_EVERY_K, REFRESH_AT_END = 0, 1, 2
OUT_PRE_REJECT, OUT_NO_MOVE, OUT_REJECT, OUT_ACCEPT = 0, 1, 2, 3

EXPORTS = [
    "kvmc_create", "kvmc_destroy", "kvmc_set_configs", "kvmc_get_configs", "kvmc_init_random",
    "kvmc_refresh", "kvmc_sweep", "kvmc_measure", "kvmc_set_orbital_derivs",
    "kvmc_measure_grad", "kvmc_get_logdet", "kvmc_get_w", "kvmc_set_sweep_count",
    "kvmc_get_sweep_count", "kvmc_w_drift", "kvmc_failed_walker", "kvmc_last_error",
    "kvmc_version", "kvmc_profile_enable", "kvmc_get_profile", "kvmc_stagger_refresh",
    "kvmc_enable_graph", "kvmc_track_drift", "kvmc_get_drift_stats", "kvmc_sweep_noalloc",
    "kvmc_group_create", "kvmc_group_sweep", "kvmc_group_destroy",
]


class KvmcDesc(C.Structure):
    _fields_ = [
        ("ns", C.c_int32), ("n_up", C.c_int32), ("n_down", C.c_int32), ("nb", C.c_int32),
        ("bonds", C.POINTER(C.c_int32)), ("u_up", C.POINTER(C.c_double)),
        ("u_down", C.POINTER(C.c_double)), ("nwalkers", C.c_int64), ("precision", C.c_int32),
        ("kernel", C.c_int32), ("refresh_policy", C.c_int32), ("refresh_k", C.c_int32),
        ("delay", C.c_int32), ("threads", C.c_int32), ("seed", C.c_uint64),
        ("orbitals", C.c_int32),
    ]


class KvmcCounters(C.Structure):
    _fields_ = [(n, C.c_int64) for n in
                ("sweeps", "moves", "accepted", "refreshes", "flushes", "launches")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class KvmcProfile(C.Structure):
    _fields_ = [("sweep_ms", C.c_double), ("refresh_ms", C.c_double), ("measure_ms", C.c_double),
                ("sweep_launches", C.c_int64), ("refresh_launches", C.c_int64),
                ("measure_launches", C.c_int64)]

    def as_dict(self):
        return {n: (float if n.endswith("ms") else int)(getattr(self, n)) for n, _ in self._fields_}


class KvmcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kagome_vmc error {code}: {msg}")
        self.code = code


class SingularError(KvmcError):
    """A tildeU was singular (the reference's SingularException path)."""


_LIB = None


def load(path: str = LIB_PATH):
    """Loads the library; raises if it is missing (no fallback exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python kagomedsl.jl_amd/build.py` "
                           "(the sampler has no CPU fallback)")
    L = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "kvmc_create": ([P(KvmcDesc), P(vp)], C.c_int),
        "kvmc_destroy": ([vp], C.c_int),
        "kvmc_set_configs": ([vp, vp, vp], C.c_int),
        "kvmc_get_configs": ([vp, vp, vp], C.c_int),
        "kvmc_init_random": ([vp, C.c_uint64, C.c_int32, P(C.c_int64)], C.c_int),
        "kvmc_refresh": ([vp], C.c_int),
        "kvmc_sweep": ([vp, C.c_int64, vp, vp, vp, P(KvmcCounters)], C.c_int),
        "kvmc_sweep_noalloc": ([vp, C.c_int64, vp, vp, vp, P(KvmcCounters)], C.c_int),
        "kvmc_group_create": ([vp, C.c_int32, P(vp)], C.c_int),
        "kvmc_group_sweep": ([vp, vp, vp], C.c_int),
        "kvmc_group_destroy": ([vp], C.c_int),
        "kvmc_measure": ([vp, vp, P(C.c_double), P(C.c_int64)], C.c_int),
        "kvmc_set_orbital_derivs": ([vp, C.c_int32, vp, vp], C.c_int),
        "kvmc_measure_grad": ([vp, vp, vp], C.c_int),
        "kvmc_get_logdet": ([vp, vp], C.c_int),
        "kvmc_get_w": ([vp, C.c_int64, C.c_int32, vp], C.c_int),
        "kvmc_set_sweep_count": ([vp, C.c_int64], C.c_int),
        "kvmc_get_sweep_count": ([vp, vp], C.c_int),
        "kvmc_w_drift": ([vp, P(C.c_double)], C.c_int),
        "kvmc_failed_walker": ([vp], C.c_int64),
        "kvmc_last_error": ([], C.c_char_p),
        "kvmc_version": ([], C.c_char_p),
        "kvmc_profile_enable": ([vp, C.c_int32], C.c_int),
        "kvmc_get_profile": ([vp, P(KvmcProfile)], C.c_int),
        "kvmc_stagger_refresh": ([vp, C.c_uint64], C.c_int),
        "kvmc_enable_graph": ([vp, C.c_int32], C.c_int),
        "kvmc_track_drift": ([vp, C.c_int32], C.c_int),
        "kvmc_get_drift_stats": ([vp, P(C.c_double), P(C.c_double), P(C.c_int64)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _LIB = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != KVMC_OK:
        msg = load().kvmc_last_error().decode()
        if rc == KVMC_ERR_SINGULAR:
            raise SingularError(rc, msg)
        raise KvmcError(rc, msg)


class DeviceSampler:
    """A batch of walkers on the current GPU (one kvmc handle)."""

    def __init__(self, bonds, u_up, u_down, nwalkers, *, precision=PREC_FP64,
                 kernel=KERNEL_REFERENCE, refresh_policy=REFRESH_REFERENCE, refresh_k=0, delay=1,
                 threads=0, seed=0, orbitals=None):
        L = load()
        self.bonds = np.ascontiguousarray(np.asarray(bonds, dtype=np.int32).reshape(-1, 2))
        # complex orbitals (flux phases, magnetic field) give a complex W;
        # orbitals with zero imaginary part run as real unless `orbitals`
        # asks for ORB_COMPLEX (e.g. real U with complex derivatives)
        if orbitals is None:
            self.complex = bool((np.iscomplexobj(u_up) and np.any(np.imag(u_up) != 0)) or
                                (np.iscomplexobj(u_down) and np.any(np.imag(u_down) != 0)))
        else:
            self.complex = orbitals == ORB_COMPLEX
            if not self.complex and (np.any(np.imag(u_up) != 0) or np.any(np.imag(u_down) != 0)):
                raise ValueError("complex orbitals need orbitals=ORB_COMPLEX")
        dt = np.complex128 if self.complex else np.float64
        self.u_up = np.asfortranarray(u_up if self.complex else np.real(u_up), dtype=dt)
        self.u_down = np.asfortranarray(u_down if self.complex else np.real(u_down), dtype=dt)
        self.ns, self.n_up = self.u_up.shape
        self.n_down = self.u_down.shape[1]
        self.nb = len(self.bonds)
        self.nwalkers = int(nwalkers)
        self.precision = precision
        d = KvmcDesc()
        d.ns, d.n_up, d.n_down, d.nb = self.ns, self.n_up, self.n_down, self.nb
        d.bonds = self.bonds.ctypes.data_as(C.POINTER(C.c_int32))
        d.u_up = self.u_up.ctypes.data_as(C.POINTER(C.c_double))
        d.u_down = self.u_down.ctypes.data_as(C.POINTER(C.c_double))
        d.nwalkers = self.nwalkers
        d.precision, d.kernel, d.refresh_policy = precision, kernel, refresh_policy
        d.refresh_k, d.delay, d.threads, d.seed = refresh_k, delay, threads, seed
        d.orbitals = ORB_COMPLEX if self.complex else ORB_REAL
        h = C.c_void_p()
        _check(L.kvmc_create(C.byref(d), C.byref(h)))
        self._h = h
        self.nparams = 0

    def close(self):
        if getattr(self, "_h", None):
            load().kvmc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- configuration -------------------------------------------------------
    def set_configs(self, kappa_up, kappa_down):
        ku = np.ascontiguousarray(kappa_up, dtype=np.int32).reshape(self.nwalkers, self.ns)
        kd = np.ascontiguousarray(kappa_down, dtype=np.int32).reshape(self.nwalkers, self.ns)
        _check(load().kvmc_set_configs(self._h, _ptr(ku), _ptr(kd)))

    def get_configs(self):
        ku = np.zeros((self.nwalkers, self.ns), dtype=np.int32)
        kd = np.zeros((self.nwalkers, self.ns), dtype=np.int32)
        _check(load().kvmc_get_configs(self._h, _ptr(ku), _ptr(kd)))
        return ku, kd

    def init_random(self, seed, max_attempts=64):
        n = C.c_int64(0)
        _check(load().kvmc_init_random(self._h, seed, max_attempts, C.byref(n)))
        return n.value

    def refresh(self):
        _check(load().kvmc_refresh(self._h))

    # --- hot path -------------------------------------------------------------
    def sweep(self, nsweeps, tape_u=None, tape_bond=None, record=False, noalloc=False):
        """Runs nsweeps sweeps on every walker; returns (counters, outcomes|None).
        noalloc: stage the tape and outcomes in the library's fixed device
        scratch instead of per-call buffers (kvmc_sweep_noalloc)."""
        cnt = KvmcCounters()
        tu = tb = None
        if tape_u is not None:
            tu = np.ascontiguousarray(tape_u, dtype=np.float64).reshape(self.nwalkers, nsweeps)
            tb = np.ascontiguousarray(tape_bond, dtype=np.int32).reshape(self.nwalkers, nsweeps)
        out = np.zeros((self.nwalkers, nsweeps), dtype=np.int8) if record else None
        fn = load().kvmc_sweep_noalloc if noalloc else load().kvmc_sweep
        _check(fn(self._h, int(nsweeps), _ptr(tu), _ptr(tb), _ptr(out), C.byref(cnt)))
        return cnt.as_dict(), out

    def measure(self):
        """Per-walker local energies (FP64) and the antiparallel-bond count."""
        e = np.zeros(self.nwalkers, dtype=np.float64)
        s = C.c_double(0.0)
        nap = C.c_int64(0)
        _check(load().kvmc_measure(self._h, _ptr(e), C.byref(s), C.byref(nap)))
        return e, nap.value

    def set_orbital_derivs(self, du_up, du_down):
        dt = np.complex128 if self.complex else np.float64
        if not self.complex and (np.iscomplexobj(du_up) or np.iscomplexobj(du_down)):
            if np.any(np.imag(du_up) != 0) or np.any(np.imag(du_down) != 0):
                raise ValueError("complex orbital derivatives need complex orbitals")
            du_up, du_down = np.real(du_up), np.real(du_down)
        du_up = np.asarray(du_up, dtype=dt)
        du_down = np.asarray(du_down, dtype=dt)
        P = du_up.shape[0]
        a = np.ascontiguousarray(np.stack([np.asfortranarray(du_up[k]).ravel(order="F")
                                           for k in range(P)]))
        b = np.ascontiguousarray(np.stack([np.asfortranarray(du_down[k]).ravel(order="F")
                                           for k in range(P)]))
        _check(load().kvmc_set_orbital_derivs(self._h, P, _ptr(a), _ptr(b)))
        self.nparams = P

    def measure_grad(self):
        """Local energies and log-derivatives O_k per walker (complex arrays
        for complex orbitals: E_L = <x|H|psi>/<x|psi>, whose real part is
        the reference's getOL)."""
        dt = np.complex128 if self.complex else np.float64
        e = np.zeros(self.nwalkers, dtype=dt)
        o = np.zeros((self.nwalkers, self.nparams), dtype=dt)
        _check(load().kvmc_measure_grad(self._h, _ptr(e), _ptr(o)))
        return e, o

    def logdet(self):
        out = np.zeros(self.nwalkers, dtype=np.float64)
        _check(load().kvmc_get_logdet(self._h, _ptr(out)))
        return out

    def get_w(self, walker, spin):
        n = self.n_up if spin == 0 else self.n_down
        out = np.zeros(self.ns * n, dtype=np.complex128 if self.complex else np.float64)
        _check(load().kvmc_get_w(self._h, int(walker), int(spin), _ptr(out)))
        return out.reshape((self.ns, n), order="F")

    def enable_graph(self, enable=True):
        """Replay a captured sweep sequence instead of launching per round."""
        _check(load().kvmc_enable_graph(self._h, int(bool(enable))))

    def stagger_refresh(self, seed):
        """Spread the EVERY_K refresh phases of the walkers (see the C header)."""
        _check(load().kvmc_stagger_refresh(self._h, int(seed)))

    def set_sweep_count(self, sweeps):
        _check(load().kvmc_set_sweep_count(self._h, int(sweeps)))

    def sweep_counts(self):
        out = np.zeros(self.nwalkers, dtype=np.int64)
        _check(load().kvmc_get_sweep_count(self._h, _ptr(out)))
        return out

    def w_drift(self):
        v = C.c_double(0.0)
        _check(load().kvmc_w_drift(self._h, C.byref(v)))
        return v.value

    def track_drift(self, enable=True):
        """Record each walker's relative W drift just before every policy
        refresh (resets the statistics)."""
        _check(load().kvmc_track_drift(self._h, int(bool(enable))))

    def drift_stats(self):
        """(max, mean, count) of the relative W drift before refresh since
        track_drift was enabled."""
        mx, mn, n = C.c_double(0.0), C.c_double(0.0), C.c_int64(0)
        _check(load().kvmc_get_drift_stats(self._h, C.byref(mx), C.byref(mn), C.byref(n)))
        return mx.value, mn.value, n.value

    def profile(self, enable=None):
        """With enable=True/False: (re)start or stop kernel timing.  Without:
        return the accumulated kernel times (ms) and launch counts."""
        if enable is not None:
            _check(load().kvmc_profile_enable(self._h, int(bool(enable))))
            return None
        p = KvmcProfile()
        _check(load().kvmc_get_profile(self._h, C.byref(p)))
        return p.as_dict()

    def failed_walker(self):
        return int(load().kvmc_failed_walker(self._h))


class SamplerGroup:
    """Several DeviceSamplers on one GPU (e.g. lattices of different sizes)
    whose sweeps run as one graph launch per call (kvmc_group_sweep): each
    sampler's sweep sequence is a parallel branch of one captured graph.
    Borrows the samplers; close it before them."""

    def __init__(self, samplers):
        self.samplers = list(samplers)
        arr = (C.c_void_p * len(self.samplers))(*[s._h.value for s in self.samplers])
        g = C.c_void_p()
        _check(load().kvmc_group_create(arr, len(self.samplers), C.byref(g)))
        self._g = g

    def sweep(self, nsweeps):
        """nsweeps: one count per sampler.  Returns one counter dict each."""
        n = np.ascontiguousarray(nsweeps, dtype=np.int64)
        if n.shape != (len(self.samplers),):
            raise ValueError("one sweep count per sampler")
        cnt = (KvmcCounters * len(self.samplers))()
        _check(load().kvmc_group_sweep(self._g, _ptr(n), C.cast(cnt, C.c_void_p)))
        return [c.as_dict() for c in cnt]

    def close(self):
        if getattr(self, "_g", None):
            load().kvmc_group_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
