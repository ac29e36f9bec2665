"""Device log-derivatives, gradient averages and the autograd wrapper against
the oracle (oracle.log_derivs / oracle.exhaustive_grad, both pinned by finite
differences in test_gradient_host.py).

Bars: 1e-9 relative for FP64 O_k and exhaustive averages; 2e-4 relative for
O_k from FP32 W; sampled 12-site energy and gradient within 5 jackknife
standard errors of the exact values.  theta1 carries a flux phase phi = 0.4,
so psi, W, E_L and O_k are complex there (complex handles); theta0 is the
real pi-flux state, whose flux-phase derivative is still complex.
"""
import numpy as np
import pytest
import torch

from kagome_vmc import DoubleKagome
from kagome_vmc import _native as nat
from kagome_vmc.gradient import (NPARAMS, THETA0, GradientSampler, ansatz_family, energy,
                                 exhaustive_average, orbital_derivatives)
from kagome_systems import BC

pytestmark = pytest.mark.gpu

THETA1 = np.array([0.4, 0.93, 0.07, -0.05])


def family(L):
    pbc, apbc = BC[L]
    return ansatz_family(DoubleKagome.build(1.0, L, L, pbc, apbc))


def orb(U, dU):
    cplx = any(np.iscomplexobj(a) and np.any(np.imag(a) != 0) for a in [U, *dU])
    return nat.ORB_COMPLEX if cplx else nat.ORB_REAL


def conditioned_configs(oracle_mod, U, nw, max_cond, seed=5):
    ns, n = U.shape
    ku = np.zeros((nw, ns), dtype=np.int32)
    kd = np.zeros((nw, ns), dtype=np.int32)
    w = attempt = 0
    while w < nw:
        a, b = oracle_mod.random_config(seed, 20_000 + attempt, 0, ns, n, ns - n)
        attempt += 1
        if (np.linalg.cond(oracle_mod.tilde_u(U, a)) < max_cond
                and np.linalg.cond(oracle_mod.tilde_u(U, b)) < max_cond):
            ku[w], kd[w] = a, b
            w += 1
    return ku, kd


@pytest.mark.parametrize("L", [2, 4, 8])
@pytest.mark.parametrize("prec", [0, 1])
def test_measure_grad_matches_oracle(oracle_mod, L, prec):
    f = family(L)
    n = 3 * L * L // 2
    U, dU, _ = orbital_derivatives(f.hamiltonian(THETA1), f.generators(THETA1), n)
    nw = 8
    ku, kd = conditioned_configs(oracle_mod, U, nw, 1e4 if prec else 1e6)
    with nat.DeviceSampler(np.array(f.nn, dtype=np.int32), U, U, nw, precision=prec, orbitals=orb(U, dU)) as ds:
        ds.set_configs(ku, kd)
        ds.set_orbital_derivs(dU, dU)
        eloc, O = ds.measure_grad()
    assert O.shape == (nw, NPARAMS)
    tol = 1e-9 if prec == 0 else 2e-4
    for w in range(nw):
        ref = oracle_mod.log_derivs(U, dU, ku[w]) + oracle_mod.log_derivs(U, dU, kd[w])
        assert np.abs(O[w] - ref).max() < tol * max(1.0, np.abs(ref).max()), (w, O[w], ref)
        Wu = oracle_mod.refresh_w(U, ku[w])
        Wd = oracle_mod.refresh_w(U, kd[w])
        e_ref = oracle_mod.getol_complex(f.nn, ku[w], kd[w], Wu, Wd)
        assert abs(eloc[w] - e_ref) < tol * max(1.0, abs(e_ref))
        assert abs(np.imag(e_ref)) > 1e-6      # complex psi: the imaginary part is live


def test_measure_grad_beyond_four_parameters(oracle_mod):
    """Five parameters (the four derivatives plus a fifth, 0.5 dU_0 - dU_3)
    take the workgroup-per-walker kernel; up to four take the site-major
    one.  Both agree with the oracle, and the fifth O equals 0.5 O_0 - O_3
    (O is linear in dU)."""
    L = 4
    f = family(L)
    n = 3 * L * L // 2
    U, dU, _ = orbital_derivatives(f.hamiltonian(THETA1), f.generators(THETA1), n)
    dU5 = list(dU) + [np.asfortranarray(0.5 * dU[0] - dU[3])]
    nw = 8
    ku, kd = conditioned_configs(oracle_mod, U, nw, 1e6, seed=9)
    with nat.DeviceSampler(np.array(f.nn, dtype=np.int32), U, U, nw, orbitals=orb(U, dU)) as ds:
        ds.set_configs(ku, kd)
        ds.set_orbital_derivs(dU, dU)
        _, O4 = ds.measure_grad()
        ds.set_orbital_derivs(dU5, dU5)
        _, O5 = ds.measure_grad()
    assert O4.shape == (nw, 4) and O5.shape == (nw, 5)
    for w in range(nw):
        ref = oracle_mod.log_derivs(U, dU5, ku[w]) + oracle_mod.log_derivs(U, dU5, kd[w])
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(O5[w] - ref).max() < 1e-9 * scale, (w, O5[w], ref)
        assert np.abs(O4[w] - ref[:4]).max() < 1e-9 * scale, (w, O4[w], ref[:4])
        assert abs(O5[w, 4] - (0.5 * O5[w, 0] - O5[w, 3])) < 1e-9 * scale


@pytest.mark.parametrize("prec,delay", [(0, 1), (1, 4)])
def test_measure_grad_after_sweeps(oracle_mod, prec, delay):
    """O_k after accepted moves (rank-1 updates, pending FP32 terms flushed by
    measure_grad) equals the oracle's value on the walkers' new configurations."""
    L = 4
    f = family(L)
    U, dU, _ = orbital_derivatives(f.hamiltonian(THETA1), f.generators(THETA1), 24)
    nw = 16
    ku, kd = conditioned_configs(oracle_mod, U, nw, 1e4)
    with nat.DeviceSampler(np.array(f.nn, dtype=np.int32), U, U, nw, precision=prec,
                           kernel=nat.KERNEL_UNBIASED, delay=delay, seed=3, orbitals=orb(U, dU)) as ds:
        ds.set_configs(ku, kd)
        ds.set_orbital_derivs(dU, dU)
        cnt, _ = ds.sweep(37)
        assert cnt["accepted"] > 0
        eloc, O = ds.measure_grad()
        ku2, kd2 = ds.get_configs()
    assert np.any(ku2 != ku)
    tol = 1e-8 if prec == 0 else 1e-3
    for w in range(nw):
        ref = oracle_mod.log_derivs(U, dU, ku2[w]) + oracle_mod.log_derivs(U, dU, kd2[w])
        assert np.abs(O[w] - ref).max() < tol * max(1.0, np.abs(ref).max()), w


def test_measure_grad_requires_derivatives():
    f = family(2)
    U, _, _ = orbital_derivatives(f.hamiltonian(THETA0), f.generators(THETA0), 6)
    with nat.DeviceSampler(np.array(f.nn, dtype=np.int32), U, U, 2) as ds:
        ds.init_random(1)
        with pytest.raises(nat.KvmcError, match="orbital derivatives"):
            ds.measure_grad()


@pytest.mark.parametrize("theta", [THETA0, THETA1])
def test_exhaustive_average_matches_oracle(oracle_mod, theta):
    f = family(2)
    U, dU, _ = orbital_derivatives(f.hamiltonian(theta), f.generators(theta), 6)
    ref = oracle_mod.exhaustive_grad(f.nn, U, U, dU, dU)
    got = exhaustive_average(f, theta, 6, 6)
    assert abs(got["E"] - ref["E"]) < 1e-9 * abs(ref["E"])
    for key in ("O", "EO", "grad"):
        assert np.abs(got[key] - ref[key]).max() < 1e-9 * max(1.0, np.abs(ref[key]).max()), key
    assert np.abs(got["S"] - ref["S"]).max() < 1e-9 * max(1.0, np.abs(ref["S"]).max())


def test_energy_gradcheck_device():
    """rrule-style check at 12 sites: the backward of energy() with the device
    evaluator against torch's finite differences of the same forward."""
    f = family(2)

    def evaluate(th):
        return exhaustive_average(f, th, 6, 6)

    theta = torch.tensor(THETA1, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: energy(t, evaluate), (theta,), eps=1e-5, atol=1e-9,
                                    rtol=1e-6)


def _jackknife(E, O, nblocks=32):
    """Jackknife errors of <E> and grad over blocks of walkers (walkers are
    independent chains; measurements of one walker are correlated)."""
    nw = E.shape[1]
    blocks = np.array_split(np.arange(nw), nblocks)

    def stats(mask):
        # E and dE/dtheta = 2 Re(<E_L conj(O)> - E <conj(O)>) (complex psi)
        e = E[:, mask].ravel()
        o = O[:, mask].reshape(-1, O.shape[-1])
        Em = float(np.real(e.mean()))
        return np.concatenate([[Em], 2.0 * np.real((e[:, None] * o.conj()).mean(0) - Em * o.conj().mean(0))])

    full = stats(np.ones(nw, dtype=bool))
    jk = []
    for b in blocks:
        m = np.ones(nw, dtype=bool)
        m[b] = False
        jk.append(stats(m))
    jk = np.array(jk)
    err = np.sqrt((nblocks - 1) / nblocks * ((jk - jk.mean(0)) ** 2).sum(0))
    return full, err


@pytest.mark.parametrize("prec", [0, 1])
def test_sampled_gradient_matches_exact(oracle_mod, prec):
    """Unbiased-kernel sampling at 12 sites reproduces the exact energy and
    gradient (stationary law |psi|^2) within 5 standard errors."""
    f = family(2)
    U, dU, _ = orbital_derivatives(f.hamiltonian(THETA1), f.generators(THETA1), 6)
    ref = oracle_mod.exhaustive_grad(f.nn, U, U, dU, dU)
    gs = GradientSampler(f, THETA1, 6, 6, 2048, seed=17, precision=prec)
    try:
        res = gs.run(seed=23, therm=400, nmeas=40, every=12)
    finally:
        gs.close()
    E, O = res["samples"]
    est, err = _jackknife(E, O)
    exact = np.concatenate([[ref["E"]], ref["grad"]])
    z = np.abs(est - exact) / err
    assert np.all(z < 5.0), (est, exact, err)
    assert err[0] < 0.02 * abs(ref["E"])


def test_reference_kernel_samples_biased_law(oracle_mod):
    """The reference kernel's stationary law is |psi|^2 / Z_mu: its sampled
    energy matches E_biased, not E (at 12 sites the two differ by > 10 sigma)."""
    f = family(2)
    U, dU, _ = orbital_derivatives(f.hamiltonian(THETA0), f.generators(THETA0), 6)
    ref = oracle_mod.exhaustive(f.nn, U, U)
    nw = 2048
    with nat.DeviceSampler(np.array(f.nn, dtype=np.int32), U, U, nw,
                           kernel=nat.KERNEL_REFERENCE, seed=29) as ds:
        ds.init_random(31)
        ds.sweep(400)
        E = []
        for _ in range(40):
            ds.sweep(12)
            E.append(ds.measure()[0])
    E = np.stack(E)
    per_walker = E.mean(0)
    mean, err = per_walker.mean(), per_walker.std(ddof=1) / np.sqrt(nw)
    assert abs(mean - ref["E_biased"]) < 5 * err, (mean, ref["E_biased"], err)
    if abs(ref["E"] - ref["E_biased"]) > 10 * err:
        assert abs(mean - ref["E"]) > 5 * err
