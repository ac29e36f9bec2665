"""Variational-gradient host math, pinned without the GPU.

The reference has no differentiation, so nothing here has a reference golden
vector; each piece is pinned by an independent route instead:
  * the generators dH/dtheta -- central differences of H(theta),
  * orbital derivatives   -- finite differences of the occupied projector,
  * oracle.log_derivs     -- finite differences of log|det tildeU(theta)|
                             (real part) and of log(det tildeU_x / det
                             tildeU_x') (the gauge-free complex part),
  * the energy gradient   -- finite differences of the exact 12-site energy
                             <psi|H|psi> / <psi|psi> (dense Heisenberg matrix),
  * the S matrix          -- the fidelity metric 1 - |<psi(a)|psi(b)>|^2,
  * the autograd wrapper  -- torch.autograd.gradcheck.
theta = (phi, ratio, m, t2); theta1 has a flux phase phi = 0.4, so psi is
complex there.
"""
import numpy as np
import pytest
import torch

from kagome_vmc import DoubleKagome, Hmat, get_nn
from kagome_vmc.lattice import site_coords
from kagome_vmc.gradient import (NPARAMS, THETA0, _torus_distance, ansatz_family, averages,
                                 energy, orbital_derivatives)
from kagome_systems import BC

THETA1 = np.array([0.4, 0.93, 0.07, -0.05])


def family(L):
    pbc, apbc = BC[L]
    return ansatz_family(DoubleKagome.build(1.0, L, L, pbc, apbc))


def occupied(H, n):
    w, v = np.linalg.eigh(H)
    return v[:, :n]


@pytest.mark.parametrize("L", [2, 4, 6])
def test_family_reduces_to_pi_flux(L):
    f = family(L)
    lat = f.lat
    H = f.hamiltonian(THETA0)
    assert np.array_equal(H, Hmat(lat).real)
    for G in f.generators(THETA1):
        assert np.abs(G - G.conj().T).max() < 1e-15
    # J1 list is the pi-flux one whatever theta is
    assert f.nn == get_nn(Hmat(lat).real)
    # at theta0 the flux direction is imaginary (time reversal maps phi to -phi)
    G0 = f.generators(THETA0)[0]
    assert np.all(np.real(G0) == 0) and np.any(np.imag(G0) != 0)


@pytest.mark.parametrize("L", [2, 4])
@pytest.mark.parametrize("theta", [THETA0, THETA1])
def test_generators_match_fd(L, theta):
    f = family(L)
    G = f.generators(theta)
    h = 1e-6
    for k in range(NPARAMS):
        e = np.zeros(NPARAMS); e[k] = h
        fd = (f.hamiltonian(theta + e) - f.hamiltonian(theta - e)) / (2 * h)
        assert np.abs(G[k] - fd).max() < 1e-9, k


@pytest.mark.parametrize("L", [4, 6])
def test_second_neighbour_and_mass_terms(L):
    f = family(L)
    ns = f.D.shape[0]
    H = f.hamiltonian(THETA0)
    G2 = f.G_nnn
    # four t2 partners per site, none of them a J1 neighbour, all at sqrt(3)
    assert np.all(np.count_nonzero(G2, axis=1) == 4)
    assert not np.any((G2 != 0) & (H != 0))
    d = _torus_distance(f.lat, site_coords(f.lat))
    assert np.allclose(d[G2 != 0], np.sqrt(3.0))
    # staggered mass: + on labels 1-3, - on labels 4-6
    assert np.array_equal(np.diag(f.D), np.where(np.arange(ns) % 6 < 3, 1.0, -1.0))
    assert ns == 3 * L * L


@pytest.mark.parametrize("L", [2, 4])
@pytest.mark.parametrize("theta", [THETA0, THETA1])
def test_orbital_derivative_matches_projector_fd(L, theta):
    f = family(L)
    n = 3 * L * L // 2
    U, dU, _ = orbital_derivatives(f.hamiltonian(theta), f.generators(theta), n)
    h = 1e-5
    for k in range(NPARAMS):
        e = np.zeros(NPARAMS); e[k] = h
        Pp = occupied(f.hamiltonian(theta + e), n)
        Pm = occupied(f.hamiltonian(theta - e), n)
        dP_fd = (Pp @ Pp.conj().T - Pm @ Pm.conj().T) / (2 * h)
        dP = dU[k] @ U.conj().T + U @ dU[k].conj().T
        assert np.abs(dP - dP_fd).max() < 1e-7, k
        assert np.abs(U.conj().T @ dU[k]).max() < 1e-12       # parallel-transport gauge


def test_open_shell_is_rejected():
    f = family(2)
    H = f.hamiltonian(THETA0)
    w = np.linalg.eigvalsh(H)
    # find a filling whose Fermi level is degenerate
    n = next(m for m in range(1, len(w)) if abs(w[m] - w[m - 1]) < 1e-10)
    with pytest.raises(ValueError, match="open shell"):
        orbital_derivatives(H, f.generators(THETA0), n)


@pytest.mark.parametrize("L", [2, 4])
def test_oracle_log_derivs_match_fd(oracle_mod, L):
    """Re O_k against the FD of log|det tildeU| (any unitary gauge of U);
    O_k(x) - O_k(x') against the FD of log(det tildeU_x / det tildeU_x'),
    which is free of the eigenvector phases (complex psi at theta1)."""
    f = family(L)
    n = 3 * L * L // 2
    ns = 3 * L * L
    theta = THETA1
    U, dU, _ = orbital_derivatives(f.hamiltonian(theta), f.generators(theta), n)
    h = 1e-5
    Up = [occupied(f.hamiltonian(theta + h * np.eye(NPARAMS)[k]), n) for k in range(NPARAMS)]
    Um = [occupied(f.hamiltonian(theta - h * np.eye(NPARAMS)[k]), n) for k in range(NPARAMS)]
    kept = []
    for trial in range(60):
        ku, _ = oracle_mod.random_config(11, trial, 0, ns, n, ns - n)
        if np.linalg.cond(oracle_mod.tilde_u(U, ku)) > 1e5:
            continue
        O = oracle_mod.log_derivs(U, dU, ku)
        assert np.iscomplexobj(O)
        for k in range(NPARAMS):
            lp = np.linalg.slogdet(oracle_mod.tilde_u(Up[k], ku))[1]
            lm = np.linalg.slogdet(oracle_mod.tilde_u(Um[k], ku))[1]
            assert abs(O[k].real - (lp - lm) / (2 * h)) < 1e-5 * max(1.0, abs(O[k])), (trial, k)
        kept.append((ku, O))
        if len(kept) == 6:
            break
    assert len(kept) == 6
    (k0, O0) = kept[0]
    for (k1, O1) in kept[1:]:
        for k in range(NPARAMS):
            rp = np.linalg.det(oracle_mod.tilde_u(Up[k], k1)) / np.linalg.det(oracle_mod.tilde_u(Up[k], k0))
            rm = np.linalg.det(oracle_mod.tilde_u(Um[k], k1)) / np.linalg.det(oracle_mod.tilde_u(Um[k], k0))
            fd = np.log(rp / rm) / (2 * h)
            assert abs((O1[k] - O0[k]) - fd) < 1e-5 * max(1.0, abs(fd)), k


def exact_energy(oracle_mod, f, theta, n):
    H = f.hamiltonian(theta)
    U = occupied(H, n)
    return oracle_mod.exhaustive(f.nn, U, U)["E_quadratic_form"]


@pytest.fixture(scope="module")
def twelve_site(oracle_mod):
    f = family(2)
    out = {}
    for name, th in (("theta0", THETA0), ("theta1", THETA1)):
        U, dU, _ = orbital_derivatives(f.hamiltonian(th), f.generators(th), 6)
        out[name] = (th, U, dU, oracle_mod.exhaustive_grad(f.nn, U, U, dU, dU))
    return f, out


@pytest.mark.parametrize("which", ["theta0", "theta1"])
def test_exhaustive_gradient_matches_energy_fd(oracle_mod, twelve_site, which):
    f, res = twelve_site
    theta, U, dU, r = res[which]
    assert abs(r["E"] - r["E_quadratic_form"]) < 1e-10
    # central differences with h = 1e-5 in FP64: truncation ~1e-11 and
    # rounding ~1e-10 absolute, so the bar is 1e-6 relative (with a 1e-9
    # absolute floor for components that vanish by symmetry)
    h = 1e-5
    for k in range(NPARAMS):
        e = np.zeros(NPARAMS); e[k] = h
        fd = (exact_energy(oracle_mod, f, theta + e, 6) - exact_energy(oracle_mod, f, theta - e, 6)) / (2 * h)
        assert abs(r["grad"][k] - fd) < 1e-6 * abs(fd) + 1e-9, (k, r["grad"][k], fd)


@pytest.mark.parametrize("which", ["theta0", "theta1"])
def test_s_matrix_is_fidelity_metric(oracle_mod, twelve_site, which):
    f, res = twelve_site
    theta, U, dU, r = res[which]
    confs = r["confs"]

    def psi(th):
        Uo = occupied(f.hamiltonian(th), 6)
        p = oracle_mod.spin_amplitudes(Uo, Uo, confs)
        return p / np.linalg.norm(p)

    # 1 - F = h^2 v^T S v + O(h^4); rounding of F stays near 1e-16 / h^2
    h = 3e-4
    S = r["S"]
    for k in range(NPARAMS):
        for l in range(k, NPARAMS):
            v = np.zeros(NPARAMS); v[k] += 1; v[l] += 1
            F = abs(np.vdot(psi(theta + 0.5 * h * v), psi(theta - 0.5 * h * v))) ** 2
            assert abs((1.0 - F) / h**2 - v @ S @ v) < 1e-5 * max(1.0, v @ S @ v), (k, l)
    assert np.all(np.linalg.eigvalsh(S) > -1e-12)


def test_averages_weighted_closed_form():
    eloc = np.array([1.0, 3.0])
    O = np.array([[2.0], [0.0]])
    r = averages(eloc, O, weights=[3.0, 1.0])
    # <E> = 1.5, <O> = 1.5, <EO> = 1.5, S = <O^2> - <O>^2 = 3 - 2.25
    assert r["E"] == 1.5 and r["O"][0] == 1.5 and r["EO"][0] == 1.5
    assert r["S"][0, 0] == 0.75 and r["grad"][0] == 2 * (1.5 - 1.5 * 1.5)


def test_energy_function_gradcheck_cpu(oracle_mod):
    """The autograd wrapper's backward against torch's own finite differences
    at 12 sites, with the oracle's exhaustive averages as the evaluator."""
    f = family(2)

    def evaluate(th):
        U, dU, _ = orbital_derivatives(f.hamiltonian(th), f.generators(th), 6)
        return oracle_mod.exhaustive_grad(f.nn, U, U, dU, dU)

    theta = torch.tensor(THETA1, dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: energy(t, evaluate), (theta,), eps=1e-5, atol=1e-9,
                                    rtol=1e-6)
    E = energy(theta, evaluate)
    (g,) = torch.autograd.grad(2.5 * E, theta)
    assert np.allclose(g.numpy(), 2.5 * evaluate(THETA1)["grad"], atol=1e-12)
