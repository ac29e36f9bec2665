"""Shared set-up for the tests: closed-shell pi-flux systems and oracle
configurations."""
import numpy as np

from kagome_vmc import DoubleKagome, Hamiltonian

# Boundary conditions that close the shell at half filling (checked by
# test_host_setup.test_shell_closure): antiperiodic along a2 for L = 4, 8, 12
# (as the reference's scripts/job.jl:17-18), plain periodic for L = 6, 10.
BC = {2: ((True, True), (False, True)), 4: ((True, True), (False, True)),
      6: ((True, True), (False, False)), 8: ((True, True), (False, True)),
      10: ((True, True), (False, False)), 12: ((True, True), (False, True))}

_CACHE = {}


def system(L):
    if L not in _CACHE:
        pbc, apbc = BC[L]
        lat = DoubleKagome.build(1.0, L, L, pbc, apbc)
        n = 3 * L * L // 2
        _CACHE[L] = Hamiltonian.build(n, n, lat)
    return _CACHE[L]


def oracle_configs(oracle, seed, nw, ham):
    ns, n = ham.ns, ham.N_up
    ku = np.zeros((nw, ns), dtype=np.int32)
    kd = np.zeros((nw, ns), dtype=np.int32)
    for w in range(nw):
        ku[w], kd[w] = oracle.random_config(seed, w, 0, ns, n, ham.N_down)
    return ku, kd


def well_conditioned_configs(oracle, seed, nw, ham, max_cond=1e6):
    """Seeded configurations whose tildeU are far from singular (keeps FP64
    agreement at the 1e-10 level a property of the code, not of the draw)."""
    ns, n = ham.ns, ham.N_up
    ku = np.zeros((nw, ns), dtype=np.int32)
    kd = np.zeros((nw, ns), dtype=np.int32)
    w = attempt = 0
    while w < nw:
        a, b = oracle.random_config(seed, 10_000 + attempt, 0, ns, n, ham.N_down)
        attempt += 1
        cu = np.linalg.cond(oracle.tilde_u(ham.U_up, a))
        cd = np.linalg.cond(oracle.tilde_u(ham.U_down, b))
        if cu < max_cond and cd < max_cond:
            ku[w], kd[w] = a, b
            w += 1
    return ku, kd
