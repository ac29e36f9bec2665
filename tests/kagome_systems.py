"""Shared set-up for the tests: closed-shell pi-flux systems and oracle
configurations."""
import numpy as np

from kagome_vmc import DoubleKagome, Hamiltonian

# Boundary conditions that close the shell at half filling (checked by
# test_host_setup.test_shell_closure): antiperiodic along a2 for L = 4, 8, 12
# (as the reference's scripts/job.jl:17-18), plain periodic for L = 6, 10, 14.
BC = {2: ((True, True), (False, True)), 4: ((True, True), (False, True)),
      6: ((True, True), (False, False)), 8: ((True, True), (False, True)),
      10: ((True, True), (False, False)), 12: ((True, True), (False, True)),
      14: ((True, True), (False, False))}

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


def tape(nw, nsweeps, nb, seed):
    """Seeded (u, bond) draws for every walker and sweep, walker-major."""
    rng = np.random.default_rng(seed)
    return rng.random((nw, nsweeps)), rng.integers(0, nb, (nw, nsweeps)).astype(np.int32)


def stagger_since(oracle, seed, nw, k):
    """Accepted-move counters after kvmc_stagger_refresh(seed): walker w
    draws Philox4x32-10 at counter (w, 0, 0, 2) under key (seed) and scales
    the first word to [0, k)."""
    key = [seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF]
    return [(oracle.philox4x32_10([w, 0, 0, 2], key)[0] * k) >> 32 for w in range(nw)]


def decision_margin(oracle, ch, u, bidx):
    """How close the reference kernel's Metropolis decision on draw (u, bond)
    is to a tie in the oracle chain's current state: the relative distance
    between min(1, |ratio|^2) and u / (Z_mu / Z_max); inf when the sweep
    returns before the ratio (src/MonteCarlo.jl:547-567).  A flip between
    an FP32 and an FP64 chain can only happen where this is near zero."""
    nb = len(ch.bonds)
    zr = oracle.z_count(ch.bonds, ch.kup, ch.kdn) / nb
    i, j = (int(x) for x in ch.bonds[bidx])
    can1 = ch.kup[i - 1] != 0 and ch.kdn[j - 1] != 0
    can2 = ch.kup[j - 1] != 0 and ch.kdn[i - 1] != 0
    if u > zr or not (can1 or can2):
        return np.inf
    if can1:
        lu, ld, Ku, Kd = ch.kup[i - 1], ch.kdn[j - 1], j, i
    else:
        lu, ld, Ku, Kd = ch.kup[j - 1], ch.kdn[i - 1], i, j
    p = min(1.0, abs(ch.Wup[Ku - 1, lu - 1] * ch.Wdn[Kd - 1, ld - 1]) ** 2)
    return abs(p - u / zr) / max(p, u / zr)


def first_divergence_is_near_tie(oracle, ham, kup0, kdn0, tu, tb, out_dev, out_or, tol=1e-3, **kw):
    """For a device outcome record that departs from the oracle's at sweep t:
    replays the oracle chain to t and returns (t, margin, margin < tol)."""
    t = int(np.flatnonzero(out_dev != out_or)[0])
    ch = oracle.Chain.start(ham.nn, ham.U_up, ham.U_down, kup0, kdn0)
    ch.acc_since = kw.pop("acc_since", 0)
    if t > 0:
        oracle.run_chain(ch, t, tape_u=tu[:t], tape_bond=tb[:t], **kw)
    m = decision_margin(oracle, ch, tu[t], tb[t])
    return t, m, m < tol
