"""Host logic of the Carlo plugin mirror (montecarlo.py) without a GPU: the
step loop's chunking and measurement cadence (src/MonteCarlo.jl:628-634 and
Carlo's sweep!/sweeps += 1/measure! order), thermalisation gating, binning
and the energy evaluable (src/MonteCarlo.jl:675-687).  The device sampler is
replaced by a recorder, so only the host bookkeeping runs."""
import math

import numpy as np
import pytest

from kagome_vmc.montecarlo import MCContext, binning_error, measure, register_evaluables, run


class RecordingSampler:
    def __init__(self, nw):
        self.nw = nw
        self.calls = []

    def sweep(self, n):
        self.calls.append(("sweep", n))
        return {"sweeps": n * self.nw, "moves": 0, "accepted": 0, "refreshes": 0}, None

    def measure(self):
        self.calls.append(("measure",))
        return np.full(self.nw, -2.0), 0


class StubMC:
    def __init__(self, nocc, nw=3):
        self.sampler = RecordingSampler(nw)
        self.N_up = self.N_down = nocc
        self.counters = {"sweeps": 0, "moves": 0, "accepted": 0, "refreshes": 0}

    @property
    def n_occupied(self):
        return min(self.N_up, self.N_down)


def reference_loop(nsteps, therm, nocc):
    """Carlo's loop one step at a time: which step counts end in a measurement."""
    out, sweeps = [], 0
    for _ in range(nsteps):
        sweeps += 1
        if sweeps > therm and sweeps % nocc == 0:
            out.append(sweeps)
    return out


@pytest.mark.parametrize("nsteps,therm,nocc", [(100, 0, 6), (100, 13, 6), (97, 30, 24), (5, 0, 6)])
def test_run_measures_at_reference_cadence(nsteps, therm, nocc):
    mc = StubMC(nocc)
    ctx = MCContext(seed=1, thermalization=therm)
    run(mc, ctx, nsteps)
    assert ctx.sweeps == nsteps
    assert sum(c[1] for c in mc.sampler.calls if c[0] == "sweep") == nsteps
    # replay the call log to find the sweep count at each measurement
    s, got = 0, []
    for c in mc.sampler.calls:
        if c[0] == "sweep":
            s += c[1]
        else:
            got.append(s)
    assert got == reference_loop(nsteps, therm, nocc)
    assert len(ctx.observables.get("OL", [])) == len(got)


def test_measure_skips_off_cadence():
    mc = StubMC(6)
    ctx = MCContext(sweeps=7)
    assert measure(mc, ctx) is None
    ctx.sweeps = 12
    assert measure(mc, ctx) is not None


def test_binning_error_closed_form():
    x = np.array([1.0, 3.0, 2.0, 2.0, 5.0, 1.0])
    # bins of 2: means 2, 2, 3 -> std 0.57735, / sqrt(3)
    assert math.isclose(binning_error(x, 2), np.std([2, 2, 3], ddof=1) / math.sqrt(3))
    assert math.isnan(binning_error(x, 6))


def test_energy_evaluable_normalises_by_sites():
    params = {"n1": 2, "n2": 2}
    ctx = MCContext(binsize=1)
    ctx.observables["OL"] = [np.array([-2.4, -2.4]), np.array([-3.6, -3.6])]
    e, err = register_evaluables(params)["energy"](ctx)
    assert math.isclose(e, -3.0 / 12)
    assert math.isclose(err, np.std([-0.2, -0.3], ddof=1) / math.sqrt(2))


# --- init_conf_qr! (src/MonteCarlo.jl:326-357), restating the properties the
# reference's own test checks (test/test-MonteCarlo.jl:14-52) ----------------
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_init_conf_qr_random_orbitals(seed, oracle_mod):
    from kagome_vmc.montecarlo import init_conf_qr
    rng = np.random.default_rng(seed)
    ns, n_up = 12, 6
    U_up, U_dn = rng.random((ns, n_up)), rng.random((ns, ns - n_up))
    ku, kd = init_conf_qr(U_up, U_dn)
    assert np.count_nonzero(ku) == n_up and np.count_nonzero(kd) == ns - n_up
    assert sorted(ku[ku != 0]) == list(range(1, n_up + 1))
    assert sorted(kd[kd != 0]) == list(range(1, ns - n_up + 1))
    assert np.all((ku != 0) ^ (kd != 0))
    assert abs(np.linalg.det(oracle_mod.tilde_u(U_up, ku))) > np.finfo(float).eps
    assert abs(np.linalg.det(oracle_mod.tilde_u(U_dn, kd))) > np.finfo(float).eps


@pytest.mark.parametrize("L", [2, 4, 8, 12])
def test_init_conf_qr_pi_flux_is_well_conditioned(L, oracle_mod):
    from kagome_vmc.montecarlo import init_conf_qr
    from kagome_systems import system
    ham = system(L)
    ku, kd = init_conf_qr(ham.U_up, ham.U_down)
    assert np.all((ku != 0) ^ (kd != 0))
    # column-pivoted QR picks well-spread rows: far from the 1e-10 redraw bar
    for U, k in ((ham.U_up, ku), (ham.U_down, kd)):
        assert np.linalg.cond(oracle_mod.tilde_u(U, k).real) < 1e8


def test_init_conf_qr_pivot_order_is_lapack_geqp3():
    """The labels follow the pivot order of LAPACK's column-pivoted QR (the
    routine behind Julia's qr(A, ColumnNorm())): the first pivot is the
    column of largest norm."""
    from kagome_vmc.montecarlo import init_conf_qr
    U = np.eye(4)[:, :2] * np.array([1.0, 3.0]) + 0.01
    ku, kd = init_conf_qr(U, np.ones((4, 2)))
    assert ku[1] == 1          # site 2 (row norm 3) is pivot 1
