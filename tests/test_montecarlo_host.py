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
