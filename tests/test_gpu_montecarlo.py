"""A job.jl-shaped run through the Carlo plugin mirror on the device
(scripts/job.jl:10-33: MC(params), init!, the step loop, energy = OL/ns),
checked against the exact 12-site energy of the reference kernel's
stationary law |psi|^2 / Z_mu (oracle.exhaustive, pinned in
test_oracle_pins.py)."""
import numpy as np
import pytest

from kagome_vmc.montecarlo import MC, MCContext, init, read_checkpoint, register_evaluables, run, \
    write_checkpoint

pytestmark = pytest.mark.gpu

PARAMS = {"n1": 2, "n2": 2, "PBC": (True, True), "antiPBC": (False, True), "N_up": 6,
          "N_down": 6, "nwalkers": 1024, "thermalization": 600, "binsize": 4, "seed": 41}


@pytest.mark.parametrize("precision,refresh,delay", [("fp64", "reference", 1), ("fp32", 16, 4)])
def test_job_energy_matches_exact(oracle_mod, precision, refresh, delay):
    params = dict(PARAMS, precision=precision, refresh=refresh, delay=delay)
    mc = MC(params)
    try:
        ctx = MCContext.from_params(params)
        init(mc, ctx, params)
        run(mc, ctx, 600 + 6 * 60)
        e, err = register_evaluables(params)["energy"](ctx)
        ham = mc.Ham
        nmeas = len(ctx.observables["OL"])
    finally:
        mc.close()
    assert nmeas == 60
    ref = oracle_mod.exhaustive(ham.nn, ham.U_up, ham.U_down)
    exact = ref["E_biased"] / 12
    assert np.isfinite(err) and err < 2e-3
    assert abs(e - exact) < 5 * err, (e, exact, err)


def test_checkpoint_resume_continues_the_chain(tmp_path):
    """write_checkpoint / read_checkpoint! in the middle of a run: the resumed
    chain continues from the same kappa and sweep counter, with W rebuilt."""
    params = dict(PARAMS, nwalkers=64)
    path = str(tmp_path / "ckpt.npz")
    mc = MC(params)
    try:
        ctx = MCContext.from_params(params)
        init(mc, ctx, params)
        run(mc, ctx, 30)
        write_checkpoint(mc, ctx, path)
        k_saved = mc.kappa
    finally:
        mc.close()
    mc2 = MC(params)
    try:
        ctx2 = MCContext.from_params(params)
        read_checkpoint(mc2, ctx2, path)
        assert ctx2.sweeps == 30
        ku, kd = mc2.kappa
        assert np.array_equal(ku, k_saved[0]) and np.array_equal(kd, k_saved[1])
        assert np.all(mc2.sampler.sweep_counts() == 30)
        run(mc2, ctx2, 12)
        assert ctx2.sweeps == 42
    finally:
        mc2.close()
