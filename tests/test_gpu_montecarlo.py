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


def test_qr_start_and_its_singular_error():
    """Carlo.init! with the reference's QR start (src/MonteCarlo.jl:382-441):
    a pi-flux cluster starts from a valid, non-singular configuration; a
    rank-1 U_up raises the reference's error text
    (test/test-MonteCarlo.jl:429-477)."""
    from kagome_vmc import _native as nat
    from kagome_vmc.montecarlo import find_initial_configuration
    params = dict(PARAMS, nwalkers=8, init="qr")
    mc = MC(params)
    try:
        ctx = MCContext.from_params(params)
        init(mc, ctx, params)
        ku, kd = mc.kappa
        assert np.all((ku != 0) ^ (kd != 0))
        assert np.all(ku == ku[0])
        run(mc, ctx, 12)
    finally:
        mc.close()
    rng = np.random.default_rng(0)
    U_bad = np.repeat(rng.random((1, 6)), 12, axis=0)
    ds = nat.DeviceSampler(np.array([[1, 2]], dtype=np.int32), U_bad, rng.random((12, 6)), 2)
    try:
        with pytest.raises(RuntimeError, match="QR-based configuration is singular"):
            find_initial_configuration(ds, U_bad, rng.random((12, 6)))
    finally:
        ds.close()
