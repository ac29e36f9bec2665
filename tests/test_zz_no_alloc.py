"""kvmc_sweep_noalloc (the sweep entry point that allocates nothing: the
tape and outcome record live in the library's fixed device scratch) against
kvmc_sweep on twin handles, through the C ABI: two calls in a row, the
second with a smaller input, with identical outcomes, configurations, W,
local energies and counters; and the size limit rejected before any GPU
work."""
import ctypes

import numpy as np
import pytest

from kagome_systems import system, tape, well_conditioned_configs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L,prec,delay", [(4, 0, 1), (8, 1, 12)])
def test_noalloc_sweep_matches_sweep(oracle_mod, L, prec, delay):
    from kagome_vmc import _native as nat
    ham = system(L)
    nw = 24
    ku, kd = well_conditioned_configs(oracle_mod, 17, nw, ham, max_cond=1e5)
    calls = [(300, 5), (120, 6)]          # (sweeps, tape seed): the second call is smaller
    res = []
    for noalloc in (False, True):
        with nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nw, precision=prec, delay=delay,
                               refresh_policy=nat.REFRESH_EVERY_K, refresh_k=16, seed=3) as ds:
            ds.set_configs(ku, kd)
            rec = []
            for n, seed in calls:
                tu, tb = tape(nw, n, len(ham.nn), seed)
                cnt, out = ds.sweep(n, tu, tb, record=True, noalloc=noalloc)
                e, _ = ds.measure()
                rec.append((cnt, out, ds.get_configs(), e))
            # a device-RNG call with an outcome record only (no tape)
            cnt, out = ds.sweep(50, record=True, noalloc=noalloc)
            rec.append((cnt, out, ds.get_configs(), ds.measure()[0]))
            W = [ds.get_w(w, 0) for w in range(0, nw, 5)]
            res.append((rec, W))
    (r0, W0), (r1, W1) = res
    for (c0, o0, k0, e0), (c1, o1, k1, e1) in zip(r0, r1):
        assert c0 == c1
        assert np.array_equal(o0, o1)
        assert np.array_equal(k0[0], k1[0]) and np.array_equal(k0[1], k1[1])
        assert np.array_equal(e0, e1)
        assert c0["accepted"] > 0
    for a, b in zip(W0, W1):
        assert np.array_equal(a, b)


def test_noalloc_rejects_oversized_call():
    from kagome_vmc import _native as nat
    ham = system(4)
    nw = 64
    with nat.DeviceSampler(ham.bonds_array(), ham.U_up, ham.U_down, nw) as ds:
        ds.init_random(1)
        L = nat.load()
        n = 2_581_071 // nw + 1                      # one walker-sweep past the limit
        tu = np.zeros(4, dtype=np.float64)           # never read: the size check comes first
        tb = np.zeros(4, dtype=np.int32)
        cnt = nat.KvmcCounters()
        rc = L.kvmc_sweep_noalloc(ds._h, n, tu.ctypes.data_as(ctypes.c_void_p),
                                  tb.ctypes.data_as(ctypes.c_void_p), None, ctypes.byref(cnt))
        assert rc == nat.KVMC_ERR_INVALID
        assert b"exceeds" in L.kvmc_last_error()
        # without a tape or record nothing is staged: any size runs
        cnt, _ = ds.sweep(10, noalloc=True)
        assert cnt["sweeps"] == 10 * nw
