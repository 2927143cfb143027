"""gm_restrict_noalloc (the restriction with its scratch in a static device
array) against gm_restrict, through the C ABI.

Both run the same kernels in the same order on the same inputs, and every
sum is taken in a fixed order, so the outputs must be equal bitwise.  Two
calls in a row, the second on a smaller pair, show that the scratch keeps
no state that changes a later result; a pair too large for the scratch is
rejected before any launch.
"""
import ctypes as C

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _pair_inputs(oracle, n, pair, seed):
    import computeFunctions as cf
    inp, _ = W.tiny(n=n)
    p = oracle.SetupProperties(inp["properties"])
    L = oracle.SetupLevels(inp, p)
    rng = np.random.default_rng(seed)
    fine = L[pair[1]]
    T = rng.uniform(300, 2500, fine["nn"]).astype(np.float32)
    S1 = (rng.uniform(0, 1, fine["nn"]) > 0.3).astype(np.float32)
    Tp, TpA, TpB = (rng.uniform(-50, 50, fine["nn"]).astype(np.float32) for _ in range(3))
    gc, gf = cf._grid_of(L[pair[0]]), cf._grid_of(fine)
    return cf, p, gc, gf, (T, S1, Tp, TpA, TpB)


def _call(fn, cf, p, gc, gf, arrays, dt):
    import _native as N
    dev = [cf._dev(a) for a in arrays]
    V = N.DeviceArray.from_numpy(np.zeros(gc.nn(), np.float32))
    fn_ = getattr(N.lib(), fn)
    N.check(fn_(C.byref(gc), C.byref(gf), C.byref(cf._material(p)), dev[0].ptr, dev[1].ptr, 40,
                dev[2].ptr, dev[3].ptr, dev[4].ptr, float(dt), V.ptr, None))
    N.sync(None)
    out = V.numpy()
    del dev
    return out


@pytest.mark.parametrize("cases", [[((12, 12, 6), (2, 3), 1), ((8, 8, 4), (1, 2), 2)],
                                   [((10, 10, 5), (1, 2), 3), ((8, 8, 4), (2, 3), 4)]])
def test_restrict_noalloc_equals_restrict(oracle, cases):
    for n, pair, seed in cases:  # the second call has a smaller input
        print(f"restrict pair {pair} on {n}, seed {seed}", flush=True)
        cf, p, gc, gf, arrays = _pair_inputs(oracle, n, pair, seed)
        ref = _call("gm_restrict", cf, p, gc, gf, arrays, 1e-3)
        got = _call("gm_restrict_noalloc", cf, p, gc, gf, arrays, 1e-3)
        assert np.any(ref != 0.0)
        assert np.array_equal(got, ref), float(np.max(np.abs(got - ref)))


def test_restrict_noalloc_rejects_oversize_pair(oracle):
    """A 400^3-element coarse cover needs far more than the 16 MiB scratch:
    the call fails on the host before any launch."""
    import _native as N
    import computeFunctions as cf
    inp, _ = W.tiny(n=(8, 8, 4))
    p = oracle.SetupProperties(inp["properties"])
    gc = N.gm_grid()
    gf = N.gm_grid()
    for d in range(3):
        gc.n[d], gc.h[d], gc.x0[d] = 401, 1.0, 0.0
        gf.n[d], gf.h[d], gf.x0[d] = 801, 0.5, 0.0
    dummy = N.DeviceArray(4)
    rc = N.lib().gm_restrict_noalloc(C.byref(gc), C.byref(gf), C.byref(cf._material(p)),
                                     dummy.ptr, dummy.ptr, 0, None, None, None, 1.0, dummy.ptr,
                                     None)
    assert rc != 0
    assert b"scratch" in N.lib().gm_last_error()
