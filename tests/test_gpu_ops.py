"""GPU parity of the single operators against the FP64 oracle.

Each test builds the same level meshes for the oracle and for the facade
(the facade functions read only node_coords, nodes and h of a level dict),
runs the operator on seeded inputs and compares.  The kernels compute in
FP32; tolerances are stated per test.
"""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cf():
    import computeFunctions as cf
    import _native as N
    N.init_device()
    return cf


def _deck(oracle, n=(8, 8, 4), props=None):
    inp, _ = W.tiny(n=n, props=props)
    p = oracle.SetupProperties(inp["properties"])
    L = oracle.SetupLevels(inp, p)
    return inp, p, L


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def test_state_properties(cf, oracle):
    rng = np.random.default_rng(0)
    props = {"thermal_conductivity_bulk_a1": 0.016, "heat_capacity_solid_a1": 0.174,
             "thermal_conductivity_powder": 0.4, "heat_capacity_mushy": 3235.0,
             "heat_capacity_fluid": 769.0, "thermal_conductivity_fluid_a0": 29.0}
    for extra in ({}, {"thermal_conductivity_bulk_table": W.K_TABLE,
                       "heat_capacity_solid_table": W.CP_TABLE}):
        _, p, _ = _deck(oracle, props=dict(props, **extra))
        T = rng.uniform(250, 3200, 5000).astype(np.float32)
        T[:50] = p["T_liquidus"]
        T[50:100] = p["T_solidus"]
        S1 = (rng.uniform(0, 1, 5000) > 0.5).astype(np.float32)
        a = cf.computeStateProperties(T, S1, p, 300)
        b = oracle.computeStateProperties(T.astype(np.float64), S1.astype(np.float64), p, 300)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert _rel(a[2], b[2]) < 2e-6 and _rel(a[3], b[3]) < 2e-6


@pytest.mark.parametrize("active_layers", [None, 2])
def test_solve_matrix_free(cf, oracle, active_layers):
    _, p, L = _deck(oracle, n=(9, 7, 5))
    lv = L[1]
    rng = np.random.default_rng(1)
    nn, ne = lv["nn"], lv["ne"]
    if active_layers:
        ne = lv["elements"][0] * lv["elements"][1] * active_layers
    T = rng.uniform(300, 2000, nn)
    k = rng.uniform(0.01, 0.03, nn)
    c = rng.uniform(3e-3, 5e-3, nn)
    Fc = rng.uniform(-1, 1, nn) * 1e-2
    Corr = rng.uniform(-1, 1, nn) * 1e-2
    dt = 1e-2
    a = cf.solveMatrixFreeFE(lv, nn, ne, k, c, dt, T, Fc, Corr)
    b = oracle.solveMatrixFreeFE(lv, nn, ne, k, c, dt, T, Fc, Corr)
    act = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), act)
    assert np.max(np.abs(a[act] - b[act])) / np.max(np.abs(b[act])) < 5e-6


def test_source_l3(cf, oracle):
    _, p, L = _deck(oracle, n=(16, 12, 8))
    v = [4.1, 3.9, 0.0]
    a = cf.computeSourcesL3(L[3], v, None, p, 250.0)
    b = oracle.computeSourcesL3(L[3], v, (0, 0, 0, 0, L[3]["nn"]), p, 250.0)
    assert _rel(a, b) < 2e-6


@pytest.mark.parametrize("n_rows", [1, 10, 37])
def test_source_l3_mean(cf, oracle, n_rows):
    """gm_source_l3_mean: the mean of computeSourcesL3 (cF:2960-3012) over a
    substep's laser rows in one launch per 32 rows, against the mean of the
    oracle's per-row sources."""
    import ctypes as C
    import _native as N
    _, p, L = _deck(oracle, n=(16, 12, 8))
    rng = np.random.default_rng(n_rows)
    rows = np.zeros((n_rows, 7), np.float32)
    rows[:, 0] = 4.0 + rng.uniform(-0.4, 0.4, n_rows)
    rows[:, 1] = 4.0 + rng.uniform(-0.4, 0.4, n_rows)
    rows[:, 6] = rng.uniform(100, 300, n_rows)
    g = cf._grid_of(L[3])
    out = N.DeviceArray(g.nn())
    N.check(N.lib().gm_source_l3_mean(C.byref(g), C.byref(cf._material(p)),
                                      N.fptr(rows), n_rows, out.ptr, None))
    N.sync(None)
    ne_nn = (0, 0, 0, 0, L[3]["nn"])
    b = sum(oracle.computeSourcesL3(L[3], r[:3].astype(np.float64), ne_nn, p, float(r[6]))
            for r in rows) / n_rows
    assert _rel(out.numpy(), b) < 2e-6


@pytest.mark.parametrize("pair", [(1, 3), (2, 3), (1, 2)])
def test_level_source(cf, oracle, pair):
    _, p, L = _deck(oracle, n=(12, 12, 6))
    rng = np.random.default_rng(2)
    rows = np.zeros((5, 7))
    rows[:, 0] = 4.0 + rng.uniform(-0.3, 0.3, 5)
    rows[:, 1] = 4.0 + rng.uniform(-0.3, 0.3, 5)
    rows[:, 6] = rng.uniform(100, 300, 5)
    sh = oracle.computeCoarseFineShapeFunctions(L[pair[0]], L[pair[1]])
    a = cf.computeLevelSource(None, None, rows, cf.Shape(L[pair[0]], L[pair[1]]), p, rows[:, 6])
    b = oracle.computeLevelSource({3: L[pair[1]]}, None, rows, sh, p, rows[:, 6])
    assert _rel(a, b) < 5e-6


def test_conv_rad(cf, oracle):
    _, p, L = _deck(oracle, n=(10, 9, 5), props={"emissivity": 0.4, "evaporation_coefficient": 0.82})
    rng = np.random.default_rng(3)
    lv = L[1]
    T = rng.uniform(300, 3500, lv["nn"])
    for ne in (lv["ne"], lv["elements"][0] * lv["elements"][1] * 3):
        a = cf.computeConvRadBC(lv, T, ne, lv["nn"], p, 0)
        b = oracle.computeConvRadBC(lv, T, ne, lv["nn"], p, 0)
        assert _rel(a, b) < 1e-5


def test_interpolation(cf, oracle):
    _, p, L = _deck(oracle, n=(8, 8, 4))
    rng = np.random.default_rng(4)
    u = rng.uniform(300, 900, L[1]["nn"])
    # targets inside, on nodes, and partly outside the source mesh
    for coords in (L[2]["node_coords"], L[3]["node_coords"],
                   [np.linspace(-1, 9, 13), np.linspace(-0.5, 8.5, 7), np.linspace(-5, 1, 9)]):
        a = cf.interpolatePoints(L[1], u, coords)
        b = oracle.interpolatePoints(L[1], u, coords)
        assert _rel(a, b) < 2e-6
        Ma = cf.interpolatePointsMatrix(L[1], coords)
        Mb = oracle.interpolatePointsMatrix(L[1], coords)
        assert np.max(np.abs(Ma[0] - Mb[0])) < 2e-6
        w = Mb[0].sum(axis=1) > 0
        assert np.array_equal(Ma[1][w], Mb[1][w])
        assert _rel(cf.interpolate_w_matrix(Ma, u), oracle.interpolate_w_matrix(Mb, u)) < 2e-6


@pytest.mark.parametrize("pair,n", [((2, 3), (8, 8, 4)), ((1, 3), (8, 8, 4)),
                                    ((1, 2), (8, 8, 4)), ((1, 2), (10, 10, 5))])
def test_restriction(cf, oracle, pair, n):
    """Flux and mass restriction (the Shapes/BCOO products).  n = (10,10,5)
    makes the L1/L2 spacing ratio 2.5, so fine elements straddle coarse
    elements and the reference's Gauss-point-0 attribution is exercised."""
    inp, _ = W.tiny(n=n)
    if n == (10, 10, 5):
        inp["Level2"]["elements"] = [12, 12, 6]   # h1/h2 = 2.4
    p = oracle.SetupProperties(inp["properties"])
    L = oracle.SetupLevels(inp, p)
    rng = np.random.default_rng(5)
    fine = L[pair[1]]
    T = rng.uniform(300, 2500, fine["nn"])
    S1 = (rng.uniform(0, 1, fine["nn"]) > 0.3).astype(np.float64)
    Tp = rng.uniform(-50, 50, fine["nn"])
    TpA = rng.uniform(-50, 50, fine["nn"])
    TpB = rng.uniform(-50, 50, fine["nn"])
    sh = oracle.computeCoarseFineShapeFunctions(L[pair[0]], fine)
    S1o, _, k, rc = oracle.computeStateProperties(T, S1, p, 40)
    fa = cf.restrictFlux(cf.Shape(L[pair[0]], fine), p, T, S1, 40, Tp)
    fb = oracle.restrictFlux(sh, fine, Tp, k)
    assert _rel(fa, fb) < 2e-5
    ma = cf.restrictMass(cf.Shape(L[pair[0]], fine), p, T, S1, 40, TpA, TpB, 1e-3)
    mb = oracle.restrictMass(sh, fine, TpA - TpB, rc, 1e-3)
    assert _rel(ma, mb) < 2e-5


def test_new_tprime(cf, oracle):
    _, p, L = _deck(oracle, n=(8, 8, 4))
    rng = np.random.default_rng(6)
    fine, coarse = L[3], L[2]
    Tf = rng.uniform(300, 900, fine["nn"])
    Tc = rng.uniform(300, 900, coarse["nn"])
    C2F = oracle.interpolatePointsMatrix(coarse, fine["node_coords"])
    a_tp, a_c = cf.getNewTprime(fine, Tf, Tc, coarse, None)
    b_tp, b_c = oracle.getNewTprime(fine, Tf, Tc, coarse, C2F)
    assert _rel(a_c, b_c) < 2e-6
    assert np.max(np.abs(a_tp - b_tp)) < 2e-3


def _pool_level(n, x0, h):
    ax = [np.float32(x0[d]) + np.arange(n[d], dtype=np.float32) * np.float32(h[d])
          for d in range(3)]
    return {"node_coords": ax, "nodes": list(n), "h": [np.float32(v) for v in h]}


def _pool_fields(ax, rng):
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    X, Y, Z = (A.astype(np.float64) for A in (X, Y, Z))
    ztop = float(ax[2][-1])
    xc, yc = float(ax[0][len(ax[0]) // 3]), float(ax[1][len(ax[1]) // 2])
    yield "gauss", 300.0 + 2500.0 * np.exp(-((X - xc) / 0.12) ** 2 - ((Y - yc) / 0.05) ** 2
                                           - ((Z - ztop) / 0.04) ** 2)
    yield "tent", 1900.0 - 2000.0 * np.abs(Y - yc) - 1000.0 * np.abs(X - xc) - 6000.0 * (ztop - Z)
    yield "random", rng.uniform(1000.0, 2000.0, size=X.shape)
    yield "cold", np.full(X.shape, 300.0)
    yield "hot", np.full(X.shape, 2500.0)
    buried = np.full(X.shape, 300.0)
    buried[min(1, X.shape[0] - 1), min(2, X.shape[1] - 1), min(3, X.shape[2] - 1)] = 2000.0
    yield "buried", buried


@pytest.mark.parametrize("n", [(41, 33, 21), (101, 101, 51), (1, 1, 1), (5, 1, 3)])
def test_melt_pool(cf, oracle, n):
    """gm_melt_pool against the oracle's grid-line walk: depth, width, length,
    ends and molten count.  FP32 coordinates and crossings: 1e-6 mm."""
    rng = np.random.default_rng(11)
    h = (0.0125, 0.0125, 0.00625)
    lv = _pool_level(n, (1.25, -0.4, -0.3), h)
    Tl = 1600.0
    for name, T in _pool_fields(lv["node_coords"], rng):
        T = np.ascontiguousarray(T, np.float32).reshape(-1)
        got = cf.meltPoolDimensions(lv, T, Tl)
        ref = oracle.melt_pool_dims(T, lv["node_coords"], Tl)
        assert got[3] == ref[3], (name, got, ref)
        for g_, r_ in zip(got[:3] + got[4:], ref[:3] + ref[4:]):
            if np.isnan(r_):
                assert np.isnan(g_), (name, got, ref)
            else:
                assert abs(g_ - r_) <= 1e-6, (name, got, ref)
    # a cold field right after the hot ones: the reduction starts afresh
    T = np.full(int(np.prod(n)), 300.0, np.float32)
    assert cf.meltPoolDimensions(lv, T, Tl)[:4] == (0.0, 0.0, 0.0, 0)


@pytest.mark.parametrize("elements", [(12, 10, 8), (60, 50, 30)])
def test_energy_balance(cf, oracle, elements):
    """gm_energy_balance against the oracle on the GPU solver's own update
    with Dirichlet faces and a floor applied: every sum to 1e-7 of the
    scale; the residual of the plain update is FP32 rounding (< 1e-4); an
    error planted at an interior node moves the free sum by rho*cp*h^3*delta."""
    inp, p, _ = _deck(oracle)
    O = oracle
    rng = np.random.default_rng(5)
    lv = {"elements": list(elements),
          "bounds": {"x": [0.0, 0.0125 * elements[0]], "y": [1.0, 1.0 + 0.0125 * elements[1]],
                     "z": [-0.00625 * elements[2], 0.0]}}
    lv["length"], lv["h"] = O.calc_length_h(lv)
    lv["nodes"] = O.calcNumNodes(lv["elements"])
    lv["ne"], lv["nn"] = int(np.prod(elements)), int(np.prod(lv["nodes"]))
    b = lv["bounds"]
    lv["node_coords"], lv["connect"] = O.createMesh3D(
        (b["x"][0], b["x"][1], lv["nodes"][0]), (b["y"][0], b["y"][1], lv["nodes"][1]),
        (b["z"][0], b["z"][1], lv["nodes"][2]))
    nn, ne = lv["nn"], lv["ne"]
    T = rng.uniform(300.0, 2000.0, nn).astype(np.float32)
    k = rng.uniform(0.01, 0.03, nn).astype(np.float32)
    rc = np.full(nn, 4e-3, np.float32)
    Fc = rng.uniform(-1e-3, 5e-3, nn).astype(np.float32)
    dt = 1e-6
    Tt = cf.solveMatrixFreeFE(lv, nn, ne, k, rc, dt, T, Fc, 0)
    nx, ny, nz = lv["nodes"]
    iz, iy, ix = np.unravel_index(np.arange(nn), (nz, ny, nx))
    faces = (ix == 0) | (ix == nx - 1) | (iy == 0) | (iy == ny - 1) | (iz == 0)
    floor = np.float32(600.0)
    Tn = np.where(faces, np.float32(300.0), np.maximum(Tt, floor)).astype(np.float32)
    got = cf.energyBalance(lv, ne, k, rc, dt, T, Fc, 0, Tn, bc_faces=31, T_clamp=floor)
    ref = O.energy_balance(lv, ne, k, rc, dt, T, Fc, 0.0, Tn, bc_faces=31, T_clamp=float(floor))
    for q in (0, 1, 2, 3, 4, 5):
        assert abs(got[q] - ref[q]) <= 1e-7 * ref[5], (q, got, ref)
    assert got[7] == ref[7] and got[7] > faces.sum()
    assert abs(got[6]) < 1e-4 and abs(got[6] - ref[6]) < 1e-7
    i = 3 + nx * (3 + ny * 3)
    Tp = Tn.copy()
    Tp[i] += np.float32(50.0)
    got2 = cf.energyBalance(lv, ne, k, rc, dt, T, Fc, 0, Tp, bc_faces=31, T_clamp=floor)
    h = [float(np.float32(v)) for v in lv["h"]]
    want = 4e-3 * h[0] * h[1] * h[2] * float(Tp[i] - Tn[i])
    assert got2[0] - got[0] == pytest.approx(want, rel=1e-6)
