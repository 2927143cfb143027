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
