"""The C++/OpenMP restatement of the element loops (oracle/cpu_port.cpp, the
benchmark's CPU baseline) against the oracle's NumPy functions in FP32.

Each ported function is compared on small meshes with random fields, and a
short three-level run of the oracle's driver with the port installed is
compared with the same run on the NumPy functions.  Sums run in other orders
(element rows by colour, atomic adds on coarse nodes), so the agreement is to
FP32 rounding, not bitwise.
"""
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def port():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    from oracle import cpu_port
    return cpu_port


@pytest.fixture()
def pair(port):
    """(NumPy oracle functions, ported functions), both FP32; restores FP64."""
    from oracle import gomelt_oracle as O
    O.set_precision("float32")
    ref = {n: getattr(O, n) for n in dir(O) if not n.startswith("_")}
    port.install(O)
    new = {n: getattr(O, n) for n in dir(O) if not n.startswith("_")}
    port.uninstall(O)
    yield O, ref, new
    port.uninstall(O)
    O.set_precision("float64")


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _deck(O, n=(8, 8, 4), props=None):
    inp, _ = W.tiny(n=n, props=props)
    p = O.SetupProperties(inp["properties"])
    L = O.SetupLevels(inp, p)
    return p, L


def test_port_level_operators(pair):
    O, ref, new = pair
    p, L = _deck(O, props={"emissivity": 0.4, "evaporation_coefficient": 0.82,
                           "thermal_conductivity_bulk_table": W.K_TABLE,
                           "heat_capacity_solid_table": W.CP_TABLE})
    rng = np.random.default_rng(11)
    lv = L[1]
    T = rng.uniform(300, 3500, lv["nn"]).astype(np.float32)
    S1 = (rng.uniform(0, 1, lv["nn"]) > 0.5).astype(np.float32)
    a = ref["computeStateProperties"](T, S1, p, 40)
    b = new["computeStateProperties"](T, S1, p, 40)
    for x, y in zip(a, b):
        assert _rel(y, x) < 1e-6
    _, _, k, rc = a
    Fc = rng.uniform(-1, 1, lv["nn"]).astype(np.float32)
    Corr = rng.uniform(-1, 1, lv["nn"]).astype(np.float32)
    for ne in (lv["ne"], lv["elements"][0] * lv["elements"][1] * 2):
        x = ref["solveMatrixFreeFE"](lv, lv["nn"], ne, k, rc, 1e-3, T, Fc, Corr)
        y = new["solveMatrixFreeFE"](lv, lv["nn"], ne, k, rc, 1e-3, T, Fc, Corr)
        ok = np.isfinite(x)
        assert np.array_equal(ok, np.isfinite(y))
        assert _rel(y[ok], x[ok]) < 1e-5
        x = ref["computeConvRadBC"](lv, T, ne, lv["nn"], p, Fc)
        y = new["computeConvRadBC"](lv, T, ne, lv["nn"], p, Fc)
        assert _rel(y, x) < 1e-5


def test_port_sources_and_restriction(pair):
    O, ref, new = pair
    p, L = _deck(O, n=(8, 8, 4))
    rng = np.random.default_rng(12)
    L3 = L[3]
    v = [4.1, 3.9, 0.0]
    ne_nn = O.getStaticNodesAndElements(L)
    assert _rel(new["computeSourcesL3"](L3, v, ne_nn, p, 500.0),
                ref["computeSourcesL3"](L3, v, ne_nn, p, 500.0)) < 1e-5
    rows = np.array([[4.0 + 0.1 * i, 4.0, 0.0, 1, 1, 1e-3, 400.0 + 10 * i] for i in range(3)])
    for c in (1, 2):
        sh_r = ref["computeCoarseFineShapeFunctions"](L[c], L3)
        sh_n = new["computeCoarseFineShapeFunctions"](L[c], L3)
        assert _rel(new["computeLevelSource"](L, None, rows, sh_n, p, rows[:, 6]),
                    ref["computeLevelSource"](L, None, rows, sh_r, p, rows[:, 6])) < 1e-5
    for c, f in ((1, 2), (1, 3), (2, 3)):
        fine = L[f]
        sh_r = ref["computeCoarseFineShapeFunctions"](L[c], fine)
        sh_n = new["computeCoarseFineShapeFunctions"](L[c], fine)
        Tp = rng.uniform(-50, 50, fine["nn"]).astype(np.float32)
        k = rng.uniform(0.01, 0.03, fine["nn"]).astype(np.float32)
        assert _rel(new["restrictFlux"](sh_n, fine, Tp, k),
                    ref["restrictFlux"](sh_r, fine, Tp, k)) < 1e-5
        assert _rel(new["restrictMass"](sh_n, fine, Tp, k, 1e-3),
                    ref["restrictMass"](sh_r, fine, Tp, k, 1e-3)) < 1e-5


def test_port_interpolation(pair):
    O, ref, new = pair
    p, L = _deck(O)
    rng = np.random.default_rng(13)
    u = rng.uniform(300, 900, L[1]["nn"]).astype(np.float32)
    for coords in (L[2]["node_coords"], L[3]["node_coords"],
                   [np.linspace(-1, 9, 13), np.linspace(-0.5, 8.5, 7), np.linspace(-5, 1, 9)]):
        assert _rel(new["interpolatePoints"](L[1], u, coords),
                    ref["interpolatePoints"](L[1], u, coords)) < 2e-6
        M = new["interpolatePointsMatrix"](L[1], coords)
        assert _rel(new["interpolate_w_matrix"](M, u),
                    ref["interpolatePoints"](L[1], u, coords)) < 2e-6


def test_port_driver_run(port):
    """The oracle's driver over a small deck (single-step and subcycled
    chunks, window moves) with the port installed, against the NumPy run."""
    from oracle import gomelt_oracle as O
    inp, g = W.tiny(n=(8, 8, 4), n2=2, n3=3, track=1.6,
                    props={"laser_power": 600.0, "emissivity": 0.4,
                           "thermal_conductivity_bulk_table": W.K_TABLE,
                           "heat_capacity_solid_table": W.CP_TABLE})
    d = tempfile.mkdtemp(prefix="gomelt_port_")
    with open(W.write_deck(inp, g, d)) as fh:
        deck = json.load(fh)
    p = O.SetupProperties(deck["properties"])
    L = O.SetupLevels(deck, p)
    nm = O.SetupNonmesh(dict(deck["nonmesh"]))
    _, text = O.parsingGcode(dict(nm, toolpath=None), p, L[2]["h"])
    rows = np.array([[float(v) for v in ln.split(",")] for ln in text.splitlines()])
    try:
        O.set_precision("float32")
        La, ia = O.go_melt(deck, rows=rows)
        port.install(O)
        Lb, ib = O.go_melt(deck, rows=rows)
    finally:
        port.uninstall(O)
        O.set_precision("float64")
    assert ia["chunks"] == ib["chunks"] and "subcycle" in ia["chunks"]
    for k in (1, 2, 3):
        a = np.asarray(La[k]["T0"], np.float64)
        b = np.asarray(Lb[k]["T0"], np.float64)
        assert float(np.max(np.abs(a - b)) / np.max(np.abs(a - 300.0))) < 1e-4, k
