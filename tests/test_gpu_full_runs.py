"""Full-length driver parity against committed oracle fixtures.

The fixtures (tests/golden/*.npz) are the final state of the CPU oracle's
driver (the reference time loop gm:134-493 over subcycleGOMELT cF:3224-3632,
stepGOMELT cF:2304-2397 and moveEverything cF:2400-2510), written by
tests/golden/make_fixtures.py in FP32 (the reference's arithmetic) and in
FP64.  The GPU driver runs the same deck here and is held to:

* against the FP64 fixture (the GPU run's error): max over nodes of
  |T_gpu - T_ref| / T_ref at most twice the FP32 oracle's own distance from
  the FP64 oracle on that level (the error FP32 arithmetic in the
  reference's order carries: 3.8e-4 on the single track, 1.3e-3 to 1.5e-3 on
  the production window, at near-ambient nodes where the mass-restriction
  terms divide T' differences by dt), and the same difference over the
  temperature rise at most 2e-4;
* against the FP32 fixture (the reference's arithmetic): within the sum of
  the two oracles' distance and the bound above (two FP32 runs that add in
  different orders agree no closer than each agrees with FP64), printed;
* melt-pool depth and width of the final L3 field within one L3 element of
  the oracle's (its own measure of the FP64 field);
* the same chunk modes (single-step / subcycled) as the oracle's driver,
  and the same Level0 melt-time maps up to a few nodes sitting on T_liquidus.

Decks:
  single_track  40x40x20 elements on every level, constant properties, one
                5 mm track at 1 m/s, P = 200 W: 5,000 toolpath rows, 50 L1
                steps (the first single-step, then 49 subcycled).
  production    the first 300 rows of the production deck (L1 200x200x50,
                L2 and L3 100x100x50 elements, 8-knot k(T), cp(T) tables,
                radiation): one single-step chunk and two subcycled L1 steps.
"""
import json
import os

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name, prec):
    path = os.path.join(GOLD, f"{name}_{prec}.npz")
    if not os.path.exists(path):
        pytest.skip(f"fixture {path} not generated (tests/golden/make_fixtures.py)")
    return dict(np.load(path))


def _run(name, tmp_path, max_rows=None):
    import go_melt as G
    inp, g = {"single_track": W.single_track, "production": W.production}[name]()
    with open(W.write_deck(inp, g, str(tmp_path))) as fh:
        deck = json.load(fh)
    run = G.TimeLoop(deck, quiet=True, write_outputs=False)
    done = 0
    for rows in run.chunks():
        if max_rows is not None and done >= max_rows:
            break
        if not run.step_chunk(rows):
            break
        done += len(rows)
    run.eng.sync()
    run.eng.sync_coords()  # the level dicts' node axes follow the engine's windows
    return run, deck


def _errors(run, fx, levels):
    out = {}
    for k in levels:
        a = np.asarray(run.Levels[k]["T0"], np.float64)
        b = fx[f"T0_{k}"].astype(np.float64)
        assert a.shape == b.shape, (k, a.shape, b.shape)
        out[k] = (float(np.max(np.abs(a - b) / np.abs(b))),
                  float(np.max(np.abs(a - b)) / max(float(np.max(b)) - 300.0, 1e-30)))
    return out


def _check(name, run, deck, f32, f64, levels, max_rows=None):
    import computeFunctions as cf
    chunks = [c == "single" for c in run.stats["chunks"]]
    assert np.array_equal(np.array(chunks, np.uint8), f32["chunks"]), name
    e32 = _errors(run, f32, levels)
    e64 = _errors(run, f64, levels)
    own = {}
    for k in levels:
        a, b = f32[f"T0_{k}"].astype(np.float64), f64[f"T0_{k}"].astype(np.float64)
        own[k] = float(np.max(np.abs(a - b) / np.abs(b)))
    print(f"{name}: FP32 oracle vs FP64 oracle (max |dT|/T) {own}")
    print(f"{name}: GPU vs FP32 oracle (max |dT|/T, max |dT|/rise) {e32}")
    print(f"{name}: GPU vs FP64 oracle (max |dT|/T, max |dT|/rise) {e64}")
    for k in levels:
        assert e64[k][0] <= 2.0 * own[k] + 1e-5, (name, k, e64[k], own[k])
        assert e64[k][1] <= 2e-4, (name, k, e64[k])
        assert e32[k][0] <= 3.0 * own[k] + 1e-5, (name, k, e32[k], own[k])
    # window positions
    for k in (2, 3):
        o = np.array([float(run.Levels[k]["node_coords"][d][0]) for d in range(3)])
        assert np.allclose(o, f64[f"origin_{k}"], atol=1e-5), (k, o, f64[f"origin_{k}"])
    # melt pool of the final L3 field: within one L3 element of the oracle's
    Tl = float(deck["properties"]["T_liquidus"])
    L3 = run.Levels[3]
    pg = cf.meltPoolDimensions(L3, L3["T0"], Tl)
    h = [float(v) for v in L3["h"]]
    for prec, fx in (("FP32", f32), ("FP64", f64)):
        po = fx["pool"]
        print(f"{name}: melt pool (depth, width, length) GPU {pg[:3]} {prec} oracle {po[:3]}")
        assert abs(pg[0] - po[0]) <= h[2] and abs(pg[1] - po[1]) <= h[1], (prec, pg, po)
    return e32, e64


def _check_accum(run, fx, dtmax):
    for key, name in (("accum", "accum_time"), ("max_accum", "max_accum_time")):
        a = np.asarray(run.Levels[0][name], np.float64)
        b = np.zeros_like(a)
        b[fx[f"{key}_idx"]] = fx[f"{key}_val"]
        d = np.abs(a - b)
        # FP32 sums of row dt over molten steps; a node sitting on T_liquidus
        # may flip phase a step earlier or later than in the oracle
        off = np.count_nonzero(d > 1e-5 * np.abs(b) + 1e-12)
        print(f"{key}: {np.count_nonzero(b)} molten nodes, {off} off by more than FP32 "
              f"rounding, max |d| {float(np.max(d)):.3g} s")
        assert off <= max(5, np.count_nonzero(b) // 200), (key, off)
        assert float(np.max(d)) <= 5 * dtmax, (key, float(np.max(d)))


def test_single_track_full_run(tmp_path):
    f32 = _fixture("single_track", "float32")
    f64 = _fixture("single_track", "float64")
    run, deck = _run("single_track", tmp_path)
    assert run.stats["chunks"].count("single") == 1 and run.stats["subcycle"] == 49
    _check("single_track", run, deck, f32, f64, (1, 2, 3))
    _check_accum(run, f64, 1e-6)
    # physics: the deck's properties and laser are Rosenthal's (constant k,
    # rho c, eta P = 70 W, 1 m/s); a Gaussian source of 50 um radius on a
    # finite body is not a point source on a half-space, so only a loose
    # agreement of the final pool with Rosenthal's is asserted (measured
    # 0.05% on depth and width)
    import computeFunctions as cf
    from rosenthal import rosenthal_pool
    _, depth, width, _ = rosenthal_pool()
    L3 = run.Levels[3]
    pg = cf.meltPoolDimensions(L3, L3["T0"], float(deck["properties"]["T_liquidus"]))
    print(f"single_track: melt pool depth, width {pg[:2]} Rosenthal {depth:.6f}, {width:.6f}")
    assert abs(pg[0] / depth - 1.0) < 0.05 and abs(pg[1] / width - 1.0) < 0.05


def test_production_window(tmp_path):
    f32 = _fixture("production", "float32")
    f64 = _fixture("production", "float64")
    rows = int(f32["rows"])
    run, deck = _run("production", tmp_path, max_rows=rows)
    assert run.stats["chunks"].count("single") == 1 and run.stats["subcycle"] == 2
    _check("production", run, deck, f32, f64, (1, 2, 3))
