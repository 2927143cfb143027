"""Pins of the oracle's orchestration: moveEverything, stepGOMELT and
subcycleGOMELT (computeFunctions.py = cF:2400-2510, 2304-2397, 3224-3632).

These functions thread state between the pinned operators (the element step,
sources, interpolation, restriction, T').  No closed form gives their output,
so each test fixes something the mathematics of the scheme does, chosen so
that a plausible sequencing slip fails it:

* a uniform field with no source and no surface flux stays uniform, and T'
  stays zero, through all three functions;
* re-moving the windows to where they are (the same laser row) leaves every
  T0 and T' unchanged: the re-interpolated composite I(L1 T0) + I(L2 T')
  + I(L3 T') (cF:2439-2444) is the fine field again only if T' was formed as
  the reference forms it (inject, then subtract the prolonged coarse field,
  cF:2086-2097) and the pre-move meshes are used;
* a move by whole spacings carries trilinear L1 T0, L2 T' and L3 T' fields
  exactly, with T' = 0 where a moved node leaves the old window (the
  outside rule, cF:1101-1102) and the window shift truncated toward zero
  (cF:1716-1718);
* subcycleGOMELT with N2 = N3 = 1 is stepGOMELT (both are the same
  predictor/corrector when no node changes phase; the reference's two code
  paths are restated separately in the oracle, through different helpers).
  It is run again with a mistake planted in the oracle (a Part-2 L2 step
  that forgets the predictor's T' history) and must then fail.

A global heat balance over an L1 step is NOT a pin: the scheme does not
conserve the composite heat content exactly (the injection of a fine peak
into coarse nodes and the predictor's T' history in the L1 mass terms move
it), so no identity of that kind holds to a tolerance that a slip would
exceed.
"""

import copy

import numpy as np
import pytest

import workloads as W


def _ctx(O, inp):
    """Oracle state of a deck after the driver's set-up (gm:29-84)."""
    p = O.SetupProperties(inp["properties"])
    L = O.SetupLevels(inp, p)
    nm = O.SetupNonmesh(inp["nonmesh"])
    c = {"p": p, "L": L, "nm": nm, "ne_nn": O.getStaticNodesAndElements(L),
         "sub": O.getStaticSubcycle(nm)}
    c["r12"] = [int(np.round(L[1]["h"][i] / L[2]["h"][i])) for i in range(2)] + \
               [int(np.round(p["layer_height"] / L[2]["h"][2]))]
    c["r23"] = [int(np.round(L[2]["h"][i] / L[3]["h"][i])) for i in range(3)]
    c["vstart"] = np.array(p["laser_center"], np.float64)
    c["LInterp"] = [O.interpolatePointsMatrix(L[1], L[2]["node_coords"]),
                    O.interpolatePointsMatrix(L[2], L[3]["node_coords"])]
    c["tmp"] = O.calcStaticTmpNodesAndElements(L, c["vstart"])
    c["move_hist"] = [0, 0, 0]
    return c


def _move(O, c, v):
    c["L"], c["Shapes"], c["LInterp"], c["move_hist"] = O.moveEverything(
        v, c["vstart"], c["L"], c["move_hist"], c["LInterp"], c["r12"], c["r23"],
        c["p"]["layer_height"])
    c["substrate"] = O.getSubstrateNodes(c["L"])


def _step(O, c, row):
    c["L"], reset = O.stepGOMELT(c["L"], c["ne_nn"], c["tmp"], c["Shapes"], c["LInterp"], row,
                                 c["p"], row[5], row[6], c["substrate"])
    return reset


def _subcycle(O, c, rows):
    n3 = c["L"][3]["nn"]
    out = O.subcycleGOMELT(c["L"], c["ne_nn"], c["Shapes"], c["substrate"], c["LInterp"],
                           c["tmp"], rows, c["p"], rows[:, 6], c["sub"], np.zeros(n3),
                           np.zeros(n3))
    c["L"] = out[0]
    return out


def _rows(n, x0=4.0, y=4.0, dx=0.1, dt=1e-3, P=2000.0):
    r = np.zeros((n, 7))
    r[:, 0] = x0 + dx * (np.arange(n) + 1)
    r[:, 1] = y
    r[:, 3] = r[:, 4] = 1
    r[:, 5] = dt
    r[:, 6] = P
    return r


def _fields(L):
    return {k: np.array(v, np.float64) for k, v in
            (("T1", L[1]["T0"]), ("T2", L[2]["T0"]), ("T3", L[3]["T0"]),
             ("Tp2", L[2]["Tprime0"]), ("Tp3", L[3]["Tprime0"]))}


def _rel(a, b, scale):
    return float(np.max(np.abs(a - b)) / scale)


# ---------------------------------------------------------------------------

def test_uniform_field_stays_uniform(oracle):
    O = oracle
    Tu = 500.0
    inp, _ = W.tiny(n=(8, 8, 4), n2=2, n3=2, props={
        "laser_power": 0.0, "h_conv": 0.0, "emissivity": 0.0, "evaporation_coefficient": 0.0})
    for f in ("x", "y", "z"):
        inp["Level1"]["conditions"][f] = [Tu, Tu]
    c = _ctx(O, inp)
    for k in (1, 2, 3):
        c["L"][k]["T0"] = np.full(c["L"][k]["nn"], Tu)
    rows = _rows(4, P=0.0)
    _move(O, c, rows[0])
    _step(O, c, rows[0])
    _move(O, c, rows[1])
    _subcycle(O, c, rows)
    _move(O, c, np.array([5.0, 3.0, 0, 1, 1, 1e-3, 0.0]))  # L3 moves 2 and -2 L2 spacings
    f = _fields(c["L"])
    for k in ("T1", "T2", "T3"):
        assert np.max(np.abs(f[k] - Tu)) < 1e-9 * Tu, k
    for k in ("Tp2", "Tp3"):
        assert np.max(np.abs(f[k])) < 1e-9 * Tu, k


def _warm_state(O, n2=2, n3=2, steps=2, P=2000.0):
    """A state with non-zero T' on every level: one single step and
    ``steps`` subcycled L1 steps of a moving laser."""
    inp, _ = W.tiny(n=(8, 8, 4), n2=n2, n3=n3, props={"laser_power": P, "h_conv": 0.0})
    c = _ctx(O, inp)
    x = 4.0
    r = _rows(1, x0=x, P=P)[0]
    _move(O, c, r)
    _step(O, c, r)
    for _ in range(steps):
        rows = _rows(n2 * n3, x0=x, P=P)
        x = rows[-1, 0]
        _move(O, c, rows[0])
        _subcycle(O, c, rows)
    return c, rows


def test_remove_to_same_position_changes_nothing(oracle):
    O = oracle
    c, rows = _warm_state(O)
    before = _fields(c["L"])
    assert np.max(np.abs(before["Tp3"])) > 1.0 and np.max(np.abs(before["Tp2"])) > 1.0
    _move(O, c, rows[0])
    after = _fields(c["L"])
    rise = np.max(before["T3"]) - 300.0
    for k in before:
        assert _rel(after[k], before[k], rise) < 1e-10, k


def _trilinear(x, y, z, a):
    return (a[0] + a[1] * x) * (a[2] + a[3] * y) * (a[4] + a[5] * z)


def _inside(ax, x, y, z, tol=1e-9):
    return ((x >= ax[0][0] - tol) & (x <= ax[0][-1] + tol) & (y >= ax[1][0] - tol)
            & (y <= ax[1][-1] + tol) & (z >= ax[2][0] - tol) & (z <= ax[2][-1] + tol))


def _mesh_xyz(ax):
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X.reshape(-1), Y.reshape(-1), Z.reshape(-1)


@pytest.mark.parametrize("v, s3, s2", [
    # laser offsets from the start position (4, 4, 0) of the tiny deck
    # (h1 = 1, h2 = 0.5 mm): the L3 shift counts L2 spacings, the L2 shift L1
    # spacings, both truncated toward zero after adding 0.01
    ((1.0, -1.0, 0.0), (2, -1, 0), (1, 0, 0)),
    ((0.74, 1.49, 0.0), (1, 2, 0), (0, 1, 0)),
    ((-1.99, 0.0, 0.0), (-3, 0, 0), (-1, 0, 0)),
])
def test_whole_spacing_move_carries_trilinear_fields(oracle, v, s3, s2):
    O = oracle
    inp, _ = W.tiny(n=(8, 8, 4), n2=2, n3=2)
    c = _ctx(O, inp)
    L = c["L"]
    a1, a2, a3 = (300.0, 2.0, 1.0, 0.5, 1.0, 0.3), (1.0, 0.4, 2.0, -0.3, 1.0, 0.2), \
        (-2.0, 0.1, 1.0, 0.6, 0.5, -0.4)
    old2 = [a.copy() for a in L[2]["node_coords"]]
    old3 = [a.copy() for a in L[3]["node_coords"]]
    L[1]["T0"] = _trilinear(*_mesh_xyz(L[1]["node_coords"]), a1)
    L[2]["Tprime0"] = _trilinear(*_mesh_xyz(old2), a2)
    L[3]["Tprime0"] = _trilinear(*_mesh_xyz(old3), a3)
    _move(O, c, c["vstart"] + np.array([v[0], v[1], v[2], 0, 0, 0, 0]))
    L = c["L"]
    for d in range(3):
        assert np.allclose(L[3]["node_coords"][d], old3[d] + s3[d] * L[2]["h"][d])
        hz = c["p"]["layer_height"] if d == 2 else L[1]["h"][d]
        assert np.allclose(L[2]["node_coords"][d], old2[d] + s2[d] * hz)
    x3, y3, z3 = _mesh_xyz(L[3]["node_coords"])
    x2, y2, z2 = _mesh_xyz(L[2]["node_coords"])
    e3 = (_trilinear(x3, y3, z3, a1) + _trilinear(x3, y3, z3, a2) * _inside(old2, x3, y3, z3)
          + _trilinear(x3, y3, z3, a3) * _inside(old3, x3, y3, z3))
    e2p = _trilinear(x2, y2, z2, a2) * _inside(old2, x2, y2, z2)
    e3p = _trilinear(x3, y3, z3, a3) * _inside(old3, x3, y3, z3)
    assert np.max(np.abs(L[3]["T0"] - e3)) < 1e-9 * np.max(np.abs(e3))
    assert np.max(np.abs(L[3]["Tprime0"] - e3p)) < 1e-9 * np.max(np.abs(e3p))
    assert np.max(np.abs(L[2]["Tprime0"] - e2p)) < 1e-9 * np.max(np.abs(e2p))
    assert np.max(np.abs(L[2]["T0"] - (_trilinear(x2, y2, z2, a1) + e2p))) < 1e-9 * 600
    # the moves really leave parts of the old windows (the outside rule acts)
    assert 0 < np.count_nonzero(e3p) < e3p.size


def _n1_pair(O, P=300.0):
    """Two copies of one warm state with N2 = N3 = 1, at the same window."""
    c, rows = _warm_state(O, n2=1, n3=1, steps=3, P=P)
    row = _rows(1, x0=rows[-1, 0], P=P)[0]
    _move(O, c, row)
    return c, copy.deepcopy(c), row


def _n1_error(O, planted=None):
    a, b, row = _n1_pair(O)
    _step(O, a, row)
    _subcycle(O, b, row[None, :])
    fa, fb = _fields(a["L"]), _fields(b["L"])
    rise = np.max(fa["T3"]) - 300.0
    return max(_rel(fa[k], fb[k], rise) for k in fa), fa, fb


def test_subcycle_with_one_substep_is_the_single_step(oracle):
    O = oracle
    err, fa, _ = _n1_error(O)
    # no node reaches the solidus, so the phase flags stay put (the two paths
    # hand L2 S1 to L1 before and after its update, cF:2547-2553 vs 3273-3278)
    assert np.max(fa["T3"]) < 1533.0
    assert np.max(np.abs(fa["Tp3"])) > 1.0
    assert err < 1e-10, err


def test_subcycle_with_one_substep_planted_history_slip(oracle, monkeypatch):
    """The Part-2 L2 mass term of the oracle's subcycle uses the predictor's
    L3 T' history (cF:3420, 3502); plant a slip that uses the start-of-step
    T' instead (a zero mass term) and the N = 1 equality must break."""
    O = oracle
    orig = O.computeL2TprimeTerms_Part2

    def slip(Levels, ne_nn, L3Tp, L3Tp0, L3rc, dt, Shapes, L2V):
        return orig(Levels, ne_nn, L3Tp0, L3Tp0, L3rc, dt, Shapes, L2V)

    monkeypatch.setattr(O, "computeL2TprimeTerms_Part2", slip)
    err, _, _ = _n1_error(O)
    assert err > 1e-4, err
