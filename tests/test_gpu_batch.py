"""Batched instances (one launch per step for many L3 windows of their own
mesh and laser) against the same steps launched one instance at a time, and
one instance against the oracle's L3 step (computeSolutions_L3, cF:3015-3054,
with the faces held at their values and the T_amb floor)."""
import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import batch as B
    import _native as N
    N.init_device()
    return B


SIZES = [(12, 10, 8), (20, 16, 10), (8, 8, 4), (40, 36, 20), (100, 100, 50), (16, 70, 6)]


def _instances(n, seed, sizes=SIZES):
    return W.batch_instances(n=n, seed=seed, sizes=sizes)


@pytest.mark.parametrize("tile", [(1, 2), (1, 3), (1, 4), (2, 2), (2, 3)])
def test_batched_equals_unbatched(B, tile):
    """Every instance of the batched launch equals its own launch bitwise
    (same kernel body), T and S1, over steps that melt."""
    inst = _instances(8, seed=1)
    a = B.Batch.from_instances(inst)
    b = B.Batch.from_instances(inst)
    a.set_tile(*tile)
    b.set_tile(*tile)
    for s in range(30):
        rows = W.batch_rows(inst, s)
        a.step(rows, batched=True)
        b.step(rows, batched=False)
    melted = 0
    for i in range(len(inst)):
        Ta, Tb = a.T(i), b.T(i)
        print("instance", i, inst[i]["elements"], "T max", float(Ta.max()), "S1", int(a.S1(i).sum()))
        assert np.array_equal(Ta, Tb), (i, float(np.max(np.abs(Ta - Tb))))
        assert np.array_equal(a.S1(i), b.S1(i)), i
        melted += int(a.S1(i).sum() > 0)
        assert np.all(np.isfinite(Ta)) and Ta.max() > 300.0
    assert melted > 0


def test_batched_equals_unbatched_past_64_instances(B):
    """A batch of more than 64 instances (the block-to-instance lookup then
    spans more than one group of 64 block starts), small meshes of mixed
    sizes: every instance equals its own launch bitwise."""
    sizes = [(8, 8, 4), (12, 10, 8), (16, 12, 6), (10, 14, 4)]
    inst = _instances(131, seed=5, sizes=sizes)
    a = B.Batch.from_instances(inst)
    b = B.Batch.from_instances(inst)
    for tile in ((1, 3), (2, 3)):
        a.set_tile(*tile)
        b.set_tile(*tile)
        for s in range(8):
            rows = W.batch_rows(inst, s)
            a.step(rows, batched=True)
            b.step(rows, batched=False)
    for i in range(len(inst)):
        Ta, Tb = a.T(i), b.T(i)
        assert np.array_equal(Ta, Tb), (i, float(np.max(np.abs(Ta - Tb))))
        assert np.array_equal(a.S1(i), b.S1(i)), i
        assert Ta.max() > 300.0, i


def test_batch_brick_shapes_agree(B):
    """Every brick shape computes the same nodes with the same arithmetic:
    equal to FP32 rounding (1e-6 of the rise)."""
    inst = _instances(6, seed=2)
    runs = []
    for tile in B.Batch.TILES:
        bt = B.Batch.from_instances(inst)
        bt.set_tile(*tile)
        for s in range(6):
            bt.step(W.batch_rows(inst, s))
        runs.append([bt.T(i) for i in range(len(inst))])
    for i in range(len(inst)):
        rise = float(runs[1][i].max() - 300.0)
        for r in range(len(runs)):
            assert float(np.max(np.abs(runs[r][i] - runs[1][i]))) <= 1e-6 * rise


def test_batch_autotune_keeps_a_valid_depth(B):
    inst = _instances(6, seed=3)
    bt = B.Batch.from_instances(inst)
    bt.step(W.batch_rows(inst, 0))
    T0 = [bt.T(i) for i in range(len(inst))]
    tile, ms = bt.autotune(iters=5)
    assert tile in B.Batch.TILES and np.all(ms > 0) and ms[B.Batch.TILES.index(tile)] == ms.min()
    # tuning does not advance the state
    for i in range(len(inst)):
        assert np.array_equal(bt.T(i), T0[i])


def _oracle_level(O, el, h):
    lv = {"elements": list(el), "bounds": {"x": [0.0, el[0] * h], "y": [0.0, el[1] * h],
                                           "z": [-el[2] * h, 0.0]}}
    lv["length"], lv["h"] = O.calc_length_h(lv)
    lv["nodes"] = O.calcNumNodes(lv["elements"])
    lv["ne"] = int(np.prod(el))
    lv["nn"] = int(np.prod(lv["nodes"]))
    lv["BC"] = O.getBCindices(lv)
    b = lv["bounds"]
    lv["node_coords"], lv["connect"] = O.createMesh3D(
        (b["x"][0], b["x"][1], lv["nodes"][0]), (b["y"][0], b["y"][1], lv["nodes"][1]),
        (b["z"][0], b["z"][1], lv["nodes"][2]))
    return lv


@pytest.mark.parametrize("table", [False, True])
def test_batch_instance_vs_oracle(B, oracle, table):
    """Instances of the batch against the FP64 oracle's L3 step with the same
    faces, source row and floor: relative max |dT| 1e-4 of the rise after 8
    steps (the engine's L3 steps meet the same bound)."""
    sizes = [(12, 10, 8), (16, 12, 6)]
    inst = _instances(2, seed=4, sizes=sizes)
    for b_ in inst:
        b_["properties"]["emissivity"] = 0.4
        b_["properties"]["laser_power"] = b_["P"]
        if table:
            b_["properties"]["thermal_conductivity_bulk_table"] = W.K_TABLE
            b_["properties"]["heat_capacity_solid_table"] = W.CP_TABLE
    bt = B.Batch.from_instances(inst)
    O = oracle
    nsteps = 8
    for s in range(nsteps):
        bt.step(W.batch_rows(inst, s))
    for i, b_ in enumerate(inst):
        p = O.SetupProperties(b_["properties"])
        lv = _oracle_level(O, b_["elements"], b_["h"])
        nn, ne = lv["nn"], lv["ne"]
        T = np.full(nn, p["T_amb"])
        S1 = np.zeros(nn)
        faces = T.copy()
        for s in range(nsteps):
            row = W.batch_rows(inst, s)[i].astype(np.float64)
            S1, _, k, rc = O.computeStateProperties(T, S1, p, 0)
            F = O.computeSourcesL3(lv, row[:3], (0, 0, 0, 0, nn), p, row[6])
            F = O.computeConvRadBC(lv, T, ne, nn, p, F)
            T = O.solveMatrixFreeFE(lv, nn, ne, k, rc, row[5], T, F, 0)
            T = O._clamp(O.assignBCsFine(T, faces, lv["BC"]), p)
        Tg = bt.T(i).astype(np.float64)
        rise = float(np.max(T) - p["T_amb"])
        assert rise > 100.0
        assert float(np.max(np.abs(Tg - T))) / rise < 1e-4, (i, float(np.max(np.abs(Tg - T))) / rise)


def test_jump_state_keeps_melted_nodes(B):
    """A caller's multi-step update written over an instance's T (an operator
    jump) keeps the melted flag of every node that is above T_liquidus at the
    jump's start or end (gm_batch_jump_state): where the plain steps over the
    same span have flagged a node and the field is above T_liquidus, the
    jumped state has it flagged too; nowhere else does the flag change."""
    import _native as N
    inst = _instances(3, seed=4, sizes=[(40, 36, 20)])
    a = B.Batch.from_instances(inst)
    b = B.Batch.from_instances(inst)
    for s in range(15):
        rows = W.batch_rows(inst, s)
        a.step(rows)
        b.step(rows)
    T0 = [a.T(i) for i in range(3)]
    S0 = [a.S1(i) for i in range(3)]
    for s in range(15, 25):
        b.step(W.batch_rows(inst, s))
    grew = 0
    for i in range(3):
        T1 = b.T(i)
        a.set_fields(i, T=T1)
        dev = N.DeviceArray.from_numpy(T0[i])
        N.check(N.lib().gm_batch_jump_state(a.h, i, dev.ptr))
        N.check(N.lib().gm_batch_refresh(a.h))
        S = a.S1(i)
        Tl = float(inst[i]["properties"]["T_liquidus"])
        want = (S0[i] != 0) | (T0[i] >= Tl) | (T1 >= Tl)
        assert np.array_equal(S != 0, want), i
        both = (b.S1(i) != 0) & (T1 >= Tl)
        assert np.all(S[both] != 0), i
        grew += int(np.count_nonzero(want & (S0[i] == 0)))
        del dev
    assert grew > 0  # the span melts nodes that were powder at its start
