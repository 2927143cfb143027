"""Dwell steps, per-layer output files and checkpoint resume of the driver.

* Dwell (go_melt.py:358-379; stepGOMELTDwellTime cF:2617-2664): a two-layer
  deck whose layer change inserts laser-off dwell rows (createPath cP:61-104)
  longer than ``wait_time``; past it the driver clears L2/L3 T' and advances
  L1 alone.  The whole run is compared with the oracle driver.
* Layer files (go_melt.py:243-261, 498-516): Level0 state and melt-time map
  at each layer change and at the end, the final temperature fields.
* Resume (go_melt.py:109-129, 390-411): the checkpoint written after the
  second layer starts holds the reference's five items; a run started from
  it (``layer_num`` = 1) ends where the uninterrupted run ends.
"""
import json
import os
import pickle

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _dwell_deck():
    inp, _ = W.tiny(n=(8, 8, 4), n2=2, n3=3, track=1.6)
    inp["Level1"]["elements"] = [8, 8, 6]
    inp["Level1"]["bounds"]["z"] = [-4, 2]
    inp["properties"]["layer_height"] = 1.0
    # 3 laser-off rows of 1 ms allowed, then dwell steps of N2*N3 ms
    inp["nonmesh"]["wait_time"] = 3
    inp["nonmesh"]["dwell_time"] = 0.02
    g = "G0 X4 Y4 Z0\nG1 X5.6 Y4 Z0\nG0 X5.6 Y4.5 Z1\nG1 X4 Y4.5 Z1\n"
    return inp, g


def _write(inp, g, d):
    with open(W.write_deck(inp, g, str(d))) as fh:
        return json.load(fh)


def _relT(a, b, Tamb=300.0):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b - Tamb)))


def test_dwell_steps_vs_oracle(oracle, tmp_path):
    import go_melt as G
    from createPath import read_rows
    inp, g = _dwell_deck()
    deck = _write(inp, g, tmp_path)
    st = {}
    Lg = G.go_melt(deck, quiet=True, stats=st)  # writes the toolpath (cP:6-188)
    rows = read_rows(deck["nonmesh"]["toolpath"])
    assert np.count_nonzero(rows[:, 4] == 0) > inp["nonmesh"]["wait_time"]
    Lo, info = oracle.go_melt(deck, rows=rows)
    assert st["chunks"] == info["chunks"]
    assert st["dwell"] == info["dwell"] and st["dwell"] >= 3, (st["dwell"], info["dwell"])
    assert st["subcycle"] >= 2
    for k in (1, 2, 3):
        a = np.asarray(Lg[k]["T0"], np.float64)
        assert _relT(a, Lo[k]["T0"]) < 5e-4, (k, _relT(a, Lo[k]["T0"]))
    for k in (2, 3):
        a = np.asarray(Lg[k]["Tprime0"], np.float64)
        b = Lo[k]["Tprime0"]
        rise = np.max(np.abs(Lo[k]["T0"] - 300.0))
        assert float(np.max(np.abs(a - b))) < 5e-4 * rise, k


def _vtr_arrays(path):
    """Decode a raw-appended VTK RectilinearGrid file (UInt64 headers)."""
    import re
    blob = open(path, "rb").read()
    head, data = blob.split(b'<AppendedData encoding="raw">\n_', 1)
    head = head.decode()
    assert 'type="RectilinearGrid" version="1.0"' in head
    assert 'header_type="UInt64"' in head and 'byte_order="LittleEndian"' in head
    out = {}
    types = {"Float32": np.float32, "Float64": np.float64}
    for m in re.finditer(r'<DataArray type="(\w+)" Name="([^"]+)" NumberOfComponents="1" '
                         r'format="appended" offset="(\d+)"/>', head):
        t, name, off = m.group(1), m.group(2), int(m.group(3))
        nbytes = int(np.frombuffer(data[off:off + 8], "<u8")[0])
        out[name] = np.frombuffer(data[off + 8:off + 8 + nbytes], types[t])
    ext = re.search(r'WholeExtent="([\d ]+)"', head).group(1)
    return out, [int(v) for v in ext.split()], head


def test_layer_files_and_resume(tmp_path):
    import go_melt as G
    inp, g = _dwell_deck()
    inp["nonmesh"]["output_files"] = 0
    deck = _write(inp, g, tmp_path)
    save = deck["nonmesh"]["save_path"]
    run = G.TimeLoop(deck, quiet=True)
    for rows in run.chunks():
        if not run.step_chunk(rows):
            break
    run.finish()
    full = {k: np.asarray(run.Levels[k]["T0"]).copy() for k in (1, 2, 3)}
    full_acc = np.asarray(run.Levels[0]["accum_time"]).copy()
    # per-layer and final files (go_melt.py:243-261, 498-516)
    for name in ("Level0_00000000.vtr", "Level0_00000001.vtr", "accum_time0000.npz",
                 "accum_time0001.npz", "FinalTemperatureFields.npz"):
        assert os.path.exists(save + name), name
    arr, ext, head = _vtr_arrays(save + "Level0_00000001.vtr")
    n0 = run.Levels[0]["nodes"]
    assert ext == [0, n0[0] - 1, 0, n0[1] - 1, 0, n0[2] - 1]
    assert '<PointData Scalars="State (Powder/Solid)">' in head
    S = arr["State (Powder/Solid)"]
    assert S.dtype == np.float32 and S.size == run.Levels[0]["nn"]
    # x-fastest node order, as reshape(nz, ny, nx).transpose(2, 1, 0) in
    # Fortran order gives (cF:2052-2054)
    assert np.array_equal(S, np.asarray(run.Levels[0]["S1"]).astype(np.float32))
    for d, key in enumerate(("x_coordinates", "y_coordinates", "z_coordinates")):
        assert np.allclose(arr[key], run.Levels[0]["node_coords"][d])
    fin = np.load(save + "FinalTemperatureFields.npz")
    assert sorted(fin.files) == ["L1T", "L2T", "L3T"]
    assert np.array_equal(fin["L3T"], full[3])
    acc1 = np.load(save + "accum_time0001.npz")["accum_time"]
    assert np.array_equal(acc1, np.maximum(full_acc, np.asarray(run.Levels[0]["max_accum_time"])))
    # the checkpoint of layer 1 (go_melt.py:390-401)
    ck = save + "checkpoint/Checkpoint0001.pkl"
    with open(ck, "rb") as fh:
        items = pickle.load(fh)
    assert len(items) == 5
    Ls, acc, macc, time_inc, record_inc = items
    assert isinstance(Ls, list) and len(Ls) == 4 and acc.size == run.Levels[0]["nn"]
    assert 0 < time_inc < run.time_inc
    # resume from it (go_melt.py:109-129): layer_num = 1
    deck2 = json.loads(json.dumps(deck))
    deck2["nonmesh"]["layer_num"] = 1
    run2 = G.TimeLoop(deck2, quiet=True)
    assert run2.time_inc == time_inc
    for rows in run2.chunks():
        if not run2.step_chunk(rows):
            break
    run2.eng.sync()
    assert run2.time_inc == run.time_inc
    assert run2.stats["chunks"] == run.stats["chunks"][-len(run2.stats["chunks"]):]
    for k in (1, 2, 3):
        a = np.asarray(run2.Levels[k]["T0"], np.float64)
        rise = float(np.max(full[k]) - 300.0)
        # the restriction sums use FP64 atomics, so runs agree to FP32
        # rounding rather than bitwise
        assert float(np.max(np.abs(a - full[k]))) < 1e-5 * rise, k
    a = np.asarray(run2.Levels[0]["accum_time"], np.float64)
    assert float(np.max(np.abs(a - full_acc))) <= 2e-3 + 1e-6
