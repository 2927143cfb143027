"""CPU tests of the host side: the C ABI library's exports, the toolpath
writer (createPath.py, cP:6-188) and the host set-up metadata (cF:115-579).

No test here calls into the GPU: the library is loaded and its symbols
resolved, nothing is launched.
"""

import os
import re

import numpy as np
import pytest

import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gomelt.h")


def _header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(gm_\w+)\s*\(", text,
                                 flags=re.M)))


def test_header_declares_the_binding_table():
    import _native as N
    assert _header_functions() == sorted(N.SYMBOLS)


def test_library_exports_every_declared_symbol():
    import ctypes
    import _native as N
    if not os.path.exists(N.LIB_PATH):
        pytest.skip("libgomelt.so is not built (make -C go-melt_amd/csrc)")
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [s for s in _header_functions() if not hasattr(lib, s)]
    assert not missing, missing


def test_native_fails_loudly_without_library(monkeypatch, tmp_path):
    import _native as N
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(N, "_lib", None)
    with pytest.raises(N.GoMeltNativeError):
        N.lib()


# ---------------------------------------------------------------------------
# toolpath (cP:6-188)
# ---------------------------------------------------------------------------

def _toolpath_both(oracle, inp, gcode, tmp_path, **nm_over):
    import computeFunctions as CF
    import createPath as CP
    g = tmp_path / "t.gcode"
    g.write_text(gcode)
    nm_in = dict(inp["nonmesh"], gcode=str(g), toolpath=str(tmp_path / "tp.txt"),
                 save_path=str(tmp_path / "res") + "/", **nm_over)
    nm = CF.SetupNonmesh(nm_in)
    p = CF.SetupProperties(inp["properties"])
    n = CP.parsingGcode(nm, p, None)
    text = (tmp_path / "tp.txt").read_text()
    on = oracle.SetupNonmesh(dict(nm_in, toolpath=None))
    op = oracle.SetupProperties(inp["properties"])
    n_o, text_o = oracle.parsingGcode(on, op, None, gcode_text=gcode)
    return n, text, n_o, text_o


def test_toolpath_single_track_counts_and_format(oracle, tmp_path):
    """A 5 mm track at 1000 mm/s and dt 1e-6: 5 mm / 1 um = 5000 rows of
    Ljump=1, Ldwell=1, P=200; the last row carries the remainder of the
    float division (5.0 // 0.001 = 4999 full rows)."""
    inp, gcode = W.single_track()
    n, text, n_o, text_o = _toolpath_both(oracle, inp, gcode, tmp_path)
    assert text == text_o and n == n_o
    lines = text.splitlines()
    assert n == len(lines) == 5000
    rows = np.array([[float(v) for v in ln.split(",")] for ln in lines])
    assert np.all(rows[:, 3] == 1) and np.all(rows[:, 4] == 1)
    assert np.all(rows[:, 6] == 200.0)
    np.testing.assert_allclose(rows[:-1, 5], 1e-6)
    assert 0.99e-6 < rows[-1, 5] <= 1e-6
    np.testing.assert_allclose(rows[-1, 0], 7.5, atol=1e-9)
    np.testing.assert_allclose(np.diff(rows[:-1, 0]), 1e-3, atol=1e-9)
    # fixed width: x, y, z as %.8e right-justified to 15 characters
    for ln in lines[:5] + lines[-5:]:
        f = ln.split(",")
        assert len(f) == 7 and all(len(v) == 15 for v in f[:3])
        assert re.fullmatch(r"\s*-?\d\.\d{8}e[+-]\d\d", f[0])


def test_toolpath_raster_matches_oracle(oracle, tmp_path):
    """Serpentine raster with G0 travel moves: travel rows have Ljump=0, P=0."""
    inp, gcode = W.production()
    n, text, n_o, text_o = _toolpath_both(oracle, inp, gcode, tmp_path)
    assert text == text_o and n == n_o
    rows = np.array([[float(v) for v in ln.split(",")] for ln in text.splitlines()])
    on = rows[:, 3] == 1
    assert np.all(rows[on, 6] == 250.0) and np.all(rows[~on, 6] == 0.0)
    assert on.sum() > 0 and (~on).sum() > 0


def test_toolpath_layers_and_dwell_match_oracle(oracle, tmp_path):
    """A z change inserts dwell rows (Ldwell=0, P=0): wait_time rows of dt3,
    then the rest in rows of dt3 * N2 * N3 and one remainder row."""
    inp, _ = W.single_track()
    gcode = ("G0 X2.5 Y5 Z0\nG1 X3.0 Y5 Z0\nG0 X3.0 Y5.5 Z0.05\nG1 X2.5 Y5.5 Z0.05\n")
    n, text, n_o, text_o = _toolpath_both(oracle, inp, gcode, tmp_path, dwell_time=3.3e-4,
                                          wait_time=20)
    assert text == text_o and n == n_o
    rows = np.array([[float(v) for v in ln.split(",")] for ln in text.splitlines()])
    dw = rows[:, 4] == 0
    assert dw.any() and np.all(rows[dw, 6] == 0.0)


# ---------------------------------------------------------------------------
# host set-up metadata (cF:115-579)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("deck", ["single_track", "production", "tiny"])
def test_setup_metadata_matches_oracle(oracle, deck, tmp_path):
    import computeFunctions as CF
    inp, _ = getattr(W, deck)()
    p = CF.SetupProperties(inp["properties"])
    po = oracle.SetupProperties(inp["properties"])
    for k in ("k_powder", "h_conv", "CM_coeff", "CT_coeff", "CP_coeff", "laser_eta"):
        assert p[k] == po[k]
    L = CF.setup_levels_host(inp, p)
    Lo = oracle.SetupLevels(inp, po)
    for l in (1, 2, 3):
        for k in ("elements", "nodes", "ne", "nn"):
            assert L[l][k] == Lo[l][k], (l, k)
        for d in range(3):
            np.testing.assert_allclose(L[l]["node_coords"][d], Lo[l]["node_coords"][d],
                                       rtol=1e-6, atol=1e-6)
        for f in range(6):
            np.testing.assert_array_equal(L[l]["BC"][f], Lo[l]["BC"][f])
    for l in (2, 3):
        for d in range(3):
            np.testing.assert_array_equal(L[l]["orig_overlap_nodes"][d],
                                          Lo[l]["orig_overlap_nodes"][d])
        assert L[l]["bounds"]["ix"] == Lo[l]["bounds"]["ix"]
    for k in ("elements", "nodes", "layer_idx_delta"):
        assert L[0][k] == Lo[0][k]
    for d in range(3):
        np.testing.assert_array_equal(L[0]["orig_overlap_nodes"][d],
                                      Lo[0]["orig_overlap_nodes"][d])
        np.testing.assert_array_equal(L[0]["orig_overlap_nodes_L2"][d],
                                      Lo[0]["orig_overlap_nodes_L2"][d])
    assert CF.getSubstrateNodes(L) == oracle.getSubstrateNodes(Lo)
    v = [0.0, 0.0, float(inp["properties"].get("laser_center", [0, 0, 0])[2])
         if inp["properties"].get("laser_center") else 0.0]
    assert CF.calcStaticTmpNodesAndElements(L, v) == oracle.calcStaticTmpNodesAndElements(Lo, v)


def test_c1_deck_is_the_documented_workload():
    """Single-track deck: 40x40x20 elements per level, h = 250/50/12.5 um,
    nn = 35,301 and ne = 32,000 per level."""
    inp, _ = W.single_track()
    for name, h in (("Level1", 0.25), ("Level2", 0.05), ("Level3", 0.0125)):
        lv = inp[name]
        assert lv["elements"] == [40, 40, 20]
        ext = lv["bounds"]["x"][1] - lv["bounds"]["x"][0]
        assert abs(ext / 40 - h) < 1e-12
    assert 41 * 41 * 21 == 35301


# ---------------------------------------------------------------------------
# device code shape
# ---------------------------------------------------------------------------

def test_device_code_has_no_calls_or_recursion(tmp_path):
    """Every device helper must inline into its kernel: an out-of-line call
    puts a stack in scratch memory, and an accidental self-call (a helper
    that calls itself) overflows it at run time.  Checked on the gfx950
    assembly of both HIP sources."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    csrc = os.path.join(ROOT, "go-melt_amd", "csrc")
    for src in ("gm_kernels.hip", "gm_engine.hip"):
        out = tmp_path / (src + ".s")
        subprocess.run([hipcc, "--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops", "--offload-device-only",
                        "-S", os.path.join(csrc, src), "-o", str(out)],
                       check=True, capture_output=True)
        asm = out.read_text()
        assert "s_swappc_b64" not in asm, f"{src}: device function call"
        assert "has_recursion, 1" not in asm, f"{src}: recursive device function"


def test_operator_loop_rejects_mixed_window_sizes():
    """run_in_loop / make_windows take one window shape for the whole batch;
    mixed mesh sizes are refused on the host before any launch."""
    import surrogate as S
    same = W.batch_instances(n=3, seed=1, sizes=[(32, 32, 16)])
    S.check_one_mesh_size(same)
    mixed = W.batch_instances(n=8, seed=0)
    assert len({tuple(b["elements"]) for b in mixed}) > 1
    with pytest.raises(ValueError, match="one mesh size"):
        S.check_one_mesh_size(mixed)
