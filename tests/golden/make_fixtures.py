"""Writes the golden fixtures of the full-length driver parity tests.

Runs the CPU oracle's driver (oracle/gomelt_oracle.py, ``go_melt``: the
reference time loop gm:134-493 over subcycleGOMELT cF:3224-3632, stepGOMELT
cF:2304-2397 and moveEverything cF:2400-2510) on a deck and stores its final
state in a compressed .npz under tests/golden/:

  single_track  the plain 3-level run: 40x40x20 elements on every level over
                10x10x5 / 2x2x1 / 0.5x0.5x0.25 mm, constant properties,
                one 5 mm track at 1 m/s, P = 200 W (5,000 toolpath rows,
                50 L1 steps: one single-step chunk, then 49 subcycled);
  production    the first 300 rows of the production deck (L1 200x200x50,
                L2 and L3 100x100x50 elements, 8-knot k(T) and cp(T) tables,
                radiation): one single-step chunk and two subcycled L1 steps.

Usage:  python tests/golden/make_fixtures.py DECK PRECISION [MAX_ROWS]
        (PRECISION float32 emulates the reference's arithmetic, float64 is
        the FP64 oracle; the file is tests/golden/DECK_PRECISION.npz).

Stored (fields as float32, so both precisions' files have the same size):
T0 of L1, L2, L3; Tprime0 of L3; S1 of L3; the window origins of L2 and L3;
the chunk modes; Level0 accum and max_accum as (index, value) pairs of their
non-zero entries; the oracle's melt-pool size of the final L3 field.
"""

import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "go-melt_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import workloads as W  # noqa: E402
from oracle import gomelt_oracle as O  # noqa: E402


def deck_rows(name):
    """The deck (solver_input with absolute paths) and its toolpath rows."""
    inp, gcode = {"single_track": W.single_track, "production": W.production}[name]()
    d = tempfile.mkdtemp(prefix=f"gomelt_{name}_")
    with open(W.write_deck(inp, gcode, d)) as fh:
        deck = json.load(fh)
    p = O.SetupProperties(deck["properties"])
    L = O.SetupLevels(deck, p)
    nm = O.SetupNonmesh(dict(deck["nonmesh"]))
    _, text = O.parsingGcode(dict(nm, toolpath=None), p, L[2]["h"])
    rows = np.array([[float(v) for v in ln.split(",")] for ln in text.splitlines()])
    return deck, rows


def main():
    name, prec = sys.argv[1], sys.argv[2]
    max_rows = int(sys.argv[3]) if len(sys.argv) > 3 else None
    deck, rows = deck_rows(name)
    O.set_precision(prec)
    t0 = time.time()

    def log(ti, Levels):
        print(f"{name} {prec}: row {ti} at {time.time() - t0:.0f} s", flush=True)

    Lo, info = O.go_melt(deck, rows=rows, max_rows=max_rows, log=log, log_rows=True)
    L0, L3 = Lo[0], Lo[3]
    out = {"rows": np.int64(min(len(rows), max_rows or len(rows))),
           "chunks": np.array([c == "single" for c in info["chunks"]], np.uint8)}
    for k in (1, 2, 3):
        out[f"T0_{k}"] = np.asarray(Lo[k]["T0"], np.float32)
    out["Tp_3"] = np.asarray(L3["Tprime0"], np.float32)
    out["S1_3"] = np.asarray(L3["S1"]).astype(np.uint8)
    for k in (2, 3):
        out[f"origin_{k}"] = np.array([Lo[k]["node_coords"][d][0] for d in range(3)],
                                      np.float64)
    for key in ("accum", "max_accum"):
        a = np.asarray(info[key])
        nz = np.flatnonzero(a)
        out[f"{key}_idx"] = nz.astype(np.int64)
        out[f"{key}_val"] = a[nz].astype(np.float32)
    Tl = float(deck["properties"]["T_liquidus"])
    out["pool"] = np.array(O.melt_pool_dims(L3["T0"], L3["node_coords"], Tl)[:4], np.float64)
    path = os.path.join(HERE, f"{name}_{prec}.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} after {time.time() - t0:.0f} s; pool {out['pool']}")


if __name__ == "__main__":
    main()
