"""GO-MELT driver on MI355X: ``python3 go-melt_amd/go_melt.py [DEVICE_ID] [input.json]``.

Same command line, JSON schema and time-loop semantics as the reference's
go_melt/go_melt.py:
  * the toolpath is read N2*N3 rows (one L1 step) at a time (gm:140-157);
  * a chunk runs in single-step mode (equal dt on all levels, stepGOMELT)
    on the first chunk, at a layer change, after a checkpoint, at the end of
    the file, after a long laser-off wait or across a jump (gm:162-172);
  * otherwise the windows move once and the chunk is one subcycled L1 step
    (moveEverything + subcycleGOMELT, gm:413-459);
  * rows past ``wait_time`` laser-off steps run the L1-only dwell step
    (gm:358-379); layer changes shift the L1 z axis and Level0 (gm:198-290).
Fields stay on the GPU; the host only reads them for output and checks.
"""

from __future__ import annotations

import copy
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from computeFunctions import *  # noqa: E402,F401,F403
import computeFunctions as cf  # noqa: E402
from createPath import count_lines, parsingGcode  # noqa: E402


def _chunks(path, n):
    """Yield lists of up to n parsed rows from the toolpath file."""
    with open(path, "r") as fh:
        while True:
            rows = []
            for _ in range(n):
                ln = fh.readline()
                if not ln:
                    break
                rows.append([float(v) for v in ln.split(",")])
            yield rows
            if len(rows) < n:
                return


def go_melt(solver_input: dict, quiet: bool = False, stats: dict = None):
    """gm:16-530.  ``stats`` (optional dict) receives counters and timings."""
    tstart = time.time()
    say = (lambda *a, **k: None) if quiet else print
    level_names = ["L1", "L2", "L3"]
    Properties = SetupProperties(solver_input.get("properties", {}))
    Levels = SetupLevels(solver_input, Properties)
    Nonmesh = SetupNonmesh(solver_input.get("nonmesh", {}))
    ne_nn = getStaticNodesAndElements(Levels)
    subcycle = getStaticSubcycle(Nonmesh)
    eng = cf._eng(Levels)
    L1L2Eratio, L2L3Eratio = eng.r12, eng.r23

    if Nonmesh["use_txt"]:
        move_mesh = count_lines(Nonmesh["toolpath"])
    else:
        move_mesh = parsingGcode(Nonmesh, Properties, Levels[2]["h"])
    total_t_inc = move_mesh

    if not Properties["laser_center"]:
        with open(Nonmesh["toolpath"], "r") as fh:
            laser_start = np.array([float(v) for v in fh.readline().split(",")])
    else:
        laser_start = np.array(Properties["laser_center"])
    LInterp = [("L1->L2",), ("L2->L3",)]

    time_inc = record_inc = wait_inc = 0
    t_output = 0.0
    saveResults(Levels, Nonmesh, int(time_inc / Nonmesh["record_step"]) + 1)
    laser_prev_z = float("inf")
    dwell_count = 0.0
    move_hist = [0, 0, 0]
    Shapes = None
    tmp_ne_nn = substrate = None
    new_checkpoint = load_chkpt = False
    layer_check = Nonmesh["layer_num"] + Nonmesh["restart_layer_num"]
    st = {"single": 0, "subcycle": 0, "dwell": 0, "loop_s": 0.0, "rows": 0, "mm": 0.0,
          "chunks": []}
    prev_row = None
    N23 = subcycle[2]
    t_loop0 = time.time()
    for rows in _chunks(Nonmesh["toolpath"], N23):
        t_loop = time.time()
        if not rows:
            break
        ongoing = len(rows) == N23
        laser_all = np.array(rows, dtype=np.float64)
        t_add = laser_all.shape[0]
        la32 = laser_all.astype(np.float32)
        jump = False
        if t_add > 1:
            jump = bool((np.abs(np.diff(la32, axis=0)[:, :2] / la32[:-1, 5].max())
                         > 100 * Nonmesh["laser_velocity"]).any())
        single_step = (bool(np.any(laser_all[:, 2] != laser_prev_z)) or load_chkpt
                       or wait_inc > max(0, Nonmesh["wait_time"] - subcycle[0] * subcycle[1] * 2)
                       or not ongoing or t_add == 1 or jump)
        st["chunks"].append("single" if single_step else "subcycle")
        if single_step:
            for laser_pos in laser_all:
                wait_inc = wait_inc + 1 if laser_pos[4] == 0 else 0
                if laser_pos[2] != laser_prev_z and time_inc > 0 and not load_chkpt:
                    new_checkpoint = True
                move_vert = False
                if laser_pos[2] != laser_prev_z:
                    tmp = copy.deepcopy(Levels[1]["orig_node_coords"])
                    kidx = 0
                    while not np.isclose(tmp[2] - laser_pos[2], 0, atol=1e-4).any():
                        tmp[2] = (tmp[2] + np.float32(Properties["layer_height"])).astype(np.float32)
                        kidx += 1
                    if not load_chkpt:
                        cf.N.check(cf.N.lib().gm_engine_layer_change(
                            eng.h, kidx, float(tmp[2][0]), float(laser_pos[2]), 1))
                        Levels[1]["node_coords"] = copy.deepcopy(tmp)
                        L0 = Levels[0]
                        L0["node_coords"][2] = (L0["orig_node_coords"][2] + laser_pos[2]
                                                - L0["orig_node_coords"][2][-1]).astype(np.float32)
                    tmp_ne_nn = calcStaticTmpNodesAndElements(Levels, laser_pos)
                    laser_prev_z = laser_pos[2]
                    wait_inc = 0
                    move_vert = True
                Levels, Shapes, LInterp, move_hist = moveEverything(
                    laser_pos, laser_start, Levels, move_hist, LInterp, L1L2Eratio, L2L3Eratio,
                    Properties["layer_height"])
                if move_vert:
                    eng.sync_coords()
                    substrate = getSubstrateNodes(Levels)
                    cf.N.check(cf.N.lib().gm_engine_level0_substrate(eng.h, substrate[0]))
                    say("Start of new layer")
                if wait_inc <= Nonmesh["wait_time"]:
                    Levels, _ = stepGOMELT(Levels, ne_nn, tmp_ne_nn, Shapes, LInterp, laser_pos,
                                           Properties, laser_pos[5], laser_pos[6], substrate)
                    st["single"] += 1
                    if Nonmesh["info_T"]:
                        say(f"Step {time_inc + 1} / {total_t_inc}")
                        printLevelMaxMin(Levels, level_names)
                else:
                    tp2 = np.asarray(Levels[2]["Tprime0"])
                    tp3 = np.asarray(Levels[3]["Tprime0"])
                    if not (tp2 == 0).all() and not (tp3 == 0).all():
                        dwell_count = Nonmesh["wait_time"] * Nonmesh["timestep_L3"]
                        cf.N.check(cf.N.lib().gm_engine_clear_tprime(eng.h))
                    Levels = stepGOMELTDwellTime(Levels, tmp_ne_nn, ne_nn, Properties,
                                                 laser_pos[5], substrate)
                    dwell_count += laser_pos[5]
                    st["dwell"] += 1
                    say(f"Dwell Time {dwell_count:.6f} s")
                time_inc += 1
                record_inc += 1
            if new_checkpoint:
                Nonmesh["layer_num"] += 1
                say(f"Saving checkpoint for Layer {Nonmesh['layer_num']}")
                cdir = Path(Nonmesh["save_path"] + "checkpoint").absolute()
                os.makedirs(cdir, exist_ok=True)
                save_object([Levels, time_inc, record_inc],
                            cdir.joinpath(f"Checkpoint{str(Nonmesh['layer_num']).zfill(4)}.pkl"))
                if Nonmesh["layer_num"] == layer_check:
                    break
                new_checkpoint = False
                load_chkpt = True
            else:
                load_chkpt = False
        else:
            wait_inc = (wait_inc + len(laser_all) - laser_all[:, 4].sum()
                        if (laser_all[:, 4] == 0).any() else 0)
            Levels, Shapes, LInterp, move_hist = moveEverything(
                laser_all[0, :], laser_start, Levels, move_hist, LInterp, L1L2Eratio, L2L3Eratio,
                Properties["layer_height"])
            Levels, _, _, _, _, _ = subcycleGOMELT(
                Levels, ne_nn, Shapes, substrate, LInterp, tmp_ne_nn, laser_all, Properties,
                laser_all[:, 6], subcycle)
            st["subcycle"] += 1
            time_inc += t_add
            record_inc += t_add
        t_output += laser_all[:, 5].sum()
        if prev_row is not None:
            seg = np.vstack([prev_row[None, :], laser_all])
        else:
            seg = laser_all
        d = np.hypot(np.diff(seg[:, 0]), np.diff(seg[:, 1]))
        st["mm"] += float(d[seg[1:, 6] > 0].sum())
        prev_row = laser_all[-1]
        if record_inc >= Nonmesh["record_step"]:
            record_inc = 0
            saveResults(Levels, Nonmesh, int(time_inc / Nonmesh["record_step"]) + 1)
        if Nonmesh["info_T"]:
            printLevelMaxMin(Levels, level_names)
        tend = time.time()
        if not quiet:
            t_now = 1000 * (tend - t_loop)
            t_avg = 1000 * (tend - tstart) / max(time_inc, 1)
            rem = ((tend - t_loop) / subcycle[2]) * (total_t_inc - time_inc) / 3600
            print("%d/%d, Real: %.6f s, Wall: %.2f s, Loop: %5.2f ms, Avg: %5.2f ms/dt"
                  % (time_inc, total_t_inc, t_output, tend - tstart, t_now, t_avg))
            print("Laser location: X: %.2f, Y: %.2f, Z: %.2f"
                  % (laser_all[-1, 0], laser_all[-1, 1], laser_all[-1, 2]))
            print(f"Estimated execution time remaining: {rem:.4f} hours")
        if not ongoing:
            break
    eng.sync()
    st["loop_s"] = time.time() - t_loop0
    st["rows"] = time_inc
    eng.sync_coords()
    if Nonmesh["output_files"] == 1:
        saveState(Levels[0], "Level0_", Nonmesh["layer_num"], Nonmesh["save_path"], 0)
    saveResultsFinal(Levels, Nonmesh)
    np.savez(f"{Nonmesh['save_path']}FinalTemperatureFields",
             L1T=np.asarray(Levels[1]["T0"]), L2T=np.asarray(Levels[2]["T0"]),
             L3T=np.asarray(Levels[3]["T0"]))
    acc = np.maximum(np.asarray(Levels[0]["accum_time"]), np.asarray(Levels[0]["max_accum_time"]))
    np.savez(Nonmesh["save_path"] + "accum_time" + str(Nonmesh["layer_num"]).zfill(4),
             accum_time=acc)
    say("End of simulation")
    if stats is not None:
        stats.update(st)
    return Levels


if __name__ == "__main__":
    DEVICE_ID = int(sys.argv[1]) if len(sys.argv) > 1 and sys.argv[1].isdigit() else 0
    input_file = sys.argv[2] if len(sys.argv) > 2 else "examples/example.json"
    if len(sys.argv) <= 1:
        print("GPU ID not provided. Setting GPU to 0.")
    if len(sys.argv) <= 2:
        print("Input file not provided. Using default: 'examples/example.json'.")
    os.environ.setdefault("HIP_VISIBLE_DEVICES", str(DEVICE_ID))
    try:
        with open(input_file, "r") as read_file:
            solver_input = json.load(read_file)
    except FileNotFoundError:
        print(f"Error: Input file '{input_file}' not found.")
        sys.exit(1)
    print("Running GO-MELT")
    print(f"GPU: {DEVICE_ID}, Input File: {input_file}")
    go_melt(solver_input)
