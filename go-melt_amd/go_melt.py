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
``TimeLoop`` exposes the loop one chunk at a time (the benchmark uses it).
"""

from __future__ import annotations

import copy
import json
import os
import pickle
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from computeFunctions import *  # noqa: E402,F401,F403
import computeFunctions as cf  # noqa: E402
from createPath import count_lines, parsingGcode  # noqa: E402


def _chunks(path, n, skip_rows=0):
    """Yield lists of up to n parsed rows from the toolpath file, from row
    ``skip_rows`` on: the rows have a fixed width, so a resumed run seeks
    ``skip_rows * line_length`` bytes (gm:124-126)."""
    with open(path, "r") as fh:
        if skip_rows:
            line_len = len(fh.readline())
            fh.seek(skip_rows * line_len)
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


class TimeLoop:
    """The set-up and per-chunk body of the reference driver (gm:16-493)."""

    level_names = ["L1", "L2", "L3"]

    def __init__(self, solver_input, quiet=False, write_outputs=True):
        """``write_outputs`` False skips the per-layer Level0 state and melt-time
        files (gm:243-261) and the final files; the timed benchmark loop uses
        it (the Level0 map of the production deck is 10^9 nodes)."""
        self.tstart = time.time()
        self.write_outputs = write_outputs
        self.say = (lambda *a, **k: None) if quiet else print
        self.quiet = quiet
        self.Properties = SetupProperties(solver_input.get("properties", {}))
        self.Levels = SetupLevels(solver_input, self.Properties)
        self.Nonmesh = SetupNonmesh(solver_input.get("nonmesh", {}))
        self.ne_nn = getStaticNodesAndElements(self.Levels)
        self.subcycle = getStaticSubcycle(self.Nonmesh)
        self.eng = cf._eng(self.Levels)
        nm = self.Nonmesh
        if nm["use_txt"]:
            self.total_t_inc = count_lines(nm["toolpath"])
        else:
            self.total_t_inc = parsingGcode(nm, self.Properties, self.Levels[2]["h"])
        if not self.Properties["laser_center"]:
            with open(nm["toolpath"], "r") as fh:
                self.laser_start = np.array([float(v) for v in fh.readline().split(",")])
        else:
            self.laser_start = np.array(self.Properties["laser_center"])
        self.LInterp = [("L1->L2",), ("L2->L3",)]
        self.time_inc = self.record_inc = self.wait_inc = 0
        self.t_output = 0.0
        saveResults(self.Levels, nm, int(self.time_inc / nm["record_step"]) + 1)
        self.laser_prev_z = float("inf")
        self.dwell_count = 0.0
        self.move_hist = [0, 0, 0]
        self.Shapes = self.tmp_ne_nn = self.substrate = None
        self.new_checkpoint = self.load_chkpt = False
        self.layer_check = nm["layer_num"] + nm["restart_layer_num"]
        self.stats = {"single": 0, "subcycle": 0, "dwell": 0, "chunks": [], "mm": 0.0}
        self.prev_row = None
        self.done = False
        self.start_row = 0
        if nm["layer_num"] > 0:
            self._resume()

    def _checkpoint_path(self, layer):
        d = Path(self.Nonmesh["save_path"] + "checkpoint").absolute()
        return d, d.joinpath(f"Checkpoint{str(layer).zfill(4)}.pkl")

    def _resume(self):
        """Load Checkpoint<layer_num>.pkl and continue after its last row
        (gm:115-129).  The payload is the reference's five items; Levels[0]
        also carries the engine's device state (window positions, fields of
        the three levels), which the reference keeps inside its Levels."""
        nm = self.Nonmesh
        self.say(f"Checkpoint loading for start of Layer {nm['layer_num']}")
        _, path = self._checkpoint_path(nm["layer_num"])
        with open(path, "rb") as fh:
            Ls, accum, max_accum, time_inc, record_inc = pickle.load(fh)
        self.eng.load_state(Ls[0]["_engine_state"])
        L = self.Levels
        L[0]["S1"] = np.asarray(Ls[0]["S1"])
        L[0]["accum_time"] = np.asarray(accum)
        L[0]["max_accum_time"] = np.asarray(max_accum)
        L[0]["node_coords"] = copy.deepcopy(Ls[0]["node_coords"])
        self.eng.sync_coords()
        self.load_chkpt = True
        self.time_inc += int(time_inc)
        self.record_inc = int(record_inc)
        self.start_row = int(time_inc)

    def _save_checkpoint(self):
        """gm:390-401: [Levels, accum_time, max_accum_time, time_inc,
        record_inc] after the chunk in which a new layer started."""
        nm, L = self.Nonmesh, self.Levels
        cdir, path = self._checkpoint_path(nm["layer_num"])
        os.makedirs(cdir, exist_ok=True)
        self.eng.sync()
        save_object([L, np.asarray(L[0]["accum_time"]), np.asarray(L[0]["max_accum_time"]),
                     self.time_inc, self.record_inc], path,
                    extra={"_engine_state": self.eng.save_state()})
        self.say("Saved Checkpoint")

    def _save_layer_outputs(self):
        """gm:243-261: Level0 state and the melt-time map of the layer that
        ends, written before Level0 shifts down one layer."""
        if not self.write_outputs:
            return
        L, nm = self.Levels, self.Nonmesh
        self.eng.sync()
        saveState(L[0], "Level0_", nm["layer_num"], nm["save_path"], 0)
        acc = np.maximum(np.asarray(L[0]["accum_time"]), np.asarray(L[0]["max_accum_time"]))
        np.savez(nm["save_path"] + "accum_time" + str(nm["layer_num"]).zfill(4), accum_time=acc)

    def chunks(self):
        return _chunks(self.Nonmesh["toolpath"], self.subcycle[2], self.start_row)

    # -- one chunk of N2*N3 rows ------------------------------------------
    def step_chunk(self, rows):
        """Advance the simulation over one chunk; returns False when the loop
        should stop (end of file or the checkpoint layer limit)."""
        L, P, nm, sub = self.Levels, self.Properties, self.Nonmesh, self.subcycle
        eng = self.eng
        t_loop = time.time()
        if not rows:
            return False
        ongoing = len(rows) == sub[2]
        la = np.array(rows, dtype=np.float64)
        la32 = la.astype(np.float32)
        jump = False
        if la.shape[0] > 1:
            jump = bool((np.abs(np.diff(la32, axis=0)[:, :2] / la32[:-1, 5].max())
                         > 100 * nm["laser_velocity"]).any())
        single = (bool(np.any(la[:, 2] != self.laser_prev_z)) or self.load_chkpt
                  or self.wait_inc > max(0, nm["wait_time"] - sub[0] * sub[1] * 2)
                  or not ongoing or la.shape[0] == 1 or jump)
        self.stats["chunks"].append("single" if single else "subcycle")
        keep_going = True
        if single:
            for pos in la:
                self._single_row(pos)
            if self.new_checkpoint:
                nm["layer_num"] += 1
                self.say(f"Saving checkpoint for Layer {nm['layer_num']}")
                self._save_checkpoint()
                if nm["layer_num"] == self.layer_check:
                    keep_going = False
                self.new_checkpoint = False
                self.load_chkpt = True
            else:
                self.load_chkpt = False
        else:
            self.wait_inc = (self.wait_inc + len(la) - la[:, 4].sum()
                             if (la[:, 4] == 0).any() else 0)
            _, self.Shapes, self.LInterp, self.move_hist = moveEverything(
                la[0, :], self.laser_start, L, self.move_hist, self.LInterp, eng.r12, eng.r23,
                P["layer_height"])
            self.run_subcycle(la)
            self.stats["subcycle"] += 1
            self.time_inc += la.shape[0]
            self.record_inc += la.shape[0]
        self.t_output += la[:, 5].sum()
        seg = la if self.prev_row is None else np.vstack([self.prev_row[None, :], la])
        d = np.hypot(np.diff(seg[:, 0]), np.diff(seg[:, 1]))
        self.stats["mm"] += float(d[seg[1:, 6] > 0].sum())
        self.prev_row = la[-1]
        if self.record_inc >= nm["record_step"]:
            self.record_inc = 0
            saveResults(L, nm, int(self.time_inc / nm["record_step"]) + 1)
        if nm["info_T"]:
            printLevelMaxMin(L, self.level_names)
        if not self.quiet:
            tend = time.time()
            t_now = 1000 * (tend - t_loop)
            t_avg = 1000 * (tend - self.tstart) / max(self.time_inc, 1)
            rem = ((tend - t_loop) / sub[2]) * (self.total_t_inc - self.time_inc) / 3600
            print("%d/%d, Real: %.6f s, Wall: %.2f s, Loop: %5.2f ms, Avg: %5.2f ms/dt"
                  % (self.time_inc, self.total_t_inc, self.t_output, tend - self.tstart,
                     t_now, t_avg))
            print("Laser location: X: %.2f, Y: %.2f, Z: %.2f" % (la[-1, 0], la[-1, 1], la[-1, 2]))
            print(f"Estimated execution time remaining: {rem:.4f} hours")
        return keep_going and ongoing

    def _single_row(self, pos):
        L, P, nm, eng = self.Levels, self.Properties, self.Nonmesh, self.eng
        self.wait_inc = self.wait_inc + 1 if pos[4] == 0 else 0
        if pos[2] != self.laser_prev_z and self.time_inc > 0 and not self.load_chkpt:
            self.new_checkpoint = True
        move_vert = False
        if pos[2] != self.laser_prev_z:
            tmp = copy.deepcopy(L[1]["orig_node_coords"])
            kidx = 0
            while not np.isclose(tmp[2] - pos[2], 0, atol=1e-4).any():
                tmp[2] = (tmp[2] + np.float32(P["layer_height"])).astype(np.float32)
                kidx += 1
            if not self.load_chkpt:
                self._save_layer_outputs()
                cf.N.check(cf.N.lib().gm_engine_layer_change(eng.h, kidx, float(tmp[2][0]),
                                                             float(pos[2]), 1))
                L[1]["node_coords"] = copy.deepcopy(tmp)
                L0 = L[0]
                L0["node_coords"][2] = (L0["orig_node_coords"][2] + pos[2]
                                        - L0["orig_node_coords"][2][-1]).astype(np.float32)
            self.tmp_ne_nn = calcStaticTmpNodesAndElements(L, pos)
            self.laser_prev_z = pos[2]
            self.wait_inc = 0
            move_vert = True
        _, self.Shapes, self.LInterp, self.move_hist = moveEverything(
            pos, self.laser_start, L, self.move_hist, self.LInterp, eng.r12, eng.r23,
            P["layer_height"])
        if move_vert:
            eng.sync_coords()
            self.substrate = getSubstrateNodes(L)
            cf.N.check(cf.N.lib().gm_engine_level0_substrate(eng.h, self.substrate[0]))
            self.say("Start of new layer")
        if self.wait_inc <= nm["wait_time"]:
            _, all_reset = stepGOMELT(L, self.ne_nn, self.tmp_ne_nn, self.Shapes, self.LInterp,
                                      pos, P, pos[5], pos[6], self.substrate)
            self.accum_single(pos, all_reset)
            self.stats["single"] += 1
            if nm["info_T"]:
                self.say(f"Step {self.time_inc + 1} / {self.total_t_inc}")
                printLevelMaxMin(L, self.level_names)
        else:
            tp2, tp3 = np.asarray(L[2]["Tprime0"]), np.asarray(L[3]["Tprime0"])
            if not (tp2 == 0).all() and not (tp3 == 0).all():
                self.dwell_count = nm["wait_time"] * nm["timestep_L3"]
                cf.N.check(cf.N.lib().gm_engine_clear_tprime(eng.h))
            stepGOMELTDwellTime(L, self.tmp_ne_nn, self.ne_nn, P, pos[5], self.substrate)
            self.dwell_count += pos[5]
            self.stats["dwell"] += 1
            self.say(f"Dwell Time {self.dwell_count:.6f} s")
        self.time_inc += 1
        self.record_inc += 1

    # -- melt-time bookkeeping (gm:339-357, 436-456) ------------------------
    # The engine keeps the melt-time maps on its own Level0 (subcycleGOMELT
    # without window arrays; stepGOMELT applies gm:339-357 on the device).
    # A driver that keeps its own maps, as the reference does, overrides
    # these two methods (tests/test_gpu_reference_driver.py runs the
    # reference's lines that way).
    def run_subcycle(self, la):
        subcycleGOMELT(self.Levels, self.ne_nn, self.Shapes, self.substrate, self.LInterp,
                       self.tmp_ne_nn, la, self.Properties, la[:, 6], self.subcycle)

    def accum_single(self, pos, all_reset):
        pass

    def finish(self):
        """Final output (gm:495-530)."""
        L, nm = self.Levels, self.Nonmesh
        self.eng.sync()
        self.eng.sync_coords()
        if not self.write_outputs:
            return
        saveState(L[0], "Level0_", nm["layer_num"], nm["save_path"], 0)
        saveResultsFinal(L, nm)
        np.savez(f"{nm['save_path']}FinalTemperatureFields", L1T=np.asarray(L[1]["T0"]),
                 L2T=np.asarray(L[2]["T0"]), L3T=np.asarray(L[3]["T0"]))
        acc = np.maximum(np.asarray(L[0]["accum_time"]), np.asarray(L[0]["max_accum_time"]))
        np.savez(nm["save_path"] + "accum_time" + str(nm["layer_num"]).zfill(4), accum_time=acc)
        self.say("End of simulation")


def go_melt(solver_input: dict, quiet: bool = False, stats: dict = None):
    """gm:16-530.  ``stats`` (optional dict) receives counters and timings."""
    run = TimeLoop(solver_input, quiet=quiet)
    t0 = time.time()
    for rows in run.chunks():
        if not run.step_chunk(rows):
            break
    run.eng.sync()
    run.stats["loop_s"] = time.time() - t0
    run.stats["rows"] = run.time_inc
    run.finish()
    if stats is not None:
        stats.update(run.stats)
    return run.Levels


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
