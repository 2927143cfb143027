"""Input decks for the multilevel heat step, in the reference JSON schema.

Each builder returns ``(solver_input, gcode_text)``.  ``solver_input`` follows
examples/README.md of the reference (Level1/2/3, properties, nonmesh); the
G-code is what ``nonmesh['gcode']`` would point at.  ``write_deck`` writes both
to disk so that ``go_melt.py`` can run them from the command line.

* ``single_track`` -- plain 3-level run, constant properties, one 5 mm track.
  Levels 40x40x20 elements over 10x10x5 / 2x2x1 / 0.5x0.5x0.25 mm.
* ``production`` -- plain 3-level run with temperature-dependent k(T), cp(T)
  tables (8 knots, 300-3000 K), serpentine raster of 20 hatch lines of 5 mm at
  0.1 mm spacing, P = 250 W, r = 40 um, radiation (emissivity 0.4) plus
  convection.  L1 200x200x50 elements over 20x20x5 mm, L2 100x100x50,
  L3 100x100x50 over 1x1x0.5 mm.  L2 spans 5x5x2.5 mm so that the L1/L2 and
  L2/L3 spacing ratios are integers (2 and 5); a 4x4x2 mm L2 would give the
  ratio 2.5, which the reference's integer window bookkeeping (go_melt.py:42-48,
  computeFunctions.py:2453, 2472-2474) cannot represent.
* ``tiny`` -- small decks for parity tests (seconds in the oracle).
"""

from __future__ import annotations

import json
import os

import numpy as np


def _props(**kw):
    base = {
        "thermal_conductivity_powder": 20.0,
        "thermal_conductivity_bulk_a0": 20.0,
        "thermal_conductivity_bulk_a1": 0.0,
        "thermal_conductivity_fluid_a0": 20.0,
        "heat_capacity_solid_a0": 500.0,
        "heat_capacity_solid_a1": 0.0,
        "heat_capacity_mushy": 500.0,
        "heat_capacity_fluid": 500.0,
        "density": 8e-6,
        "laser_radius": 0.05,
        "laser_depth": 0.05,
        "laser_power": 200.0,
        "laser_absorptivity": 0.35,
        "T_amb": 300.0,
        "T_solidus": 1533.0,
        "T_liquidus": 1609.0,
        "T_boiling": 3038.0,
        "h_conv": 1e-5,
        "emissivity": 0.0,
        "evaporation_coefficient": 0.0,
        "boltzmann_constant": 1.38e-23,
        "atomic_mass": 9.746e-26,
        "latent_heat_evap": 6457000.0,
        "molar_mass": 58.69,
        "layer_height": 0.05,
    }
    base.update(kw)
    return base


def _nonmesh(**kw):
    base = {
        "timestep_L3": 1e-6,
        "subcycle_num_L2": 10,
        "subcycle_num_L3": 10,
        "dwell_time": 0.0,
        "wait_time": 0,
        "output_files": 0,
        "record_step": 100,
        "Level1_record_step": 1,
        "laser_velocity": 1000.0,
        "save_path": "./results/",
        "gcode": "./track.gcode",
        "toolpath": "./results/toolpath.txt",
        "layer_num": 0,
        "info_T": 0,
        "dwell_time_multiplier": 1,
        "use_txt": 0,
    }
    base.update(kw)
    return base


def _level(el, x, y, z, cond=None):
    d = {"elements": list(el), "bounds": {"x": list(x), "y": list(y), "z": list(z)}}
    if cond is not None:
        d["conditions"] = {"x": [cond, cond], "y": [cond, cond], "z": [cond, cond]}
    return d


def single_track():
    """5 mm straight track at 1 m/s, P = 200 W, 3 x 40x40x20 elements."""
    inp = {
        "Level1": _level((40, 40, 20), (0, 10), (0, 10), (-5, 0), 300.0),
        "Level2": _level((40, 40, 20), (1.5, 3.5), (4, 6), (-1, 0)),
        "Level3": _level((40, 40, 20), (2.25, 2.75), (4.75, 5.25), (-0.25, 0)),
        "properties": _props(laser_center=[2.5, 5.0, 0.0, 0, 0, 0, 0]),
        "nonmesh": _nonmesh(),
    }
    gcode = "G0 X2.5 Y5 Z0\nG1 X7.5 Y5 Z0\n"
    return inp, gcode


# 8-knot tables, 300-3000 K (nickel-alloy-like; W/m K and J/kg K)
K_TABLE = [[300.0, 500.0, 800.0, 1100.0, 1400.0, 1700.0, 2300.0, 3000.0],
           [9.8, 12.5, 16.6, 20.5, 24.5, 28.0, 30.0, 31.5]]
CP_TABLE = [[300.0, 500.0, 800.0, 1100.0, 1400.0, 1700.0, 2300.0, 3000.0],
            [410.0, 460.0, 520.0, 580.0, 640.0, 700.0, 720.0, 730.0]]


def raster_gcode(x0, x1, y0, hatch, n_lines, z=0.0):
    """Serpentine raster: G1 scan lines joined by G0 hatch jumps."""
    lines = [f"G0 X{x0:.4f} Y{y0:.4f} Z{z:g}"]
    x_end = x1
    for i in range(n_lines):
        y = y0 + i * hatch
        if i > 0:
            lines.append(f"G0 X{x_start:.4f} Y{y:.4f} Z{z:g}")
        lines.append(f"G1 X{x_end:.4f} Y{y:.4f} Z{z:g}")
        x_start, x_end = x_end, (x0 if x_end == x1 else x1)
    return "\n".join(lines) + "\n"


def production():
    """20-line raster on a 200x200x50 part-scale mesh (see module doc)."""
    props = _props(
        thermal_conductivity_bulk_a0=20.0, thermal_conductivity_fluid_a0=30.0,
        thermal_conductivity_bulk_table=K_TABLE, heat_capacity_solid_table=CP_TABLE,
        heat_capacity_mushy=800.0, heat_capacity_fluid=730.0,
        laser_radius=0.04, laser_depth=0.04, laser_power=250.0, laser_absorptivity=0.35,
        emissivity=0.4, laser_center=[7.5, 9.0, 0.0, 0, 0, 0, 0])
    inp = {
        "Level1": _level((200, 200, 50), (0, 20), (0, 20), (-5, 0), 300.0),
        "Level2": _level((100, 100, 50), (5.0, 10.0), (6.5, 11.5), (-2.5, 0)),
        "Level3": _level((100, 100, 50), (7.0, 8.0), (8.5, 9.5), (-0.5, 0)),
        "properties": props,
        "nonmesh": _nonmesh(),
    }
    return inp, raster_gcode(7.5, 12.5, 9.0, 0.1, 20)


def production_mirror():
    """The production meshes with the part-scale mesh centred on y = 9 mm and
    one straight track along y = 9 mm: every level is mirror-symmetric about
    that line, so the temperature fields are too (a full-size check)."""
    inp, _ = production()
    inp["Level1"]["bounds"]["y"] = [-1.0, 19.0]
    return inp, "G0 X7.5 Y9 Z0\nG1 X12.5 Y9 Z0\n"


def tiny(n=(8, 8, 4), n2=2, n3=2, track=1.0, props=None, velocity=100.0, dt3=1e-3):
    """Small 3-level deck: L1 over 8x8x4 mm, L2 4x4x2 mm, L3 2x2x1 mm.

    dt3 = 1 ms by default; at the default 100 mm/s the track then takes 10
    rows per mm.  Finer meshes need a smaller dt3 for the explicit scheme."""
    inp = {
        "Level1": _level(n, (0, 8), (0, 8), (-4, 0), 300.0),
        "Level2": _level(n, (2, 6), (2, 6), (-2, 0)),
        "Level3": _level(n, (3, 5), (3, 5), (-1, 0)),
        "properties": _props(**dict(dict(laser_center=[4.0, 4.0, 0.0, 0, 0, 0, 0],
                                         laser_radius=0.5, laser_depth=0.5, laser_power=2000.0),
                                    **(props or {}))),
        "nonmesh": _nonmesh(subcycle_num_L2=n2, subcycle_num_L3=n3, timestep_L3=dt3,
                            laser_velocity=velocity),
    }
    gcode = f"G0 X4 Y4 Z0\nG1 X{4 + track:g} Y4 Z0\n"
    return inp, gcode


def write_deck(inp, gcode, directory):
    """Write ``input.json`` and the G-code under ``directory``; paths are made
    absolute so the deck runs from any working directory."""
    os.makedirs(directory, exist_ok=True)
    d = os.path.abspath(directory)
    inp = json.loads(json.dumps(inp))
    inp["nonmesh"]["gcode"] = os.path.join(d, "path.gcode")
    inp["nonmesh"]["toolpath"] = os.path.join(d, "toolpath.txt")
    inp["nonmesh"]["save_path"] = os.path.join(d, "results") + "/"
    with open(inp["nonmesh"]["gcode"], "w") as fh:
        fh.write(gcode)
    path = os.path.join(d, "input.json")
    with open(path, "w") as fh:
        json.dump(inp, fh, indent=2)
    return path


def scan_length_mm(rows):
    """Laser path length with P > 0: sum of |dxy| over rows whose power is on."""
    rows = np.asarray(rows, dtype=np.float64)
    if rows.shape[0] < 2:
        return 0.0
    d = np.hypot(np.diff(rows[:, 0]), np.diff(rows[:, 1]))
    return float(d[rows[1:, 6] > 0].sum())


# L3 element counts of the batched configuration (uniform draw per instance)
BATCH_SIZES = [(32, 32, 32), (48, 48, 24), (64, 64, 32), (96, 96, 48)]


def batch_instances(n=64, seed=0, h=0.0125, dt=1e-6, sizes=None):
    """Independent single-track instances of an L3 window for the batched
    step: mesh size drawn from BATCH_SIZES, P uniform 100-400 W, speed
    uniform 0.5-2 m/s, laser radius uniform 30-70 um, from a fixed seed.
    Spacing h (mm) in all axes, top face at z = 0, origin at 0 in x and y;
    the track runs along x on the centre line from a quarter of the length.
    Returns a list of dicts: elements, properties (reference schema), speed
    (mm/s), P, x0 (track start), y, dt."""
    rng = np.random.default_rng(seed)
    sizes = sizes or BATCH_SIZES
    out = []
    for _ in range(n):
        el = sizes[int(rng.integers(len(sizes)))]
        P = float(rng.uniform(100.0, 400.0))
        v = float(rng.uniform(500.0, 2000.0))
        r = float(rng.uniform(0.030, 0.070))
        props = _props(laser_power=P, laser_radius=r, laser_depth=0.05)
        out.append({"elements": list(el), "properties": props, "speed": v, "P": P,
                    "h": h, "x0": 0.25 * el[0] * h, "y": 0.5 * el[1] * h, "dt": dt})
    return out


def batch_rows(instances, step):
    """Toolpath rows (x, y, z, Ljump, Ldwell, dt, P) of step `step` (0-based)
    of every instance: the laser at x0 + v (step + 1) dt, power on."""
    rows = np.zeros((len(instances), 7), np.float32)
    for i, b in enumerate(instances):
        rows[i] = (b["x0"] + b["speed"] * (step + 1) * b["dt"], b["y"], 0.0, 1, 1, b["dt"], b["P"])
    return rows


# ---------------------------------------------------------------------------
# Decks of the learned operator: training windows and the part-scale run
# ---------------------------------------------------------------------------

# L3 window of the operator: 64 x 64 x 32 nodes (63 x 63 x 31 elements) at
# 12.5 um; L2 63 x 63 x 31 elements at 50 um; L1 at 100 um
OP_L3_ELEMENTS = (63, 63, 31)
OP_H3, OP_H2, OP_H1 = 0.0125, 0.05, 0.1


def _op_windows(x, y):
    """L2 and L3 windows of the operator decks with the laser at (x, y) on the
    top surface: L3 starts 32 L3 spacings before the laser in x and y, L2 at
    an L1 node 16 L1 spacings before it."""
    e3 = OP_L3_ELEMENTS
    x3, y3 = round(x - 32 * OP_H3, 6), round(y - 32 * OP_H3, 6)
    # L3 origin on an L2 node, L2 origin on an L1 node
    x2 = round(np.floor((x3 - 24 * OP_H2) / OP_H1 + 1e-9) * OP_H1, 6)
    y2 = round(np.floor((y3 - 24 * OP_H2) / OP_H1 + 1e-9) * OP_H1, 6)
    x3 = round(x2 + round((x3 - x2) / OP_H2) * OP_H2, 6)
    y3 = round(y2 + round((y3 - y2) / OP_H2) * OP_H2, 6)
    L2 = _level(e3, (x2, x2 + e3[0] * OP_H2), (y2, y2 + e3[1] * OP_H2), (-e3[2] * OP_H2, 0))
    L3 = _level(e3, (x3, x3 + e3[0] * OP_H3), (y3, y3 + e3[1] * OP_H3), (-e3[2] * OP_H3, 0))
    return L2, L3


def window_deck(seed, n_l1=(80, 80, 40), length=4.0):
    """A single-track (even seed) or two-line raster (odd seed) run whose L3
    window is the operator's 64 x 64 x 32 nodes; P uniform 150-350 W and speed
    uniform 0.5-1.5 m/s from the seed.  The runs the training windows come
    from (the plain solver at every L2 substep)."""
    rng = np.random.default_rng(1000 + seed)
    P = float(rng.uniform(150.0, 350.0))
    v = float(rng.uniform(500.0, 1500.0))
    ext = [n_l1[0] * OP_H1, n_l1[1] * OP_H1, n_l1[2] * OP_H1]
    x0, y0 = 0.3 * ext[0], 0.5 * ext[1] - 0.1
    L2, L3 = _op_windows(x0, y0)
    props = _props(laser_power=P, laser_radius=0.05, laser_depth=0.05, emissivity=0.4,
                   laser_center=[x0, y0, 0.0, 0, 0, 0, 0])
    inp = {
        "Level1": _level(n_l1, (0, ext[0]), (0, ext[1]), (-ext[2], 0), 300.0),
        "Level2": L2,
        "Level3": L3,
        "properties": props,
        "nonmesh": _nonmesh(laser_velocity=v),
    }
    if seed % 2 == 0:
        g = f"G0 X{x0:.4f} Y{y0:.4f} Z0\nG1 X{x0 + length:.4f} Y{y0:.4f} Z0\n"
    else:
        g = raster_gcode(x0, x0 + 0.5 * length, y0, 0.1, 2)
    return inp, g


def part_scale(lines=30, layers=5, line_mm=10.0, hatch=0.1, P=250.0):
    """The surrogate-in-the-loop part: L1 400 x 400 x 100 elements over
    40 x 40 x 10 mm, the operator's L2/L3 windows, `layers` layers of 50 um
    (each activating its L1 elements as the laser reaches it), each a
    serpentine raster of `lines` lines of `line_mm` mm at `hatch` spacing,
    1 m/s, P = 250 W."""
    x0, y0 = 15.0, 15.0
    L2, L3 = _op_windows(x0, y0)
    props = _props(laser_power=P, laser_radius=0.05, laser_depth=0.05, emissivity=0.4,
                   laser_center=[x0, y0, 0.0, 0, 0, 0, 0], layer_height=0.05)
    inp = {
        "Level1": _level((400, 400, 100), (0, 40), (0, 40), (-10, 0), 300.0),
        "Level2": L2,
        "Level3": L3,
        "properties": props,
        "nonmesh": _nonmesh(),
    }
    g = "".join(raster_gcode(x0, x0 + line_mm, y0, hatch, lines, z=0.05 * k)
                for k in range(layers))
    return inp, g


def stack_decks(n=64, seed=0, h=0.0125, dt=1e-6, sizes=None, track_frac=0.5):
    """The batched configuration as three-level stacks: n single-track runs,
    each with its own L1/L2/L3 meshes of the L3 element counts drawn from
    BATCH_SIZES (the same counts on every level), L2 and L1 spanning 4x and
    16x the L3 extent (spacings h, 4h, 16h), P uniform 100-400 W, speed
    uniform 0.5-2 m/s and laser radius uniform 30-70 um, from a fixed seed
    (the same draws as batch_instances).  The track runs along x through the
    middle of L1 for `track_frac` of its length.  Returns a list of
    (solver_input, gcode)."""
    out = []
    for b in batch_instances(n=n, seed=seed, h=h, dt=dt, sizes=sizes):
        ex, ey, ez = b["elements"]
        h1, h2 = 16 * h, 4 * h
        ext1 = [ex * h1, ey * h1, ez * h1]
        # laser start a quarter of the way along x, centred in y, on L1 nodes
        x0 = round(0.25 * ex) * h1
        y0 = round(0.5 * ey) * h1
        # L2 and L3 centred on the laser, origins on the next coarser lattice
        x2 = x0 - round(0.5 * ex) * h2
        y2 = y0 - round(0.5 * ey) * h2
        x3 = x0 - round(0.5 * ex) * h
        y3 = y0 - round(0.5 * ey) * h
        x0, y0, x2, y2, x3, y3 = (round(v, 9) for v in (x0, y0, x2, y2, x3, y3))
        ext1 = [round(v, 9) for v in ext1]
        props = dict(b["properties"])
        props["laser_center"] = [x0, y0, 0.0, 0, 0, 0, 0]
        r9 = lambda v: round(v, 9)  # noqa: E731
        inp = {
            "Level1": _level((ex, ey, ez), (0, ext1[0]), (0, ext1[1]), (-ext1[2], 0), 300.0),
            "Level2": _level((ex, ey, ez), (x2, r9(x2 + ex * h2)), (y2, r9(y2 + ey * h2)),
                             (r9(-ez * h2), 0)),
            "Level3": _level((ex, ey, ez), (x3, r9(x3 + ex * h)), (y3, r9(y3 + ey * h)),
                             (r9(-ez * h), 0)),
            "properties": props,
            "nonmesh": _nonmesh(laser_velocity=b["speed"], timestep_L3=dt),
        }
        g = f"G0 X{x0:.6f} Y{y0:.6f} Z0\nG1 X{x0 + track_frac * ext1[0]:.6f} Y{y0:.6f} Z0\n"
        out.append((inp, g))
    return out
