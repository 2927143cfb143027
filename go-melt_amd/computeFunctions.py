"""GO-MELT call surface on MI355X: the functions go_melt.py imports.

This module keeps the names, positional arguments and return tuples of the
reference's go_melt/computeFunctions.py (the boundary go_melt.py imports with
``from computeFunctions import *``, go_melt.py:10), so the driver and user
scripts written against it run unchanged.  Underneath:

* set-up (``SetupProperties``, ``SetupLevels``, ...) is host Python that
  reproduces the reference's JSON schema, defaults and index bookkeeping, and
  creates a device-resident three-level state (``gm_engine``) in
  libgomelt.so;
* the time loop (``moveEverything``, ``subcycleGOMELT``, ``stepGOMELT``,
  ``stepGOMELTDwellTime``) runs as hand-written gfx950 kernels through the C
  ABI of include/gomelt.h; a subcycled L1 step is one hipGraph launch;
* the single operators (``solveMatrixFreeFE``, ``interpolatePoints``, ...) are
  one C-ABI call each; they take and return host NumPy arrays.

Fields live on the GPU.  ``Levels[k]["T0"]`` etc. are :class:`DeviceArray`
views of the engine's buffers (``np.asarray(...)`` downloads them).  There is
no CPU fallback: without libgomelt.so or a GPU every kernel call raises.
"""

from __future__ import annotations

import copy
import ctypes as C
import json
import math
import os
import pickle
import sys
import struct

import numpy as np

import _native as N
from _native import DeviceArray, GoMeltNativeError  # noqa: F401

F32 = np.float32

# ---------------------------------------------------------------------------
# Configuration (cF:67-112, 267-444)
# ---------------------------------------------------------------------------


class obj:
    """Attribute view of a dict (cF:67-82)."""

    def __init__(self, d):
        self.__dict__.update(d)


def dict2obj(d):
    """cF:85-98."""
    return json.loads(json.dumps(d), object_hook=obj)


def structure_to_dict(struct_):
    """cF:422-444."""
    return {k: (v if (not hasattr(v, "__dict__") or hasattr(v, "tolist"))
                else structure_to_dict(v)) for k, v in struct_.__dict__.items()}


def save_object(o, filename, extra=None):
    """cF:101-112 (the reference uses dill; fields are downloaded first).
    ``extra`` (a dict) is merged into the first level dict of a Levels list
    at position 0 of ``o``: the driver's checkpoint stores the engine state
    there."""
    h = _host_copy(o)
    if extra:
        h[0][0].update(extra)
    with open(filename, "wb") as fh:
        pickle.dump(h, fh, pickle.HIGHEST_PROTOCOL)


def _host_copy(o):
    if isinstance(o, DeviceArray):
        return o.numpy()
    if isinstance(o, dict):
        # the engine handle and the engine's own Level0 melt-time maps (the
        # driver checkpoints those as separate items) are not level data
        return {k: _host_copy(v) for k, v in o.items()
                if k not in ("_eng", "accum_time", "max_accum_time")}
    if isinstance(o, (list, tuple)):
        return type(o)(_host_copy(v) for v in o)
    return o


def SetupProperties(prop_obj):
    """Material, laser and surface constants with the reference's defaults and
    unit fixes (cF:267-345): h_conv x1e6, molar mass x1e-3 and the derived
    evaporation coefficients.  Optional extension: piecewise-linear tables
    ``thermal_conductivity_bulk_table`` / ``heat_capacity_solid_table``
    ([[T knots], [values]]) replace the linear bulk-k / solid-cp laws."""
    p = dict(prop_obj)
    g = prop_obj.get
    p["k_powder"] = g("thermal_conductivity_powder", 0.4)
    p["k_bulk_coeff_a0"] = g("thermal_conductivity_bulk_a0", 4.23)
    p["k_bulk_coeff_a1"] = g("thermal_conductivity_bulk_a1", 0.016)
    p["k_fluid_coeff_a0"] = g("thermal_conductivity_fluid_a0", 29.0)
    p["cp_solid_coeff_a0"] = g("heat_capacity_solid_a0", 383.1)
    p["cp_solid_coeff_a1"] = g("heat_capacity_solid_a1", 0.174)
    p["cp_mushy"] = g("heat_capacity_mushy", 3235.0)
    p["cp_fluid"] = g("heat_capacity_fluid", 769.0)
    p["rho"] = g("density", 8.0e-6)
    p["laser_radius"] = g("laser_radius", 0.110)
    p["laser_depth"] = g("laser_depth", 0.05)
    p["laser_power"] = g("laser_power", 300.0)
    p["laser_eta"] = g("laser_absorptivity", 0.25)
    p["laser_center"] = g("laser_center", [])
    p["T_amb"] = g("T_amb", 353.15)
    p["T_solidus"] = g("T_solidus", 1554.0)
    p["T_liquidus"] = g("T_liquidus", 1625.0)
    p["T_boiling"] = g("T_boiling", 3038.0)
    p["h_conv"] = g("h_conv", 1.473e-5) * 1e6
    p["vareps"] = g("emissivity", 0.600)
    p["evc"] = g("evaporation_coefficient", 0.82)
    p["kb"] = g("boltzmann_constant", 1.38e-23)
    p["mA"] = g("atomic_mass", 7.9485017e-26)
    p["Lev"] = g("latent_heat_evap", 4.22e6)
    p["molar_mass"] = g("molar_mass", 58.69) * 1e-3
    p["layer_height"] = g("layer_height", 0.04)
    p["sigma_sb"] = 5.67e-8
    p["gas_const"] = 8.314
    p["atmospheric_pressure"] = 101325
    p["CM_coeff"] = p["molar_mass"] / (2.0 * math.pi * p["gas_const"])
    p["CT_coeff"] = p["Lev"] * p["molar_mass"] / p["gas_const"]
    p["CP_coeff"] = 0.54 * p["atmospheric_pressure"]
    p["k_table"] = g("thermal_conductivity_bulk_table")
    p["cp_table"] = g("heat_capacity_solid_table")
    return p


def SetupNonmesh(nonmesh_input):
    """cF:348-419 (creates save_path like the reference)."""
    n = dict(nonmesh_input)
    g = nonmesh_input.get
    n["timestep_L3"] = g("timestep_L3", 1e-5)
    n["subcycle_num_L2"] = g("subcycle_num_L2", 1)
    n["subcycle_num_L3"] = g("subcycle_num_L3", 1)
    n["dwell_time"] = g("dwell_time", 0.1)
    n["Level1_record_step"] = g("Level1_record_step", 1)
    n["save_path"] = g("save_path", "results/")
    n["output_files"] = g("output_files", 1)
    n["toolpath"] = g("toolpath", "laserPath.txt")
    n["wait_time"] = g("wait_time", 500.0)
    n["layer_num"] = g("layer_num", 0)
    n["restart_layer_num"] = g("restart_layer_num", 10000)
    n["info_T"] = g("info_T", 0)
    n["laser_velocity"] = g("laser_velocity", 500)
    n["wait_track"] = g("wait_track", 0.0)
    n["record_step"] = g("record_step", n["subcycle_num_L2"] * n["subcycle_num_L3"])
    n["gcode"] = g("gcode", "./examples/gcodefiles/defaultName.gcode")
    n["dwell_time_multiplier"] = g("dwell_time_multiplier", 1)
    n["use_txt"] = g("use_txt", 0)
    if not os.path.exists(n["save_path"]):
        os.makedirs(n["save_path"])
    return n


def _material(p):
    """gm_material of a properties dict (FP32, kernel units)."""
    m = N.gm_material()
    m.T_amb, m.T_solidus, m.T_liquidus, m.T_boiling = (p["T_amb"], p["T_solidus"],
                                                        p["T_liquidus"], p["T_boiling"])
    m.k_powder, m.k_bulk_a0, m.k_bulk_a1, m.k_fluid = (p["k_powder"], p["k_bulk_coeff_a0"],
                                                       p["k_bulk_coeff_a1"],
                                                       p["k_fluid_coeff_a0"])
    m.cp_solid_a0, m.cp_solid_a1, m.cp_mushy, m.cp_fluid = (p["cp_solid_coeff_a0"],
                                                            p["cp_solid_coeff_a1"],
                                                            p["cp_mushy"], p["cp_fluid"])
    m.rho = p["rho"]
    for key, nfield, tf, vf in (("k_table", "k_table_n", "k_table_T", "k_table_v"),
                                ("cp_table", "cp_table_n", "cp_table_T", "cp_table_v")):
        tab = p.get(key)
        if tab is not None:
            kn, kv = list(tab[0]), list(tab[1])
            if len(kn) != len(kv) or not 2 <= len(kn) <= N.GM_MAX_KNOTS:
                raise ValueError(f"{key}: need 2..{N.GM_MAX_KNOTS} knots")
            setattr(m, nfield, len(kn))
            for i, (t, v) in enumerate(zip(kn, kv)):
                getattr(m, tf)[i] = t
                getattr(m, vf)[i] = v
    m.h_conv = p["h_conv"]
    m.sigma_eps = p["sigma_sb"] * p["vareps"]
    m.evc_cp = p["evc"] * p["CP_coeff"]
    m.CT = p["CT_coeff"]
    m.CM = p["CM_coeff"]
    m.Lev = p["Lev"]
    m.laser_eta = p["laser_eta"]
    m.laser_radius = p["laser_radius"]
    m.laser_depth = p["laser_depth"]
    return m


# ---------------------------------------------------------------------------
# Mesh set-up (cF:18-64, 115-264, 447-579, 858-925, 1642-1669, 1886-1942)
# ---------------------------------------------------------------------------


def calcNumNodes(elements):
    """cF:18-29."""
    return [elements[0] + 1, elements[1] + 1, elements[2] + 1]


def createMesh3D(x, y, z):
    """cF:32-64: FP32 node axes and the per-axis connectivity (8 local nodes,
    order 0(000) 1(100) 2(110) 3(010) 4(001) 5(101) 6(111) 7(011))."""
    axes = [np.linspace(a[0], a[1], int(a[2])).astype(F32) for a in (x, y, z)]
    bits = (np.array([0, 1, 1, 0, 0, 1, 1, 0]), np.array([0, 0, 1, 1, 0, 0, 1, 1]),
            np.array([0, 0, 0, 0, 1, 1, 1, 1]))
    conn = [np.arange(int(a[2]) - 1)[:, None] + b[None, :] for a, b in zip((x, y, z), bits)]
    return axes, conn


def calc_length_h(A):
    """cF:1916-1942 (A: dict with bounds and elements)."""
    b = A["bounds"]
    L = [b["x"][1] - b["x"][0], b["y"][1] - b["y"][0], b["z"][1] - b["z"][0]]
    return L, [L[i] / A["elements"][i] for i in range(3)]


def getBCindices(x):
    """cF:520-559: node index sets [west, east, south, north, bottom, top]."""
    nx, ny, nz = x["nodes"]
    nn = x["nn"]
    s = (np.arange(nx)[:, None] + (nx * ny * np.arange(nz))[None, :]).reshape(-1)
    n_ = (np.arange(nx * (ny - 1), nx * ny)[:, None] + (nx * ny * np.arange(nz))[None, :]).reshape(-1)
    return [np.arange(0, nn, nx), np.arange(nx - 1, nn, nx), s, n_, np.arange(nx * ny),
            np.arange(nx * ny * (nz - 1), nn)]


def find_max_const(CoarseLevel, FinerLevel):
    """cF:1886-1913: how far the finer window may move inside the coarse bounds."""
    c, f = CoarseLevel["bounds"], FinerLevel["bounds"]
    return ([c["x"][0] - f["x"][0], c["x"][1] - f["x"][1]],
            [c["y"][0] - f["y"][0], c["y"][1] - f["y"][1]],
            [c["z"][0] - f["z"][0], c["z"][1] - f["z"][1]])


def getCoarseNodesInFineRegion(xnf, xnc):
    """cF:858-887."""
    hc = (xnc.max() - xnc.min()) / (xnc.size - 1)
    lo = np.round((xnf.min() - xnc.min()) / hc)
    hi = np.round((xnf.max() - xnc.min()) / hc) + 1
    return np.arange(lo, hi).astype(np.int64)


def getCoarseNodesInLargeFineRegion(xnc, xnf):
    """cF:890-925."""
    hf = (xnf.max() - xnf.min()) / (xnf.size - 1)
    hc = (xnc.max() - xnc.min()) / (xnc.size - 1)
    lo = np.round((xnc.min() - xnf.min()) / hf)
    hi = np.round((xnc.max() - xnf.min()) / hf) + 1
    return np.arange(lo, hi, int(np.round(hc / hf))).astype(np.int64)


def getOverlapRegion(node_coords, nx, ny):
    """cF:1642-1669: flat x-fastest indices of a tensor index box."""
    ax, ay, az = [np.asarray(a, dtype=np.int64) for a in node_coords]
    return (ax[None, None, :] + ay[None, :, None] * nx + az[:, None, None] * nx * ny).reshape(-1)


def _grid_of(lv):
    """gm_grid of a level dict's current node axes."""
    nc = lv["node_coords"]
    h = [F32(lv["h"][d]) if "h" in lv else (nc[d][1] - nc[d][0]) for d in range(3)]
    return N.gm_grid.make(lv["nodes"], [nc[d][0] for d in range(3)], h)


def _level_dict(src, p):
    lv = {"elements": list(src["elements"]),
          "bounds": {k: list(v) for k, v in src["bounds"].items()}}
    if "conditions" in src:
        lv["conditions"] = {k: list(v) for k, v in src["conditions"].items()}
    lv["length"], lv["h"] = calc_length_h(lv)
    lv["nodes"] = calcNumNodes(lv["elements"])
    lv["ne"] = lv["elements"][0] * lv["elements"][1] * lv["elements"][2]
    lv["nn"] = lv["nodes"][0] * lv["nodes"][1] * lv["nodes"][2]
    lv["BC"] = getBCindices(lv)
    b = lv["bounds"]
    lv["node_coords"], lv["connect"] = createMesh3D(
        (b["x"][0], b["x"][1], lv["nodes"][0]), (b["y"][0], b["y"][1], lv["nodes"][1]),
        (b["z"][0], b["z"][1], lv["nodes"][2]))
    return lv


class LevelDict(dict):
    """A level of ``Levels`` bound to the device state of an engine.

    Keys that name engine fields (``T0``, ``Tprime0``, ``S1``, ``S2`` and the
    Level0 melt-time maps) hold device views of the engine's buffers.
    Assigning to such a key, as the reference's driver code does
    (e.g. ``Levels[2]["Tprime0"] = ...`` at go_melt.py:366-367), copies the
    new values INTO the engine buffer; the key keeps the device view, so the
    assignment cannot detach the level from the state the kernels advance.
    The overlap index sets that moveEverything updates (cF:2472-2497) are
    read from the engine on access: ``overlapNodes`` (all windows) and, on
    Level0, ``overlapNodes_L2``, ``idx`` and ``idx_L2`` (flat x-fastest
    indices, getOverlapRegion cF:1642-1669)."""

    _BOXES = {(0, "overlapNodes"): 0, (0, "overlapNodes_L2"): 1, (2, "overlapNodes"): 2,
              (3, "overlapNodes"): 3}

    def __getitem__(self, key):
        eng = dict.get(self, "_eng")
        lev = dict.get(self, "level")
        if eng is not None:
            if lev == 0 and key in ("idx", "idx_L2"):
                nodes = eng.overlap_nodes(0 if key == "idx" else 1)
                return getOverlapRegion(nodes, self["nodes"][0], self["nodes"][1])
            which = self._BOXES.get((lev, key))
            if which is not None:
                return eng.overlap_nodes(which)
        return dict.__getitem__(self, key)

    def get(self, key, default=None):
        try:
            return self[key]
        except KeyError:
            return default

    def __setitem__(self, key, value):
        cur = dict.get(self, key)
        if isinstance(cur, DeviceArray) and dict.get(self, "_eng") is not None:
            if value is not cur:
                v = value.numpy() if isinstance(value, DeviceArray) else np.asarray(value)
                cur.upload(v.astype(cur.dtype).reshape(-1))
            return
        dict.__setitem__(self, key, value)


def SetupLevels(solver_input, properties):
    """cF:115-264: Level0 (state grid) and Levels 1-3, their overlap index
    boxes, and the device state (T = T_amb, S1 = 0) in a new gm_engine."""
    Levels = [LevelDict(lv) for lv in setup_levels_host(solver_input, properties)]
    eng = Engine(Levels, properties, solver_input.get("nonmesh", {}))
    for lv in Levels:
        dict.__setitem__(lv, "_eng", eng)
    eng.bind(Levels)
    return Levels


def setup_levels_host(solver_input, properties):
    """The host part of SetupLevels: mesh metadata and overlap boxes only."""
    p = properties
    L1, L2, L3 = (_level_dict(solver_input["Level1"], p), _level_dict(solver_input["Level2"], p),
                  _level_dict(solver_input["Level3"], p))
    L1["S1_storage_rows"] = int(round(L1["h"][2] / p["layer_height"]))
    for lv in (L2, L3):
        ix, iy, iz = find_max_const(L1, lv)
        lv["bounds"]["ix"], lv["bounds"]["iy"], lv["bounds"]["iz"] = ix, iy, iz
        lv["init_node_coors"] = copy.deepcopy(lv["node_coords"])
    L1["orig_node_coords"] = copy.deepcopy(L1["node_coords"])
    tmp = copy.deepcopy(L1["orig_node_coords"])
    while not np.isclose(tmp[2] - L2["node_coords"][2][-1], 0, atol=1e-4).any():
        tmp[2] = (tmp[2] + F32(p["layer_height"])).astype(F32)
    L1["node_coords"] = tmp

    def overlap(fine, coarse):
        nodes = [getCoarseNodesInFineRegion(fine["node_coords"][i], coarse["node_coords"][i])
                 for i in range(3)]
        return nodes, [coarse["node_coords"][i][nodes[i]] for i in range(3)]

    L2["orig_overlap_nodes"], L2["orig_overlap_coors"] = overlap(L2, L1)
    L3["orig_overlap_nodes"], L3["orig_overlap_coors"] = overlap(L3, L2)
    for lv in (L2, L3):
        lv["overlapNodes"] = copy.deepcopy(lv["orig_overlap_nodes"])
        lv["overlapCoords"] = copy.deepcopy(lv["orig_overlap_coors"])

    L0 = {"elements": [round(L1["length"][0] / L3["h"][0]), round(L1["length"][1] / L3["h"][1]),
                       round(L2["length"][2] / L3["h"][2])]}
    L0["nodes"] = calcNumNodes(L0["elements"])
    L0["ne"] = L0["elements"][0] * L0["elements"][1] * L0["elements"][2]
    L0["nn"] = L0["nodes"][0] * L0["nodes"][1] * L0["nodes"][2]
    L0["node_coords"], L0["connect"] = createMesh3D(
        (L1["bounds"]["x"][0], L1["bounds"]["x"][1], L0["nodes"][0]),
        (L1["bounds"]["y"][0], L1["bounds"]["y"][1], L0["nodes"][1]),
        (L2["bounds"]["z"][0], L2["bounds"]["z"][1], L0["nodes"][2]))
    L0["orig_node_coords"] = copy.deepcopy(L0["node_coords"])
    L0["orig_overlap_nodes"], L0["orig_overlap_coors"] = overlap(L3, L0)
    L0["orig_overlap_nodes_L2"] = [
        getCoarseNodesInLargeFineRegion(L2["node_coords"][i], L0["node_coords"][i])
        for i in range(3)]
    L0["orig_overlap_coors_L2"] = [L0["node_coords"][i][L0["orig_overlap_nodes_L2"][i]]
                                   for i in range(3)]
    L0["overlapNodes"] = copy.deepcopy(L0["orig_overlap_nodes"])
    L0["overlapNodes_L2"] = copy.deepcopy(L0["orig_overlap_nodes_L2"])
    L0["layer_idx_delta"] = int(round(p["layer_height"] / L3["h"][2]))
    for i, lv in enumerate((L0, L1, L2, L3)):
        lv["level"] = i
    return [L0, L1, L2, L3]


def getStaticNodesAndElements(L):
    """cF:447-470: (ne2, ne3, nn1, nn2, nn3)."""
    return (L[2]["ne"], L[3]["ne"], L[1]["nn"], L[2]["nn"], L[3]["nn"])


def getStaticSubcycle(Nm):
    """cF:473-492."""
    N2, N3 = Nm["subcycle_num_L2"], Nm["subcycle_num_L3"]
    return (N2, N3, N2 * N3, float(N2), float(N3), float(N2 * N3))


def calcStaticTmpNodesAndElements(L, v):
    """cF:495-517: active L1 elements and nodes at or below the laser z."""
    n = int(np.sum(L[1]["node_coords"][2] <= v[2] + 1e-5))
    return (L[1]["elements"][0] * L[1]["elements"][1] * (n - 1),
            L[1]["nodes"][0] * L[1]["nodes"][1] * n)


def getSubstrateNodes(Levels):
    """cF:562-579."""
    return tuple(int(np.sum(L["node_coords"][2] < 1e-5) * L["nodes"][0] * L["nodes"][1])
                 for L in Levels[:4])


# ---------------------------------------------------------------------------
# The device-resident state
# ---------------------------------------------------------------------------


_ENGINE_STREAM = []  # a stream new engines share (engines_on_stream)


class engines_on_stream:
    """Context in which SetupLevels creates its engines on the given stream
    (an N.Stream), for runs advanced together (batch.Stacks)."""

    def __init__(self, stream):
        self.stream = stream

    def __enter__(self):
        _ENGINE_STREAM.append(self.stream)
        return self.stream

    def __exit__(self, *exc):
        _ENGINE_STREAM.pop()
        return False


class Engine:
    """One gm_engine: the three levels, Level0 and the scratch of the L1 step."""

    def __init__(self, Levels, p, nonmesh):
        N.init_device()
        L0, L1, L2, L3 = Levels
        self.props = p
        self.stream = _ENGINE_STREAM[-1] if _ENGINE_STREAM else N.Stream()
        self.mat = _material(p)
        cfg = N.gm_config()
        for l, lv in enumerate(Levels):
            if l == 0:
                nc = lv["node_coords"]
                h = [(nc[d][-1] - nc[d][0]) / (lv["nodes"][d] - 1) for d in range(3)]
                cfg.level[0] = N.gm_grid.make(lv["nodes"], [nc[d][0] for d in range(3)], h)
            else:
                cfg.level[l] = _grid_of(lv)
        cond = L1["conditions"]
        for i, v in enumerate((cond["y"][0], cond["y"][1], cond["x"][0], cond["x"][1],
                               cond["z"][0])):
            cfg.l1_bc[i] = v
        cfg.layer_height = p["layer_height"]
        cfg.n2 = int(nonmesh.get("subcycle_num_L2", 1))
        cfg.n3 = int(nonmesh.get("subcycle_num_L3", 1))
        self.r12 = [int(np.round(L1["h"][i] / L2["h"][i])) for i in range(2)] + \
                   [int(np.round(p["layer_height"] / L2["h"][2]))]
        self.r23 = [int(np.round(L2["h"][i] / L3["h"][i])) for i in range(3)]
        for d in range(3):
            cfg.ratio12[d] = self.r12[d]
            cfg.ratio23[d] = self.r23[d]
        for w, lv in enumerate((L2, L3)):
            b = lv["bounds"]
            for d, k in enumerate(("ix", "iy", "iz")):
                cfg.move_lo[w][d] = b[k][0]
                cfg.move_hi[w][d] = b[k][1]
        boxes = {2: (L2["orig_overlap_nodes"], L2["orig_overlap_coors"]),
                 3: (L3["orig_overlap_nodes"], L3["orig_overlap_coors"]),
                 0: (L0["orig_overlap_nodes"], L0["orig_overlap_coors"]),
                 1: (L0["orig_overlap_nodes_L2"], L0["orig_overlap_coors_L2"])}
        for k, (nodes, coors) in boxes.items():
            for d in range(3):
                cfg.ov_lo[k][d] = int(nodes[d][0]) if len(nodes[d]) else 0
                cfg.ov_n[k][d] = len(nodes[d])
                cfg.ov_x0[k][d] = float(coors[d][0]) if len(coors[d]) else 0.0
        for d in range(3):
            st = L0["orig_overlap_nodes_L2"][d]
            cfg.l0_stride_l2[d] = int(st[1] - st[0]) if len(st) > 1 else 1
            if cfg.ov_n[0][d] != L3["nodes"][d] or cfg.ov_n[1][d] != L2["nodes"][d]:
                raise ValueError("Level0 must align with the L3 and L2 nodes "
                                 "(the reference's Level0 index sets would not match)")
        cfg.l0_layer_delta = L0["layer_idx_delta"]
        self.cfg = cfg
        self.N2, self.N3 = cfg.n2, cfg.n3
        h = C.c_void_p()
        N.check(N.lib().gm_engine_create(C.byref(cfg), C.byref(self.mat), self.stream.ptr,
                                         C.byref(h)))
        self.h = h.value
        self.layer = None
        if os.environ.get("GOMELT_NO_GRAPH"):
            N.check(N.lib().gm_engine_use_graph(self.h, 0))

    def __del__(self):
        try:
            if self.h:
                N.lib().gm_engine_destroy(self.h)
        except Exception:
            pass

    def field(self, level, field):
        p = C.c_void_p()
        n = C.c_int64()
        N.check(N.lib().gm_engine_field(self.h, level, field, C.byref(p), C.byref(n)))
        dt = np.float32
        if (field in (N.GM_S1, N.GM_S2) and not (field == N.GM_S1 and level == 1)) or \
                field == N.GM_RESET_L3:
            dt = np.uint8
        return DeviceArray(n.value, dt, ptr=p.value, owner=False)

    def grid(self, level):
        g = N.gm_grid()
        N.check(N.lib().gm_engine_grid(self.h, level, C.byref(g)))
        return g

    def bind(self, Levels):
        """Put device views of the engine's fields into the level dicts."""
        self.Levels = Levels
        put = dict.__setitem__
        put(Levels[1], "T0", self.field(1, N.GM_T0))
        put(Levels[1], "S1", self.field(1, N.GM_S1))
        for l in (2, 3):
            put(Levels[l], "T0", self.field(l, N.GM_T0))
            put(Levels[l], "Tprime0", self.field(l, N.GM_TPRIME0))
            put(Levels[l], "S1", self.field(l, N.GM_S1))
        put(Levels[3], "S2", self.field(3, N.GM_S2))
        put(Levels[0], "S1", self.field(0, N.GM_S1))
        put(Levels[0], "accum_time", self.field(0, N.GM_ACCUM))
        put(Levels[0], "max_accum_time", self.field(0, N.GM_MAX_ACCUM))

    def overlap_nodes(self, which):
        """Index axes of an overlap box after the last move (see
        gm_engine_overlap_box): 0 Level0 in L3, 1 Level0 in L2 (strided),
        2 L1 in L2, 3 L2 in L3."""
        lo, n, st = (np.zeros(3, np.int32) for _ in range(3))
        N.check(N.lib().gm_engine_overlap_box(self.h, int(which), lo.ctypes.data_as(N._IP),
                                              n.ctypes.data_as(N._IP),
                                              st.ctypes.data_as(N._IP)))
        return [int(lo[d]) + int(st[d]) * np.arange(int(n[d]), dtype=np.int64)
                for d in range(3)]

    def save_state(self):
        """The engine's device state as bytes (gm_engine_save_state)."""
        n = C.c_int64()
        N.check(N.lib().gm_engine_state_bytes(self.h, C.byref(n)))
        buf = np.empty(n.value, np.uint8)
        N.check(N.lib().gm_engine_save_state(self.h, buf.ctypes.data))
        return buf.tobytes()

    def load_state(self, blob):
        buf = np.frombuffer(blob, np.uint8)
        N.check(N.lib().gm_engine_load_state(self.h, buf.ctypes.data, buf.size))
        self.layer = None

    def set_accum_external(self, external):
        if bool(external) != getattr(self, "_acc_ext", False):
            N.check(N.lib().gm_engine_set_accum_mode(self.h, int(bool(external))))
            self._acc_ext = bool(external)

    def set_layer(self, tmp_ne_nn, substrate):
        key = (tuple(int(v) for v in tmp_ne_nn), tuple(int(v) for v in substrate))
        if key != self.layer:
            sub = (C.c_int64 * 4)(*key[1])
            N.check(N.lib().gm_engine_set_layer(self.h, key[0][0], key[0][1], sub))
            self.layer = key

    def sync_coords(self):
        for l in (1, 2, 3):
            self.Levels[l]["node_coords"] = self.grid(l).axes()

    def sync(self):
        self.stream.sync()


def attach_l3_operator(Levels, operator):
    """Let ``operator`` (surrogate.EngineOperator, or None to detach) advance
    the L3 window in subcycleGOMELT: the subcycle then runs stage by stage and
    the operator takes the N3 L3 solves of each L2 substep (an addition with no
    reference counterpart; the reference always runs them, cF:3367-3412,
    3530-3590)."""
    _eng(Levels).l3_operator = operator


def _eng(Levels):
    for lv in Levels:
        if isinstance(lv, dict) and "_eng" in lv:
            return lv["_eng"]
    raise ValueError("Levels were not created by SetupLevels")


# ---------------------------------------------------------------------------
# Single operators (one C-ABI call each; host arrays in and out)
# ---------------------------------------------------------------------------


def _dev(x, dtype=np.float32):
    return N.as_device(x, dtype)


def computeStateProperties(T, S1, properties, Level_nodes_substrate):
    """cF:2567-2614 -> (S1, S2, k, rhocp)."""
    T = _dev(T)
    n = T.n
    S1d = _dev(S1)
    S1o, S2o = DeviceArray(n), DeviceArray(n, np.uint8)
    k, c = DeviceArray(n), DeviceArray(n)
    N.check(N.lib().gm_state_properties(C.byref(_material(properties)), T.ptr, S1d.ptr, n,
                                        int(Level_nodes_substrate), S1o.ptr, S2o.ptr, k.ptr,
                                        c.ptr, None))
    return S1o.numpy(), S2o.numpy().astype(bool), k.numpy(), c.numpy()


def solveMatrixFreeFE(Level, nn, ne, k, rhocp, dt, T, Fc, Corr):
    """cF:582-642: one explicit lumped-mass update of a level (host arrays)."""
    g = _grid_of(Level)
    out = DeviceArray(nn)
    kd, cd, Td = _dev(k), _dev(rhocp), _dev(T)
    Fd = None if (np.isscalar(Fc) and Fc == 0) else _dev(Fc)
    Cd = None if (np.isscalar(Corr) and Corr == 0) else _dev(Corr)
    N.check(N.lib().gm_solve_matrix_free(C.byref(g), int(ne), kd.ptr, cd.ptr, float(dt), Td.ptr,
                                         Fd.ptr if Fd else None, Cd.ptr if Cd else None, out.ptr,
                                         None))
    return out.numpy()


def computeSourceFunction_jax(x, y, z, v, properties, P):
    """cF:991-1025 (pointwise Gaussian; host evaluation of a formula, used for
    diagnostics -- the kernels integrate it in closed separable form)."""
    p = properties
    pc = 6 * np.sqrt(3) * P * p["laser_eta"]
    rc = 1 / (p["laser_radius"] * np.sqrt(np.pi))
    dc = 1 / (p["laser_depth"] * np.sqrt(np.pi))
    return (pc * rc * np.exp(-3 * (x - v[0]) ** 2 / p["laser_radius"] ** 2)
            * rc * np.exp(-3 * (y - v[1]) ** 2 / p["laser_radius"] ** 2)
            * dc * np.exp(-3 * (z - v[2]) ** 2 / p["laser_depth"] ** 2))


def computeSourcesL3(Level, v, ne_nn, properties, laserP):
    """cF:2960-3012."""
    g = _grid_of(Level)
    out = DeviceArray(g.nn())
    vv = np.asarray(v, dtype=np.float32)[:3].copy()
    N.check(N.lib().gm_source_l3(C.byref(g), C.byref(_material(properties)), N.fptr(vv),
                                 float(laserP), out.ptr, None))
    return out.numpy()


class Shape:
    """Fine -> coarse restriction handle, in the slot of the reference's
    ``Shapes[k]`` (cF:1213-1358).  The coarse shape functions at the fine
    Gauss points are evaluated inside the kernels, never stored."""

    def __init__(self, coarse, fine):
        self.coarse, self.fine = coarse, fine


def computeCoarseFineShapeFunctions(Coarse, Fine):
    """cF:1213-1358 -> a Shape handle (nothing is tabulated in memory)."""
    return Shape(Coarse, Fine)


def computeLevelSource(Levels, ne_nn, laser_position, LevelShape, properties, laserP):
    """cF:2667-2730."""
    rows = np.array(laser_position, dtype=np.float32).reshape(-1, 7).copy()
    rows[:, 6] = np.asarray(laserP, dtype=np.float32).reshape(-1)
    gc, gf = _grid_of(LevelShape.coarse), _grid_of(LevelShape.fine)
    out = DeviceArray(gc.nn())
    N.check(N.lib().gm_level_source(C.byref(gc), C.byref(gf), C.byref(_material(properties)),
                                    N.fptr(rows), rows.shape[0], out.ptr, None))
    return out.numpy()


def computeSources(Level, v, Shapes, ne_nn, properties, laserP):
    """cF:928-988 -> (Fc on L1, Fm on L2, Ff on L3)."""
    rows = np.zeros((1, 7), dtype=np.float32)
    rows[0, :3] = np.asarray(v, dtype=np.float32)[:3]
    Fc = computeLevelSource(None, ne_nn, rows, Shapes[1], properties, [laserP])
    Fm = computeLevelSource(None, ne_nn, rows, Shapes[2], properties, [laserP])
    return Fc, Fm, computeSourcesL3(Level, v, ne_nn, properties, laserP)


def computeConvRadBC(Level, LevelT0, ne, nn, properties, F):
    """cF:2207-2301: F + top-face convection/radiation/evaporation flux."""
    g = _grid_of(Level)
    Fd = DeviceArray.from_numpy(np.zeros(nn, np.float32) + np.asarray(F, np.float32))
    Td = _dev(LevelT0)
    N.check(N.lib().gm_conv_rad_bc(C.byref(g), C.byref(_material(properties)), Td.ptr,
                                   int(ne), Fd.ptr, None))
    return Fd.numpy()


def _target_grid(coords):
    c = [np.asarray(a, dtype=np.float32) for a in coords]
    h = [(a[1] - a[0]) if a.size > 1 else F32(1.0) for a in c]
    return N.gm_grid.make([a.size for a in c], [a[0] for a in c], h)


def interpolatePointsMatrix(Level, node_coords_new):
    """cF:1028-1107 -> [Nc (n, 8), node (n, 8)]."""
    gs, gt = _grid_of(Level), _target_grid(node_coords_new)
    n = gt.nn()
    Nc, node = DeviceArray(8 * n), DeviceArray(8 * n, np.int32)
    N.check(N.lib().gm_interpolate_matrix(C.byref(gs), C.byref(gt), Nc.ptr, node.ptr, None))
    return [Nc.numpy().reshape(n, 8), node.numpy().reshape(n, 8)]


def interpolate_w_matrix(C2F, T):
    """cF:1110-1128."""
    Nc, node = _dev(np.asarray(C2F[0]).reshape(-1)), _dev(np.asarray(C2F[1]).reshape(-1), np.int32)
    n = Nc.n // 8
    out = DeviceArray(n)
    Td = _dev(T)
    N.check(N.lib().gm_interpolate_w_matrix(Nc.ptr, node.ptr, n, Td.ptr, out.ptr, None))
    return out.numpy()


def interpolatePoints(Level, u, node_coords_new):
    """cF:1131-1210."""
    gs, gt = _grid_of(Level), _target_grid(node_coords_new)
    out = DeviceArray(gt.nn())
    ud = _dev(u)
    N.check(N.lib().gm_interpolate_points(C.byref(gs), ud.ptr, C.byref(gt), out.ptr, None))
    return out.numpy()


def _restrict(Shape_, properties, T, S1, n_sub, Tp=None, TpA=None, TpB=None, dt=1.0):
    gc, gf = _grid_of(Shape_.coarse), _grid_of(Shape_.fine)
    V = DeviceArray.from_numpy(np.zeros(gc.nn(), np.float32))
    # keep every temporary device buffer referenced until the call returns
    held = [None if a is None else _dev(a) for a in (T, S1, Tp, TpA, TpB)]
    ptrs = [None if h is None else h.ptr for h in held]
    N.check(N.lib().gm_restrict(C.byref(gc), C.byref(gf), C.byref(_material(properties)),
                                ptrs[0], ptrs[1], int(n_sub), ptrs[2], ptrs[3], ptrs[4],
                                float(dt), V.ptr, None))
    out = V.numpy()
    del held
    return out


def restrictFlux(Shape_, properties, T, S1, n_sub, Tp):
    """-sum_e sum_q grad Nc . (k_e grad T') w (one Shapes pair of
    cF:1513-1533, 2776-2783, 2897-2914); k from (T, S1)."""
    return _restrict(Shape_, properties, T, S1, n_sub, Tp=Tp)


def restrictMass(Shape_, properties, T, S1, n_sub, TpA, TpB, dt):
    """-(1/dt) sum_e sum_q Nc rho*cp_e (TpA - TpB) w (cF:1440-1468, 3106-3109,
    3211-3216); rho*cp from (T, S1)."""
    return _restrict(Shape_, properties, T, S1, n_sub, TpA=TpA, TpB=TpB, dt=dt)


def getNewTprime(Fine, FineT0, CoarseT, Coarse, C2F):
    """cF:2060-2099 -> (Tprime, CoarseT with the fine field injected)."""
    gf, gc = _grid_of(Fine), _grid_of(Coarse)
    Cd = DeviceArray.from_numpy(np.asarray(CoarseT, np.float32))
    Fd = _dev(FineT0)
    Tp = DeviceArray(gf.nn())
    ov = Fine["overlapNodes"]
    lo = np.array([int(np.asarray(ov[d])[0]) for d in range(3)], np.int32)
    nb = np.array([len(ov[d]) for d in range(3)], np.int32)
    N.check(N.lib().gm_new_tprime(C.byref(gf), Fd.ptr, C.byref(gc), Cd.ptr,
                                  lo.ctypes.data_as(N._IP), nb.ctypes.data_as(N._IP), Tp.ptr,
                                  None))
    return Tp.numpy(), Cd.numpy()


def assignBCs(RHS, Levels):
    """cF:1568-1595 (host helper; the kernels apply it fused)."""
    out = np.array(RHS, dtype=np.float32, copy=True)
    bc, cond = Levels[1]["BC"], Levels[1]["conditions"]
    for f, v in ((2, cond["y"][0]), (3, cond["y"][1]), (0, cond["x"][0]), (1, cond["x"][1]),
                 (4, cond["z"][0])):
        out[bc[f]] = v
    return out


def assignBCsFine(RHS, TfAll, BC):
    """cF:1598-1620 (host helper; the kernels apply it fused)."""
    out = np.array(RHS, dtype=np.float32, copy=True)
    for f in (2, 3, 0, 1, 4):
        out[BC[f]] = np.asarray(TfAll)[BC[f]]
    return out


def substitute_Tbar(Tbar, _idx, _val):
    """cF:1848-1865."""
    out = np.array(Tbar, dtype=np.float32, copy=True)
    out[_idx:] = _val
    return out


# ---------------------------------------------------------------------------
# Time loop (engine calls)
# ---------------------------------------------------------------------------


def moveEverything(v, vstart, Levels, move_v, LInterp, L1L2Eratio, L2L3Eratio, height):
    """cF:2400-2510: move L3/L2 with the laser, re-interpolate T0 and T' onto
    the moved meshes and pull S1/S2 from Level0.  Shapes and LInterp are
    handles (the kernels evaluate both on the fly)."""
    e = _eng(Levels)
    vv = np.asarray(v, dtype=np.float32).reshape(-1)[:3].copy()
    vs = np.asarray(vstart, dtype=np.float32).reshape(-1)[:3].copy()
    mv = np.zeros(3, np.int32)
    N.check(N.lib().gm_engine_move(e.h, N.fptr(vv), N.fptr(vs), mv.ctypes.data_as(N._IP)))
    Shapes = [Shape(Levels[1], Levels[2]), Shape(Levels[1], Levels[3]),
              Shape(Levels[2], Levels[3])]
    LInterp = [("L1->L2",), ("L2->L3",)]
    return Levels, Shapes, LInterp, [int(x) for x in mv]


def subcycleGOMELT(Levels, ne_nn, Shapes, substrate, LInterp, tmp_ne_nn, laser_position,
                   properties, laserP, subcycle, max_accum_L3=None, accum_L3=None):
    """cF:3224-3632: one L1 step with N2 x N3 subcycling (predictor, corrector,
    melt-time accumulation) as one hipGraph launch.
    Returns (Levels, L2all, L3all, L3pall, max_accum_L3, accum_L3) as the
    reference does; L2all/L3all/L3pall are device views of the corrector's
    per-L2-substep fields.

    Melt time, as in the reference: ``max_accum_L3`` and ``accum_L3`` are the
    caller's maps on the L3 window nodes (the driver passes
    ``max_accum_time[Levels[0]["idx"]]``, go_melt.py:447-448); the updated
    window values come back as host arrays (the driver scatters them,
    go_melt.py:453-455).  Passing None for both keeps melt time in the
    engine's own Level0 maps (``Levels[0]["accum_time"]`` etc., updated in
    place on the device), and the last two returns are those maps."""
    e = _eng(Levels)
    e.set_layer(tmp_ne_nn, substrate)
    rows = np.array(laser_position, dtype=np.float32).reshape(-1, 7).copy()
    if laserP is not None:
        rows[:, 6] = np.asarray(laserP, dtype=np.float32).reshape(-1)
    external = max_accum_L3 is not None or accum_L3 is not None
    e.set_accum_external(external)
    if external:
        nn3 = int(Levels[3]["nn"])
        for fld, arr in ((N.GM_MAX_ACCUM_L3, max_accum_L3), (N.GM_ACCUM_L3, accum_L3)):
            dst = e.field(3, fld)
            src = np.zeros(nn3, np.float32) if arr is None else np.asarray(arr, np.float32)
            if src.size != nn3:
                raise ValueError(f"melt-time window arrays need {nn3} values (the L3 nodes), "
                                 f"got {src.size}")
            N.check(N.lib().gm_memcpy_h2d(dst.ptr, np.ascontiguousarray(src).ctypes.data,
                                          dst.nbytes, e.stream.ptr))
    op = getattr(e, "l3_operator", None)
    if op is None:
        N.check(N.lib().gm_engine_subcycle(e.h, N.fptr(rows), rows.shape[0]))
    else:
        op.subcycle(e, rows)
    L2all = [e.field(k, N.GM_HIST_L2) for k in range(e.N2)]
    L3all = [e.field(k, N.GM_HIST_L3) for k in range(e.N2)]
    L3pall = [e.field(k, N.GM_HIST_TP3) for k in range(e.N2)]
    if external:
        return (Levels, L2all, L3all, L3pall, e.field(3, N.GM_MAX_ACCUM_L3).numpy(e.stream.ptr),
                e.field(3, N.GM_ACCUM_L3).numpy(e.stream.ptr))
    return (Levels, L2all, L3all, L3pall, Levels[0]["max_accum_time"],
            Levels[0]["accum_time"])


def stepGOMELT(Levels, ne_nn, tmp_ne_nn, Shapes, LInterp, v, properties, dt, laserP,
               substrate):
    """cF:2304-2397: equal-dt predictor/corrector over all levels.
    Returns (Levels, _resetmask) as the reference does: the L3 nodes whose S2
    went from 0 to 1 in this step (cF:2394-2397), a device view of uint8 that
    reads as a host array (``all_reset > 0``, go_melt.py:343).  The view
    holds the mask of the latest step.  While the engine keeps melt time
    itself (subcycleGOMELT never given window arrays), the step also applies
    the driver's melt-time bookkeeping (go_melt.py:339-357) to its Level0
    maps."""
    e = _eng(Levels)
    e.set_layer(tmp_ne_nn, substrate)
    row = np.zeros(7, np.float32)
    row[:len(v)] = np.asarray(v, dtype=np.float32)[:7]
    row[5] = dt
    row[6] = laserP
    N.check(N.lib().gm_engine_step(e.h, N.fptr(row)))
    return Levels, e.field(3, N.GM_RESET_L3)


def stepGOMELTDwellTime(Levels, tmp_ne_nn, ne_nn, properties, dt, substrate):
    """cF:2617-2664: L1-only step."""
    e = _eng(Levels)
    e.set_layer(tmp_ne_nn, substrate)
    N.check(N.lib().gm_engine_step_dwell(e.h, float(dt)))
    return Levels


def melting_temp(temps, delt_T, T_melt, accum_time, idx):
    """cF:3696-3712: accum_time[idx] += (temps > T_melt) * delt_T (host form
    for a driver that keeps its own maps; the engine applies it on the
    device otherwise).  An ``accum_time`` with a ``.at`` functional update
    (jax.Array style) is updated through it, as the reference does."""
    add = np.array(np.asarray(temps) > T_melt) * delt_T
    if hasattr(accum_time, "at"):
        return accum_time.at[idx].add(add)
    out = np.array(accum_time, copy=True)
    np.add.at(out, idx, add)
    return out


def energyBalance(Level, ne, k, rhocp, dt, T, Fc, Corr, Tnew, bc_faces=0, T_clamp=-np.inf,
                  stream=None):
    """Energy audit of an explicit update T -> Tnew of a level (8 FP64 sums).

    An addition with no reference counterpart; it checks the identity of
    solveMatrixFreeFE (cF:582-642): the unconstrained update
    Tt = T + dt (Fc + Corr - K T) / C satisfies sum C (Tt - T) = dt sum F.
    Arguments as solveMatrixFreeFE, plus the Dirichlet faces (bits x0, x1,
    y0, y1, z0) and the floor T_clamp whose nodes count as held.  Returns
    [free C dT, held C dTt, dt sum F, dt sum K T, held reaction, scale,
    residual, held count] (one call of gm_energy_balance)."""
    g = _grid_of(Level)
    kd, cd, Td, Tnd = _dev(k), _dev(rhocp), _dev(T), _dev(Tnew)
    Fd = None if (np.isscalar(Fc) and Fc == 0) else _dev(Fc)
    Cd = None if (np.isscalar(Corr) and Corr == 0) else _dev(Corr)
    out = DeviceArray(8, np.float64)
    sp = stream.ptr if stream is not None else None
    N.check(N.lib().gm_energy_balance(C.byref(g), int(ne), kd.ptr, cd.ptr, float(dt), Td.ptr,
                                      Fd.ptr if Fd else None, Cd.ptr if Cd else None, Tnd.ptr,
                                      int(bc_faces), float(T_clamp), out.ptr, sp))
    N.sync(sp)
    return out.numpy(sp)


def meltPoolDimensions(Level, T, T_liquidus, stream=None):
    """Melt-pool size of a level's field, in mesh units (mm).

    An addition with no reference counterpart: the reference only forms
    S2 = T >= T_liquidus (cF:2585).  Returns (depth, width, length, n_molten,
    ylo, yhi, xlo, xhi): depth below the top node plane of the lowest
    T_liquidus crossing on a vertical grid line, and the y and x extents of
    the molten top plane, with linear interpolation along edges (one call of
    gm_melt_pool).  ``Level`` gives the mesh (its current ``node_coords``);
    ``T`` is a device or host array of its nodes."""
    g = _grid_of(Level)
    Td = _dev(T)
    out = DeviceArray(8)
    sp = stream.ptr if stream is not None else None
    N.check(N.lib().gm_melt_pool(C.byref(g), Td.ptr, float(T_liquidus), out.ptr, sp))
    N.sync(sp)
    d = out.numpy(sp)
    return (float(d[0]), float(d[1]), float(d[2]), int(round(float(d[3]))), float(d[4]),
            float(d[5]), float(d[6]), float(d[7]))


def printLevelMaxMin(Ls, Lnames):
    """cF:3635-3665: print per-level min/max; exit(1) on NaN, <= 0 or > 1e5."""
    e = _eng(Ls)
    mm = np.zeros(6, np.float32)
    N.check(N.lib().gm_engine_minmax(e.h, N.fptr(mm)))
    print("Temps:", end=" ")
    bad = False
    for i in range(3):
        lo, hi = float(mm[2 * i]), float(mm[2 * i + 1])
        for v, tag in ((hi, "Lmax"), (lo, "Lmin")):
            if not np.isfinite(v) or v <= 0 or v > 1e5:
                print(f"\nTerminating program: {tag} for {Lnames[i]} is NaN, 0, or invalid.")
                bad = True
        print(f"{Lnames[i]}: [{lo:.2f}, {hi:.2f}]", end=" ")
    if bad:
        sys.exit(1)
    print("")
    return mm


# ---------------------------------------------------------------------------
# Output (cF:1945-2057, 3668-3693): VTK rectilinear grids, own writer
# ---------------------------------------------------------------------------


_VTK_TYPES = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64",
              np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64",
              np.dtype(np.uint8): "UInt8", np.dtype(np.int8): "Int8"}


def _write_vtr(path, coords, point_data):
    """VTK XML RectilinearGrid with raw appended data, the layout
    pyevtk.hl.gridToVTK writes (cF:1982 calls it; pyevtk is not needed here):
    version 1.0, UInt64 block headers, the three coordinate axes first, then
    the point data in the order given, each array in its own dtype with its
    values in Fortran order (x fastest); PointData names its first array as
    the Scalars attribute.  Writes ``path + ".vtr"``."""
    nx, ny, nz = (len(c) for c in coords)
    axes = [np.ascontiguousarray(np.asarray(c)) for c in coords]
    arrays = [(n, a) for n, a in zip(("x_coordinates", "y_coordinates", "z_coordinates"), axes)]
    for name, a in point_data.items():
        a = np.asarray(a)
        arrays.append((name, a.reshape(-1, order="F") if a.ndim > 1 else a))
    offs, off = [], 0
    for _, a in arrays:
        offs.append(off)
        off += 8 + a.nbytes
    ext = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"

    def darr(name, a, o):
        return (f'<DataArray type="{_VTK_TYPES[a.dtype]}" Name="{name}" '
                f'NumberOfComponents="1" format="appended" offset="{o}"/>\n')

    with open(path + ".vtr", "wb") as fh:
        w = lambda t: fh.write(t.encode())  # noqa: E731
        w('<?xml version="1.0"?>\n<VTKFile type="RectilinearGrid" version="1.0" '
          'byte_order="LittleEndian" header_type="UInt64">\n')
        w(f'<RectilinearGrid WholeExtent="{ext}">\n<Piece Extent="{ext}">\n<Coordinates>\n')
        for (name, a), o in zip(arrays[:3], offs[:3]):
            w(darr(name, a, o))
        w("</Coordinates>\n")
        if len(arrays) > 3:
            w(f'<PointData Scalars="{arrays[3][0]}">\n')
            for (name, a), o in zip(arrays[3:], offs[3:]):
                w(darr(name, a, o))
            w("</PointData>\n")
        w('</Piece>\n</RectilinearGrid>\n<AppendedData encoding="raw">\n_')
        for _, a in arrays:
            fh.write(struct.pack("<Q", a.nbytes))
            fh.write(np.ascontiguousarray(a).astype(a.dtype.newbyteorder("<")).tobytes())
        w("\n</AppendedData>\n</VTKFile>\n")


def _fields(Level):
    """T0 and S1 as (nx, ny, nz) float32 arrays (cF:1972-1977)."""
    nx, ny, nz = Level["nodes"]
    T = np.asarray(Level["T0"], np.float32).reshape(nz, ny, nx).transpose(2, 1, 0)
    S = np.asarray(Level["S1"]).astype(np.float32).reshape(nz, ny, nx).transpose(2, 1, 0)
    return T, S


def _axes(Level, zoffset):
    nc = Level["node_coords"]
    return [np.asarray(nc[0], np.float32), np.asarray(nc[1], np.float32),
            (np.asarray(nc[2], np.float32) - np.float32(zoffset)).astype(np.float32)]


def saveResult(Level, save_str, record_lab, save_path, zoffset):
    """cF:1945-1982: <save_path><save_str><record_lab:08>.vtr with the point
    data "Temperature (K)" and "State (Powder/Solid)"."""
    T, S = _fields(Level)
    _write_vtr(f"{save_path}{save_str}{record_lab:08}", _axes(Level, zoffset),
               {"Temperature (K)": T, "State (Powder/Solid)": S})


def saveFinalResult(Level, save_str, save_path, zoffset):
    """cF:1985-2021: <save_path><save_str>Final.vtr."""
    T, S = _fields(Level)
    _write_vtr(f"{save_path}{save_str}Final", _axes(Level, zoffset),
               {"Temperature (K)": T, "State (Powder/Solid)": S})


def saveState(Level, save_str, record_lab, save_path, zoffset):
    """cF:2024-2057: the state field alone."""
    nx, ny, nz = Level["nodes"]
    S = np.asarray(Level["S1"]).astype(np.float32).reshape(nz, ny, nx).transpose(2, 1, 0)
    _write_vtr(f"{save_path}{save_str}{record_lab:08}", _axes(Level, zoffset),
               {"State (Powder/Solid)": S})


def saveResults(Levels, Nonmesh, savenum):
    """cF:3668-3682."""
    if Nonmesh["output_files"] == 1:
        _eng(Levels).sync_coords()
        if savenum == 1 or (np.mod(savenum, Nonmesh["Level1_record_step"]) == 1
                            or Nonmesh["Level1_record_step"] == 1):
            saveResult(Levels[1], "Level1_", savenum, Nonmesh["save_path"], 2e-3)
        saveResult(Levels[2], "Level2_", savenum, Nonmesh["save_path"], 1e-3)
        saveResult(Levels[3], "Level3_", savenum, Nonmesh["save_path"], 0)
        print(f"Saved Levels_{savenum:08}")


def saveResultsFinal(Levels, Nonmesh):
    """cF:3685-3693."""
    if Nonmesh["output_files"] == 1:
        _eng(Levels).sync_coords()
        saveFinalResult(Levels[1], "Level1_", Nonmesh["save_path"], 2e-3)
        saveFinalResult(Levels[2], "Level2_", Nonmesh["save_path"], 1e-3)
        saveFinalResult(Levels[3], "Level3_", Nonmesh["save_path"], 0)
        print("Saved Final Results")
