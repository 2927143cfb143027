"""CPU oracle: a NumPy restatement of GO-MELT's multilevel explicit FE heat step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or, in bench.py, as the timed CPU port).  The product path
(``go-melt_amd/``) never imports it and fails loudly without its HIP library.

What it is
----------
A line-by-line *semantic* restatement of the reference's JAX hot path, written
with NumPy vectorisation instead of ``jax.vmap``/``lax.scan``.  It keeps the
reference's element-loop structure on purpose (gather the 8 element nodes,
evaluate the 8x8 element operator, scatter-add with ``bincount``), and it
tabulates the coarse shape functions at fine quadrature points exactly as the
reference's ``Shapes`` tensors do, so that it is an independent check of the
GPU design, which evaluates the same quantities node-centrically and
separably.  Every function cites the reference lines it follows
(``cF`` = go_melt/computeFunctions.py, ``gm`` = go_melt/go_melt.py,
``cP`` = go_melt/createPath.py).

Precision
---------
``set_precision(np.float64)`` (default) is the FP64 oracle.
``set_precision(np.float32)`` emulates the reference, which never enables x64
and so runs in FP32 with integer indices.  Scatter-adds go through
``np.bincount`` (accumulates in FP64, then the result is cast back).

How it is pinned (tests/test_oracle_*.py)
-----------------------------------------
The reference ships no tests or golden outputs and JAX is not installed, so
each function is pinned by closed forms, brute force and invariants:
  * element operator: exact box-element stiffness/mass (tensor products of the
    1-D linear element), patch test, constant-field invariance, the discrete
    Fourier symbol of the explicit update, dense brute-force assembly;
  * sources: closed-form Gaussian integral (eta*P into the half-space);
  * surface flux: uniform-temperature closed form;
  * interpolation: exact reproduction of trilinear fields, outside -> 0;
  * restriction (Shapes): partition of unity, gradient reproduction and a
    brute-force triple loop on tiny meshes;
  * T': zero when the fine field is the coarse interpolant;
  * window shift: integer truncation arithmetic;
  * toolpath: hand-derived row counts and fixed-width format.
Orchestration (subcycleGOMELT / stepGOMELT / moveEverything / driver) can only
be pinned by invariants (uniform field stays uniform, energy identity,
single-level limit): its sequencing is restated from the reference text and is
"parity unpinned" beyond those invariants.
"""

from __future__ import annotations

import copy
import json
import math
import os
import re

import numpy as np

FLOAT = np.float64


def set_precision(dtype) -> None:
    """Select FP64 (oracle) or FP32 (reference emulation) arithmetic."""
    global FLOAT
    FLOAT = np.dtype(dtype).type


def F(x):
    """Cast to the working float type (keeps NumPy scalars from promoting)."""
    return np.asarray(x, dtype=FLOAT)


# ---------------------------------------------------------------------------
# Mesh and configuration set-up  (cF:18-64, 115-579, 858-925, 1886-1942)
# ---------------------------------------------------------------------------

def calcNumNodes(elements):
    """cF:18-29."""
    return [elements[0] + 1, elements[1] + 1, elements[2] + 1]


def createMesh3D(x, y, z):
    """cF:32-64: 1-D node axes and per-axis connectivity (8 local nodes).

    Local node order 0(000) 1(100) 2(110) 3(010) 4(001) 5(101) 6(111) 7(011).
    """
    axes = [np.linspace(a[0], a[1], int(a[2])).astype(FLOAT) for a in (x, y, z)]
    xb = np.array([0, 1, 1, 0, 0, 1, 1, 0])
    yb = np.array([0, 0, 1, 1, 0, 0, 1, 1])
    zb = np.array([0, 0, 0, 0, 1, 1, 1, 1])
    conn = []
    for a, bits in zip((x, y, z), (xb, yb, zb)):
        ne = int(a[2]) - 1
        conn.append(np.arange(ne)[:, None] + bits[None, :])
    return axes, conn


def calc_length_h(lv):
    """cF:1916-1942."""
    b = lv["bounds"]
    L = [b["x"][1] - b["x"][0], b["y"][1] - b["y"][0], b["z"][1] - b["z"][0]]
    h = [L[i] / lv["elements"][i] for i in range(3)]
    return L, h


def getBCindices(lv):
    """cF:520-559: [west, east, south, north, bottom, top] node index sets."""
    nx, ny, nz = lv["nodes"]
    nn = lv["nn"]
    bidx = np.arange(0, nx * ny)
    tidx = np.arange(nx * ny * (nz - 1), nn)
    widx = np.arange(0, nn, nx)
    eidx = np.arange(nx - 1, nn, nx)
    sidx = (np.arange(0, nx)[:, None] + (nx * ny * np.arange(0, nz))[None, :]).reshape(-1)
    nidx = (np.arange(nx * (ny - 1), nx * ny)[:, None]
            + (nx * ny * np.arange(0, nz))[None, :]).reshape(-1)
    return [widx, eidx, sidx, nidx, bidx, tidx]


def find_max_const(coarse, fine):
    """cF:1886-1913: window move limits, relative to the coarse (L1) bounds."""
    cb, fb = coarse["bounds"], fine["bounds"]
    iE = cb["x"][1] - fb["x"][1]
    iN = cb["y"][1] - fb["y"][1]
    iT = cb["z"][1] - fb["z"][1]
    iW = cb["x"][0] - fb["x"][0]
    iS = cb["y"][0] - fb["y"][0]
    iB = cb["z"][0] - fb["z"][0]
    return [iW, iE], [iS, iN], [iB, iT]


def getCoarseNodesInFineRegion(xnf, xnc):
    """cF:858-887: coarse node indices covering the fine axis extent."""
    xfmin, xfmax = xnf.min(), xnf.max()
    xcmin, xcmax = xnc.min(), xnc.max()
    hc = (xcmax - xcmin) / (xnc.size - 1)
    lo = np.round((xfmin - xcmin) / hc)
    hi = np.round((xfmax - xcmin) / hc) + 1
    return np.arange(lo, hi).astype(np.int64)


def getCoarseNodesInLargeFineRegion(xnc, xnf):
    """cF:890-925: fine node indices that coincide with coarse nodes."""
    xfmin, xfmax = xnf.min(), xnf.max()
    xcmin, xcmax = xnc.min(), xnc.max()
    hf = (xfmax - xfmin) / (xnf.size - 1)
    hc = (xcmax - xcmin) / (xnc.size - 1)
    lo = np.round((xcmin - xfmin) / hf)
    hi = np.round((xcmax - xfmin) / hf) + 1
    step = int(np.round(hc / hf))
    return np.arange(lo, hi, step).astype(np.int64)


def getOverlapRegion(node_axes, nx, ny):
    """cF:1642-1669: flattened (x-fastest) indices of a tensor index box."""
    ax, ay, az = [np.asarray(a, dtype=np.int64) for a in node_axes]
    return (ax[None, None, :] + ay[None, :, None] * nx
            + az[:, None, None] * nx * ny).reshape(-1)


def SetupProperties(prop):
    """cF:267-345: material/laser/surface properties with the reference defaults.

    Extension (not in the reference): optional piecewise-linear tables
    ``thermal_conductivity_bulk_table`` and ``heat_capacity_solid_table``
    ([[T knots], [values]]) replace the linear bulk-k and solid-cp laws.
    """
    p = dict(prop)
    p["k_powder"] = prop.get("thermal_conductivity_powder", 0.4)
    p["k_bulk_coeff_a0"] = prop.get("thermal_conductivity_bulk_a0", 4.23)
    p["k_bulk_coeff_a1"] = prop.get("thermal_conductivity_bulk_a1", 0.016)
    p["k_fluid_coeff_a0"] = prop.get("thermal_conductivity_fluid_a0", 29.0)
    p["cp_solid_coeff_a0"] = prop.get("heat_capacity_solid_a0", 383.1)
    p["cp_solid_coeff_a1"] = prop.get("heat_capacity_solid_a1", 0.174)
    p["cp_mushy"] = prop.get("heat_capacity_mushy", 3235.0)
    p["cp_fluid"] = prop.get("heat_capacity_fluid", 769.0)
    p["rho"] = prop.get("density", 8.0e-6)
    p["laser_radius"] = prop.get("laser_radius", 0.110)
    p["laser_depth"] = prop.get("laser_depth", 0.05)
    p["laser_power"] = prop.get("laser_power", 300.0)
    p["laser_eta"] = prop.get("laser_absorptivity", 0.25)
    p["laser_center"] = prop.get("laser_center", [])
    p["T_amb"] = prop.get("T_amb", 353.15)
    p["T_solidus"] = prop.get("T_solidus", 1554.0)
    p["T_liquidus"] = prop.get("T_liquidus", 1625.0)
    p["T_boiling"] = prop.get("T_boiling", 3038.0)
    p["h_conv"] = prop.get("h_conv", 1.473e-5) * 1e6
    p["vareps"] = prop.get("emissivity", 0.600)
    p["evc"] = prop.get("evaporation_coefficient", 0.82)
    p["kb"] = prop.get("boltzmann_constant", 1.38e-23)
    p["mA"] = prop.get("atomic_mass", 7.9485017e-26)
    p["Lev"] = prop.get("latent_heat_evap", 4.22e6)
    p["molar_mass"] = prop.get("molar_mass", 58.69) * 1e-3
    p["layer_height"] = prop.get("layer_height", 0.04)
    p["sigma_sb"] = 5.67e-8
    p["gas_const"] = 8.314
    p["atmospheric_pressure"] = 101325
    p["CM_coeff"] = p["molar_mass"] / (2.0 * math.pi * p["gas_const"])
    p["CT_coeff"] = p["Lev"] * p["molar_mass"] / p["gas_const"]
    p["CP_coeff"] = 0.54 * p["atmospheric_pressure"]
    p["k_table"] = prop.get("thermal_conductivity_bulk_table")
    p["cp_table"] = prop.get("heat_capacity_solid_table")
    return p


def SetupNonmesh(nm):
    """cF:348-419 (the output directory is not created here)."""
    n = dict(nm)
    n["timestep_L3"] = nm.get("timestep_L3", 1e-5)
    n["subcycle_num_L2"] = nm.get("subcycle_num_L2", 1)
    n["subcycle_num_L3"] = nm.get("subcycle_num_L3", 1)
    n["dwell_time"] = nm.get("dwell_time", 0.1)
    n["Level1_record_step"] = nm.get("Level1_record_step", 1)
    n["save_path"] = nm.get("save_path", "results/")
    n["output_files"] = nm.get("output_files", 1)
    n["toolpath"] = nm.get("toolpath", "laserPath.txt")
    n["wait_time"] = nm.get("wait_time", 500.0)
    n["layer_num"] = nm.get("layer_num", 0)
    n["restart_layer_num"] = nm.get("restart_layer_num", 10000)
    n["info_T"] = nm.get("info_T", 0)
    n["laser_velocity"] = nm.get("laser_velocity", 500)
    n["wait_track"] = nm.get("wait_track", 0.0)
    n["record_step"] = nm.get("record_step", n["subcycle_num_L2"] * n["subcycle_num_L3"])
    n["gcode"] = nm.get("gcode", "./examples/gcodefiles/defaultName.gcode")
    n["dwell_time_multiplier"] = nm.get("dwell_time_multiplier", 1)
    n["use_txt"] = nm.get("use_txt", 0)
    return n


def SetupLevels(solver_input, properties):
    """cF:115-264: Level0 (state grid) and Levels 1-3 with fields and overlaps."""
    levels = {}
    for name in ("Level1", "Level2", "Level3"):
        src = solver_input.get(name, {})
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
            (b["x"][0], b["x"][1], lv["nodes"][0]),
            (b["y"][0], b["y"][1], lv["nodes"][1]),
            (b["z"][0], b["z"][1], lv["nodes"][2]))
        nn = lv["nn"]
        lv["T"] = np.full(nn, properties["T_amb"], dtype=FLOAT)
        lv["T0"] = np.full(nn, properties["T_amb"], dtype=FLOAT)
        lv["S1"] = np.zeros(nn, dtype=FLOAT)
        lv["S2"] = np.zeros(nn, dtype=bool)
        lv["k"] = np.full(nn, properties["k_powder"], dtype=FLOAT)
        lv["rhocp"] = np.full(nn, properties["cp_solid_coeff_a0"] * properties["rho"], dtype=FLOAT)
        levels[name] = lv
    L1, L2, L3 = levels["Level1"], levels["Level2"], levels["Level3"]
    L1["S1_storage"] = np.zeros((int(round(L1["h"][2] / properties["layer_height"])), L1["nn"]),
                                dtype=FLOAT)
    for lv in (L2, L3):
        ix, iy, iz = find_max_const(L1, lv)
        lv["bounds"]["ix"], lv["bounds"]["iy"], lv["bounds"]["iz"] = ix, iy, iz
        lv["init_node_coors"] = copy.deepcopy(lv["node_coords"])
        lv["Tprime"] = np.zeros(lv["nn"], dtype=FLOAT)
        lv["Tprime0"] = np.zeros(lv["nn"], dtype=FLOAT)
    # L1 z axis is shifted by whole layer heights until a node meets the L2 top
    L1["orig_node_coords"] = copy.deepcopy(L1["node_coords"])
    tmp = copy.deepcopy(L1["orig_node_coords"])
    while not np.isclose(tmp[2] - L2["node_coords"][2][-1], 0, atol=1e-4).any():
        tmp[2] = (tmp[2] + properties["layer_height"]).astype(FLOAT)
    L1["node_coords"] = copy.deepcopy(tmp)

    def get_overlap(fine, coarse):
        nodes = [getCoarseNodesInFineRegion(fine["node_coords"][i], coarse["node_coords"][i])
                 for i in range(3)]
        coors = [coarse["node_coords"][i][nodes[i]].astype(FLOAT) for i in range(3)]
        return nodes, coors

    L2["orig_overlap_nodes"], L2["orig_overlap_coors"] = get_overlap(L2, L1)
    L3["orig_overlap_nodes"], L3["orig_overlap_coors"] = get_overlap(L3, L2)
    for lv in (L2, L3):
        lv["overlapNodes"] = copy.deepcopy(lv["orig_overlap_nodes"])
        lv["overlapCoords"] = copy.deepcopy(lv["orig_overlap_coors"])

    L0 = {}
    L0["elements"] = [round(L1["length"][0] / L3["h"][0]), round(L1["length"][1] / L3["h"][1]),
                      round(L2["length"][2] / L3["h"][2])]
    L0["nodes"] = calcNumNodes(L0["elements"])
    L0["ne"] = L0["elements"][0] * L0["elements"][1] * L0["elements"][2]
    L0["nn"] = L0["nodes"][0] * L0["nodes"][1] * L0["nodes"][2]
    L0["node_coords"], L0["connect"] = createMesh3D(
        (L1["bounds"]["x"][0], L1["bounds"]["x"][1], L0["nodes"][0]),
        (L1["bounds"]["y"][0], L1["bounds"]["y"][1], L0["nodes"][1]),
        (L2["bounds"]["z"][0], L2["bounds"]["z"][1], L0["nodes"][2]))
    L0["orig_node_coords"] = copy.deepcopy(L0["node_coords"])
    L0["orig_overlap_nodes"], L0["orig_overlap_coors"] = get_overlap(L3, L0)
    L0["overlapNodes"] = copy.deepcopy(L0["orig_overlap_nodes"])
    L0["overlapCoords"] = copy.deepcopy(L0["orig_overlap_coors"])
    L0["orig_overlap_nodes_L2"] = [
        getCoarseNodesInLargeFineRegion(L2["node_coords"][i], L0["node_coords"][i])
        for i in range(3)]
    L0["orig_overlap_coors_L2"] = [L0["node_coords"][i][L0["orig_overlap_nodes_L2"][i]]
                                   for i in range(3)]
    L0["overlapNodes_L2"] = copy.deepcopy(L0["orig_overlap_nodes_L2"])
    L0["overlapCoords_L2"] = copy.deepcopy(L0["orig_overlap_coors_L2"])
    L0["S1"] = np.zeros(L0["nn"], dtype=FLOAT)
    L0["S2"] = np.zeros(L0["nn"], dtype=bool)
    L0["idx"] = getOverlapRegion(L0["overlapNodes"], L0["nodes"][0], L0["nodes"][1])
    L0["idx_L2"] = getOverlapRegion(L0["overlapNodes_L2"], L0["nodes"][0], L0["nodes"][1])
    L0["layer_idx_delta"] = int(round(properties["layer_height"] / L3["h"][2]))
    return [L0, L1, L2, L3]


def getStaticNodesAndElements(L):
    """cF:447-470: (ne2, ne3, nn1, nn2, nn3)."""
    return (L[2]["ne"], L[3]["ne"], L[1]["nn"], L[2]["nn"], L[3]["nn"])


def getStaticSubcycle(N):
    """cF:473-492: (N2, N3, N2*N3, float N2, float N3, float N23)."""
    N2, N3 = N["subcycle_num_L2"], N["subcycle_num_L3"]
    return (N2, N3, N2 * N3, float(N2), float(N3), float(N2 * N3))


def calcStaticTmpNodesAndElements(L, v):
    """cF:495-517: active L1 elements/nodes at or below the laser z."""
    n_active_z = int(np.sum(L[1]["node_coords"][2] <= v[2] + 1e-5))
    tmp_ne = L[1]["elements"][0] * L[1]["elements"][1] * (n_active_z - 1)
    tmp_nn = L[1]["nodes"][0] * L[1]["nodes"][1] * n_active_z
    return (int(tmp_ne), int(tmp_nn))


def getSubstrateNodes(Levels):
    """cF:562-579: per level, (#node layers with z < 1e-5) * nx * ny."""
    return tuple(int(np.sum(L["node_coords"][2] < 1e-5) * L["nodes"][0] * L["nodes"][1])
                 for L in Levels[:4])


# ---------------------------------------------------------------------------
# Reference element  (cF:645-855, 3135-3166)
# ---------------------------------------------------------------------------

_XB = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=np.float64)
_YB = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=np.float64)
_ZB = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=np.float64)


def convert2XYZ(i, ne_x, ne_y, nn_x, nn_y):
    """cF:645-689: element id -> (ix, iy, iz) and its 8 node ids, shape (8, ...)."""
    i = np.asarray(i)
    iz = i // (ne_x * ne_y)
    iy = (i // ne_x) - iz * ne_y
    ix = i % ne_x
    base = ix + iy * nn_x + iz * nn_x * nn_y
    dx, dy, dz = 1, nn_x, nn_x * nn_y
    idx = np.stack([base, base + dx, base + dx + dy, base + dy,
                    base + dz, base + dx + dz, base + dx + dy + dz, base + dy + dz])
    return ix, iy, iz, idx


def computeQuad3dFemShapeFunctions(coords):
    """cF:692-775: N (8q x 8a), dNdx (8q x 8a x 3), wq (8,1) at 2-pt Gauss.

    The Jacobian is diagonal and taken from the first Gauss point.
    """
    g = 1.0 / math.sqrt(3.0)
    kq, eq, zq = g * _XB, g * _YB, g * _ZB
    _k = 1 + kq[:, None] * _XB[None, :]
    _e = 1 + eq[:, None] * _YB[None, :]
    _z = 1 + zq[:, None] * _ZB[None, :]
    N = 0.125 * _k * _e * _z
    dNdk = 0.125 * _XB[None, :] * _e * _z
    dNde = 0.125 * _YB[None, :] * _k * _z
    dNdz = 0.125 * _ZB[None, :] * _k * _e
    c = np.asarray(coords, dtype=np.float64)
    jx = (dNdk @ c[:, 0])[0]
    jy = (dNde @ c[:, 1])[0]
    jz = (dNdz @ c[:, 2])[0]
    dNdx = np.stack([dNdk / jx, dNde / jy, dNdz / jz], axis=2)
    wq = np.full((8, 1), jx * jy * jz)
    return N.astype(FLOAT), dNdx.astype(FLOAT), wq.astype(FLOAT)


def computeQuad2dFemShapeFunctions(coords):
    """cF:778-855: 4-node face element at 2x2 Gauss; Jacobian from nodes 4-7."""
    g = 1.0 / math.sqrt(3.0)
    ki = np.array([-1, 1, 1, -1], dtype=np.float64)
    ei = np.array([-1, -1, 1, 1], dtype=np.float64)
    _k = 1 + (g * ki)[:, None] * ki[None, :]
    _e = 1 + (g * ei)[:, None] * ei[None, :]
    N = 0.25 * _k * _e
    dNdk = 0.25 * ki[None, :] * _e
    dNde = 0.25 * ei[None, :] * _k
    c = np.asarray(coords, dtype=np.float64)
    jx = (dNdk @ c[4:, 0])[0]
    jy = (dNde @ c[4:, 1])[0]
    dNdx = np.stack([dNdk / jx, dNde / jy], axis=2)
    wq = np.full((4, 1), jx * jy)
    return N.astype(FLOAT), dNdx.astype(FLOAT), wq.astype(FLOAT)


def getSampleCoords(lv):
    """cF:3135-3148: node coordinates of element 0, shape (8, 3)."""
    x = lv["node_coords"][0][lv["connect"][0][0, :]]
    y = lv["node_coords"][1][lv["connect"][1][0, :]]
    z = lv["node_coords"][2][lv["connect"][2][0, :]]
    return np.stack([x, y, z], axis=1)


def _element_nodes(lv, ne):
    """Node ids (8, ne) and element axis indices for elements 0..ne-1."""
    ex, ey = lv["elements"][0], lv["elements"][1]
    nx, ny = lv["nodes"][0], lv["nodes"][1]
    return convert2XYZ(np.arange(ne), ex, ey, nx, ny)


def getQuadratureCoords(lv, ix, iy, iz, Nf):
    """cF:3151-3166: physical Gauss point coordinates, each (8q, n_elem)."""
    cx = lv["node_coords"][0][lv["connect"][0][ix, :]]   # (n, 8)
    cy = lv["node_coords"][1][lv["connect"][1][iy, :]]
    cz = lv["node_coords"][2][lv["connect"][2][iz, :]]
    return Nf @ cx.T, Nf @ cy.T, Nf @ cz.T


def bincount(idx, data, nn):
    """cF:1623-1639 (jnp.bincount scatter-add)."""
    return np.bincount(np.asarray(idx).reshape(-1), weights=np.asarray(data).reshape(-1),
                       minlength=nn)[:nn].astype(FLOAT)


# ---------------------------------------------------------------------------
# Level operator  (cF:582-642)
# ---------------------------------------------------------------------------

def solveMatrixFreeFE(lv, nn, ne, k, rhocp, dt, T, Fc, Corr):
    """cF:582-642: explicit lumped-mass update.

    T+ = [sum_e (diag(M_e) - K_e) T_e + Fc + Corr] / sum_e M_e, with element
    means of the nodal k and rho*cp.  Nodes in no active element divide 0/0.
    """
    N, dNdx, wq = computeQuad3dFemShapeFunctions(getSampleCoords(lv))
    w = wq[0, 0]
    NTN = N.T @ N
    BTB = sum(dNdx[:, :, d].T @ dNdx[:, :, d] for d in range(3))
    _, _, _, idx = _element_nodes(lv, ne)            # (8, ne)
    kvec = k[idx].mean(axis=0)                        # (ne,)
    mvec = rhocp[idx].mean(axis=0) / F(dt)
    Mrow = (NTN * w).sum(axis=0)                      # (8,)
    Me = mvec[:, None] * Mrow[None, :]                # (ne, 8)
    Te = T[idx].T                                     # (ne, 8)
    KT = (Te @ (BTB * w).T) * kvec[:, None]
    aT = Me * Te - KT
    newT = bincount(idx.T, aT, nn)
    newM = bincount(idx.T, Me, nn)
    with np.errstate(invalid="ignore", divide="ignore"):
        return ((newT + Fc + Corr) / newM).astype(FLOAT)


# ---------------------------------------------------------------------------
# Material state  (cF:2513-2614)
# ---------------------------------------------------------------------------

def _table(T, table):
    """Piecewise-linear table with linear extension beyond the end knots."""
    knots = np.asarray(table[0], dtype=np.float64)
    vals = np.asarray(table[1], dtype=np.float64)
    j = np.clip(np.searchsorted(knots, T, side="right") - 1, 0, len(knots) - 2)
    t0, t1 = knots[j], knots[j + 1]
    v0, v1 = vals[j], vals[j + 1]
    slope = (v1 - v0) / (t1 - t0)
    return (v0 + slope * (T - t0)).astype(FLOAT)


def computeStateProperties(T, S1, p, n_sub):
    """cF:2567-2614: phase flags and nodal k (W/mm K) and rho*cp."""
    S2 = T >= p["T_liquidus"]
    S3 = (T > p["T_solidus"]) & (T < p["T_liquidus"])
    S1 = (1.0 * ((S1 > 0.499) | S2)).astype(FLOAT)
    S1[:n_sub] = 1
    fS2 = S2.astype(FLOAT)
    fS3 = S3.astype(FLOAT)
    if p.get("k_table") is not None:
        kb = _table(T, p["k_table"])
    else:
        kb = p["k_bulk_coeff_a1"] * T + p["k_bulk_coeff_a0"]
    k = ((1 - S1) * (1 - fS2) * p["k_powder"] + S1 * (1 - fS2) * kb
         + fS2 * p["k_fluid_coeff_a0"]) / 1000
    if p.get("cp_table") is not None:
        cs = _table(T, p["cp_table"])
    else:
        cs = p["cp_solid_coeff_a1"] * T + p["cp_solid_coeff_a0"]
    cp = (1 - fS2) * (1 - fS3) * cs + fS3 * p["cp_mushy"] + fS2 * p["cp_fluid"]
    return S1, S2, k.astype(FLOAT), (p["rho"] * cp).astype(FLOAT)


def updateStateProperties(Levels, p, substrate):
    """cF:2513-2564."""
    Levels[3]["S1"], Levels[3]["S2"], L3k, L3rc = computeStateProperties(
        Levels[3]["T0"], Levels[3]["S1"], p, substrate[3])
    Levels[2]["S1"], _, L2k, L2rc = computeStateProperties(
        Levels[2]["T0"], Levels[2]["S1"], p, substrate[2])
    val = interpolatePoints(Levels[2], Levels[2]["S1"], Levels[2]["overlapCoords"])
    idx = getOverlapRegion(Levels[2]["overlapNodes"], Levels[1]["nodes"][0], Levels[1]["nodes"][1])
    Levels[1]["S1"] = Levels[1]["S1"].copy()
    Levels[1]["S1"][idx] = val
    Levels[1]["S1"][:substrate[1]] = 1
    _, _, L1k, L1rc = computeStateProperties(Levels[1]["T0"], Levels[1]["S1"], p, substrate[1])
    return Levels, [0, L1k, L2k, L3k], [0, L1rc, L2rc, L3rc]


# ---------------------------------------------------------------------------
# Sources and surface flux  (cF:928-1025, 2207-2301, 2667-2730, 2960-3012)
# ---------------------------------------------------------------------------

def computeSourceFunction(x, y, z, v, p, P):
    """cF:991-1025: separable Gaussian volumetric source (W/mm^3)."""
    pc = F(6 * math.sqrt(3.0)) * F(P) * F(p["laser_eta"])
    rc = F(1.0 / (p["laser_radius"] * math.sqrt(math.pi)))
    dc = F(1.0 / (p["laser_depth"] * math.sqrt(math.pi)))
    rsq, dsq = F(p["laser_radius"] ** 2), F(p["laser_depth"] ** 2)
    Qx = rc * np.exp(F(-3) * (x - F(v[0])) ** 2 / rsq)
    Qy = rc * np.exp(F(-3) * (y - F(v[1])) ** 2 / rsq)
    Qz = dc * np.exp(F(-3) * (z - F(v[2])) ** 2 / dsq)
    return (pc * Qx * Qy * Qz).astype(FLOAT)


def _fine_quadrature_source(L3, v, p, P):
    """Q(x_q)*w_q at every L3 Gauss point, shape (8q, ne3); plus Nf, wq, idx."""
    Nf, _, wq = computeQuad3dFemShapeFunctions(getSampleCoords(L3))
    ix, iy, iz, idx = _element_nodes(L3, L3["ne"])
    x, y, z = getQuadratureCoords(L3, ix, iy, iz, Nf)
    Q = computeSourceFunction(x, y, z, v, p, P)
    return Q * wq, Nf, idx


def computeSourcesL3(L3, v, ne_nn, p, P):
    """cF:2960-3012: L3 nodal source vector."""
    Qw, Nf, idx = _fine_quadrature_source(L3, v, p, P)
    return bincount(idx, Nf @ Qw, ne_nn[4])


def computeSources(L3, v, Shapes, ne_nn, p, P):
    """cF:928-988: L3 source, and its projection onto L1 and L2 via Shapes."""
    Qw, Nf, idx = _fine_quadrature_source(L3, v, p, P)
    Ff = bincount(idx, Nf @ Qw, ne_nn[4])
    d1 = (Shapes[1][0] * Qw.T[:, :, None]).sum(axis=1)
    d2 = (Shapes[2][0] * Qw.T[:, :, None]).sum(axis=1)
    Fc = bincount(Shapes[1][2], d1, Shapes[1][3])
    Fm = bincount(Shapes[2][2], d2, Shapes[2][3])
    return Fc, Fm, Ff


def computeLevelSource(Levels, ne_nn, laser_position, LevelShape, p, laserP):
    """cF:2667-2730: coarse source averaged over the given laser positions."""
    acc = None
    for i in range(laser_position.shape[0]):
        Qw, _, _ = _fine_quadrature_source(Levels[3], laser_position[i], p, laserP[i])
        d = (LevelShape[0] * Qw.T[:, :, None]).sum(axis=1)
        acc = d if acc is None else acc + d
    acc = acc / F(laser_position.shape[0])
    return bincount(LevelShape[2], acc, LevelShape[3])


def computeConvRadBC(lv, T, ne, nn, p, F_in):
    """cF:2207-2301: top-face convection + radiation + evaporation flux.

    Applied on the top layer of the first ``ne`` elements (the active ones).
    """
    cx, cy = lv["connect"][0], lv["connect"][1]
    ne_x, ne_y = cx.shape[0], cy.shape[0]
    coords = np.stack([lv["node_coords"][0][cx[0, :]], lv["node_coords"][1][cy[0, :]]], axis=1)
    N, _, wq = computeQuad2dFemShapeFunctions(coords)
    elems = np.arange(ne - ne_x * ne_y, ne)
    _, _, _, idx = convert2XYZ(elems, ne_x, ne_y, ne_x + 1, ne_y + 1)
    top = idx[4:]                                   # (4, n_top)
    Tq = N @ T[top]                                 # (4q, n_top)
    Tq = np.minimum(Tq, F(p["T_boiling"] + 1000))
    invT = F(1.0) / Tq
    E_pv = F(p["Lev"]) + F(p["cp_fluid"]) * (Tq - F(p["T_amb"]))
    MolMot = np.sqrt(F(p["CM_coeff"]) * invT)
    S = (F(p["evc"]) * F(p["CP_coeff"]) * np.exp(-F(p["CT_coeff"]) * (invT - F(1.0 / p["T_boiling"])))
         * MolMot * E_pv)
    q = (F(p["h_conv"]) * (F(p["T_amb"]) - Tq)
         + F(p["sigma_sb"]) * F(p["vareps"]) * (F(p["T_amb"]) ** 4 - Tq ** 4) - S)
    q = q * F(1e-6)
    contrib = N.T @ (q * wq)                        # (4a, n_top)
    return (F_in + bincount(top, contrib, nn)).astype(FLOAT)


# ---------------------------------------------------------------------------
# Dirichlet conditions  (cF:1568-1620, 1848-1883)
# ---------------------------------------------------------------------------

def assignBCs(RHS, Levels):
    """cF:1568-1595: L1 faces from JSON, order y0, y1, x0, x1, z0 (later wins)."""
    out = RHS.copy()
    bc, cond = Levels[1]["BC"], Levels[1]["conditions"]
    out[bc[2]] = cond["y"][0]
    out[bc[3]] = cond["y"][1]
    out[bc[0]] = cond["x"][0]
    out[bc[1]] = cond["x"][1]
    out[bc[4]] = cond["z"][0]
    return out


def assignBCsFine(RHS, TfAll, BC):
    """cF:1598-1620: W, E, S, N, bottom faces from the prolonged coarse field."""
    out = RHS.copy()
    for f in (2, 3, 0, 1, 4):
        out[BC[f]] = TfAll[BC[f]]
    return out


def substitute_Tbar(Tbar, idx, val):
    """cF:1848-1865."""
    out = Tbar.copy()
    out[idx:] = val
    return out


# ---------------------------------------------------------------------------
# Interpolation / prolongation  (cF:1028-1210, 1361-1393)
# ---------------------------------------------------------------------------

def compute3DN(q, x, y, z, h):
    """cF:1361-1393: trilinear shape functions of a box element, shape (8, ...)."""
    inv_vol = F(1.0) / (h[0] * h[1] * h[2])
    a = [x[1] - q[0], q[0] - x[0]]
    b = [y[1] - q[1], q[1] - y[0]]
    c = [z[1] - q[2], q[2] - z[0]]
    xb = [0, 1, 1, 0, 0, 1, 1, 0]
    yb = [0, 0, 1, 1, 0, 0, 1, 1]
    zb = [0, 0, 0, 0, 1, 1, 1, 1]
    return np.stack([a[xb[i]] * b[yb[i]] * c[zb[i]] * inv_vol for i in range(8)])


def interpolatePointsMatrix(lv, coords_new):
    """cF:1028-1107: weights (n, 8) and coarse node ids (n, 8) for a target grid.

    Element index is clipped into the mesh; if any weight leaves [-0.01, 1.01]
    all 8 weights become 0, otherwise they are clipped to [0, 1].
    """
    nc = lv["node_coords"]
    ne = [lv["connect"][i].shape[0] for i in range(3)]
    nnx, nny = ne[0] + 1, ne[1] + 1
    h = [nc[i][1] - nc[i][0] for i in range(3)]
    tx, ty, tz = [np.asarray(c, dtype=FLOAT) for c in coords_new]
    X = np.broadcast_to(tx[None, None, :], (tz.size, ty.size, tx.size)).reshape(-1)
    Y = np.broadcast_to(ty[None, :, None], (tz.size, ty.size, tx.size)).reshape(-1)
    Z = np.broadcast_to(tz[:, None, None], (tz.size, ty.size, tx.size)).reshape(-1)
    e = [np.clip(np.floor((q - nc[i][0]) / h[i]).astype(np.int64), 0, ne[i] - 1)
         for i, q in enumerate((X, Y, Z))]
    nodex = lv["connect"][0][e[0]]
    nodey = lv["connect"][1][e[1]]
    nodez = lv["connect"][2][e[2]]
    node = nodex + nodey * nnx + nodez * nnx * nny          # (n, 8)
    x0, x1 = nc[0][nodex[:, 0]], nc[0][nodex[:, 1]]
    y0, y1 = nc[1][nodey[:, 0]], nc[1][nodey[:, 3]]
    z0, z1 = nc[2][nodez[:, 0]], nc[2][nodez[:, 5]]
    Nc = compute3DN([X, Y, Z], [x0, x1], [y0, y1], [z0, z1], h).T   # (n, 8)
    valid = np.all(Nc >= -1e-2, axis=1) & np.all(Nc <= 1 + 1e-2, axis=1)
    Nc = np.where(valid[:, None], np.clip(Nc, 0.0, 1.0), 0.0).astype(FLOAT)
    return [Nc, node]


def interpolate_w_matrix(C2F, T):
    """cF:1110-1128."""
    return (C2F[0] * T[C2F[1]]).sum(axis=1).astype(FLOAT)


def interpolatePoints(lv, u, coords_new):
    """cF:1131-1210: interpolatePointsMatrix followed by the weighted gather."""
    return interpolate_w_matrix(interpolatePointsMatrix(lv, coords_new), u)


# ---------------------------------------------------------------------------
# Cross-level tabulation and restriction  (cF:1213-1358, 1396-1565,
# 2733-2810, 2857-2914, 3057-3132, 3169-3221)
# ---------------------------------------------------------------------------

def computeCoarseFineShapeFunctions(Coarse, Fine):
    """cF:1213-1358: coarse N and grad N at every fine Gauss point.

    Returns [Nc (nef, 8q, 8a), [dNcdx, dNcdy, dNcdz], nodes (nef, 8), nnc].
    ``nodes`` are the coarse nodes of the coarse element holding Gauss point 0
    of each fine element (the BCOO projection of the reference); each Gauss
    point's values are taken in its own (clipped) coarse element.
    """
    cc = Coarse["node_coords"]
    nec = [Coarse["connect"][i].shape[0] for i in range(3)]
    nnc = [cc[i].shape[0] for i in range(3)]
    hc = [cc[i][1] - cc[i][0] for i in range(3)]
    hxyz = hc[0] * hc[1] * hc[2]
    nef = [Fine["connect"][i].shape[0] for i in range(3)]
    Nf, _, _ = computeQuad3dFemShapeFunctions(getSampleCoords(Fine))
    ix, iy, iz, _ = convert2XYZ(np.arange(nef[0] * nef[1] * nef[2]), nef[0], nef[1],
                                nef[0] + 1, nef[1] + 1)
    x, y, z = getQuadratureCoords(Fine, ix, iy, iz, Nf)     # (8q, nef)
    x, y, z = x.T, y.T, z.T                                  # (nef, 8q)
    e = [np.clip(np.floor((q - cc[i][0]) / hc[i]).astype(np.int64), 0, nec[i] - 1)
         for i, q in enumerate((x, y, z))]
    xc0, xc1 = cc[0][e[0]], cc[0][e[0] + 1]
    yc0, yc3 = cc[1][e[1]], cc[1][e[1] + 1]
    zc0, zc5 = cc[2][e[2]], cc[2][e[2] + 1]
    Nc = np.moveaxis(compute3DN([x, y, z], [xc0, xc1], [yc0, yc3], [zc0, zc5], hc), 0, 2)
    sx = np.array([-1, 1, 1, -1, -1, 1, 1, -1], dtype=FLOAT)
    sy = np.array([-1, -1, 1, 1, -1, -1, 1, 1], dtype=FLOAT)
    sz = np.array([-1, -1, -1, -1, 1, 1, 1, 1], dtype=FLOAT)
    px = [(xc1 - x), (x - xc0)]
    py = [(yc3 - y), (y - yc0)]
    pz = [(zc5 - z), (z - zc0)]
    xb = [0, 1, 1, 0, 0, 1, 1, 0]
    yb = [0, 0, 1, 1, 0, 0, 1, 1]
    zb = [0, 0, 0, 0, 1, 1, 1, 1]
    dNcdx = np.stack([sx[a] * py[yb[a]] * pz[zb[a]] for a in range(8)], axis=2) / hxyz
    dNcdy = np.stack([px[xb[a]] * sy[a] * pz[zb[a]] for a in range(8)], axis=2) / hxyz
    dNcdz = np.stack([px[xb[a]] * py[yb[a]] * sz[a] for a in range(8)], axis=2) / hxyz
    cx = Coarse["connect"][0][e[0][:, 0]]
    cy = Coarse["connect"][1][e[1][:, 0]]
    cz = Coarse["connect"][2][e[2][:, 0]]
    nodes = cx + cy * nnc[0] + cz * nnc[0] * nnc[1]          # (nef, 8)
    return [Nc.astype(FLOAT), [dNcdx.astype(FLOAT), dNcdy.astype(FLOAT), dNcdz.astype(FLOAT)],
            nodes, nnc[0] * nnc[1] * nnc[2]]


def _fine_flux_terms(Fine, Tp, k):
    """k_mean_e * grad T' at the fine Gauss points, each (nef, 8q); plus wq."""
    Nf, dNdx, wq = computeQuad3dFemShapeFunctions(getSampleCoords(Fine))
    _, _, _, idx = _element_nodes(Fine, Fine["ne"])
    kmean = (Nf @ k[idx]).mean(axis=0)
    d = [(kmean * (dNdx[:, :, i] @ Tp[idx])).T for i in range(3)]
    return d, wq


def _fine_mass_term(Fine, dTp, rhocp):
    """rho*cp_mean_e * dT' at the fine Gauss points (nef, 8q); plus wq."""
    Nf, _, wq = computeQuad3dFemShapeFunctions(getSampleCoords(Fine))
    _, _, _, idx = _element_nodes(Fine, Fine["ne"])
    return ((Nf @ dTp[idx]) * (Nf @ rhocp[idx]).mean(axis=0)).T, wq


def restrictFlux(Shape, Fine, Tp, k):
    """-sum_e sum_q grad N_c . (k_e grad T') w_q, scattered to coarse nodes.

    cF:1513-1533 / 2776-2783 / 2897-2914 (one fine->coarse pair).
    """
    d, wq = _fine_flux_terms(Fine, Tp, k)
    acc = sum((-Shape[1][i] * d[i][:, :, None] * wq[None, :, :]).sum(axis=1) for i in range(3))
    return bincount(Shape[2], acc, Shape[3])


def restrictMass(Shape, Fine, dTp, rhocp, dt):
    """-(1/dt) sum_e sum_q N_c rho*cp_e dT' w_q, scattered to coarse nodes.

    cF:1440-1468 / 3106-3109 / 3211-3216 (one fine->coarse pair).
    """
    m, wq = _fine_mass_term(Fine, dTp, rhocp)
    acc = (-Shape[0] * m[:, :, None] * (F(1.0) / F(dt) * wq)[None, :, :]).sum(axis=1)
    return bincount(Shape[2], acc, Shape[3])


def computeCoarseTprimeTerm(Levels, L3k, L2k, Shapes):
    """cF:1477-1565: predictor flux terms (Vcu on L1, Vmu on L2)."""
    Vcu = (restrictFlux(Shapes[1], Levels[3], Levels[3]["Tprime0"], L3k)
           + restrictFlux(Shapes[0], Levels[2], Levels[2]["Tprime0"], L2k))
    Vmu = restrictFlux(Shapes[2], Levels[3], Levels[3]["Tprime0"], L3k)
    return Vcu, Vmu


def computeCoarseTprimeMassTerm(Levels, Tpf, Tpm, L3rc, L2rc, dt, Shapes, Vcu, Vmu):
    """cF:1396-1474."""
    dTf = Tpf - Levels[3]["Tprime0"]
    dTm = Tpm - Levels[2]["Tprime0"]
    Vcu = Vcu + (restrictMass(Shapes[1], Levels[3], dTf, L3rc, dt)
                 + restrictMass(Shapes[0], Levels[2], dTm, L2rc, dt))
    Vmu = Vmu + restrictMass(Shapes[2], Levels[3], dTf, L3rc, dt)
    return Vcu, Vmu


def computeL1TprimeTerms_Part1(Levels, ne_nn, L3k, Shapes, L2k):
    """cF:2733-2810."""
    return (restrictFlux(Shapes[1], Levels[3], Levels[3]["Tprime0"], L3k)
            + restrictFlux(Shapes[0], Levels[2], Levels[2]["Tprime0"], L2k))


def computeL1TprimeTerms_Part2(Levels, ne_nn, L3Tp, L2Tp, L3rc, L2rc, dt, Shapes, Vcu):
    """cF:3057-3132."""
    return Vcu + (restrictMass(Shapes[1], Levels[3], L3Tp - Levels[3]["Tprime0"], L3rc, dt)
                  + restrictMass(Shapes[0], Levels[2], L2Tp - Levels[2]["Tprime0"], L2rc, dt))


def computeL2TprimeTerms_Part1(Levels, ne_nn, L3Tp0, L3k, Shapes):
    """cF:2857-2914."""
    return restrictFlux(Shapes[2], Levels[3], L3Tp0, L3k)


def computeL2TprimeTerms_Part2(Levels, ne_nn, L3Tp, L3Tp0, L3rc, dt, Shapes, L2V):
    """cF:3169-3221."""
    return L2V + restrictMass(Shapes[2], Levels[3], L3Tp - L3Tp0, L3rc, dt)


def getNewTprime(Fine, FineT0, CoarseT, Coarse, C2F):
    """cF:2060-2099: inject the fine field into the coarse overlap nodes, then
    T' = fine - prolong(coarse)."""
    val = interpolatePoints(Fine, FineT0, Fine["overlapCoords"])
    idx = getOverlapRegion(Fine["overlapNodes"], Coarse["nodes"][0], Coarse["nodes"][1])
    CoarseT = CoarseT.copy()
    CoarseT[idx] = val
    return (FineT0 - interpolate_w_matrix(C2F, CoarseT)).astype(FLOAT), CoarseT


def getBothNewTprimes(Levels, FineT, MesoT, M2F, CoarseT, C2M):
    """cF:2102-2132."""
    lTp, mT0 = getNewTprime(Levels[3], FineT, MesoT, Levels[2], M2F)
    mTp, uT0 = getNewTprime(Levels[2], mT0, CoarseT, Levels[1], C2M)
    return lTp, mTp, mT0, uT0


# ---------------------------------------------------------------------------
# Level solves  (cF:2135-2204, 2813-2854, 2917-2957, 3015-3054)
# ---------------------------------------------------------------------------

def computeL1Temperature(Levels, ne_nn, tmp_ne_nn, L1F, L1V, L1k, L1rc, dt, p):
    """cF:2813-2854."""
    T = solveMatrixFreeFE(Levels[1], ne_nn[2], tmp_ne_nn[0], L1k, L1rc, dt,
                          Levels[1]["T0"], L1F, L1V)
    T = substitute_Tbar(T, tmp_ne_nn[1], p["T_amb"])
    return assignBCs(T, Levels)


def computeL2Temperature(L1T, L1L2Interp, Levels, ne_nn, L2T0, L2F, L2V, L2k, L2rc, dt):
    """cF:2917-2957."""
    TfAll = interpolate_w_matrix(L1L2Interp, L1T)
    T = solveMatrixFreeFE(Levels[2], ne_nn[3], ne_nn[0], L2k, L2rc, dt, L2T0, L2F, L2V)
    return assignBCsFine(T, TfAll, Levels[2]["BC"])


def computeSolutions_L3(FinalL2, L2L3Interp, Levels, ne_nn, L3T0, L3F, L3k, L3rc, dt):
    """cF:3015-3054."""
    TfAll = interpolate_w_matrix(L2L3Interp, FinalL2)
    T = solveMatrixFreeFE(Levels[3], ne_nn[4], ne_nn[1], L3k, L3rc, dt, L3T0, L3F, 0)
    return assignBCsFine(T, TfAll, Levels[3]["BC"])


def computeSolutions(Levels, ne_nn, tmp_ne_nn, LF, L1V, LInterp, Lk, Lrc, L2V, dt, p):
    """cF:2135-2204: L1 -> L2 -> L3 with BCs prolonged from the new coarse field."""
    L1T = solveMatrixFreeFE(Levels[1], ne_nn[2], tmp_ne_nn[0], Lk[1], Lrc[1], dt,
                            Levels[1]["T0"], LF[1], L1V)
    L1T = substitute_Tbar(L1T, tmp_ne_nn[1], p["T_amb"])
    FinalL1 = assignBCs(L1T, Levels)
    TfAll = interpolate_w_matrix(LInterp[0], FinalL1)
    L2T = solveMatrixFreeFE(Levels[2], ne_nn[3], ne_nn[0], Lk[2], Lrc[2], dt,
                            Levels[2]["T0"], LF[2], L2V)
    FinalL2 = assignBCsFine(L2T, TfAll, Levels[2]["BC"])
    TfAll = interpolate_w_matrix(LInterp[1], FinalL2)
    L3T = solveMatrixFreeFE(Levels[3], ne_nn[4], ne_nn[1], Lk[3], Lrc[3], dt,
                            Levels[3]["T0"], LF[3], 0)
    FinalL3 = assignBCsFine(L3T, TfAll, Levels[3]["BC"])
    return FinalL1, FinalL2, FinalL3


def _clamp(T, p):
    """max(T_amb, T), the reference's single-precision guard (cF:2359-2362)."""
    return np.maximum(F(p["T_amb"]), T).astype(FLOAT)


# ---------------------------------------------------------------------------
# Window move  (cF:1672-1845, 2400-2510)
# ---------------------------------------------------------------------------

def jit_constrain_v(vtot, lv):
    """cF:1672-1695."""
    b = lv["bounds"]
    return [np.clip(vtot[0], b["ix"][0], b["ix"][1]),
            np.clip(vtot[1], b["iy"][0], b["iy"][1]),
            np.clip(vtot[2], b["iz"][0], b["iz"][1])]


def _shift(v, h):
    """Window quantisation (v/h + 1e-2) truncated toward zero (cF:1716-1718)."""
    return int(np.trunc(F(v) / F(h) + F(1e-2)))


def move_fine_mesh(node_coords, h, v):
    """cF:1698-1724."""
    s = [_shift(v[i], h[i]) for i in range(3)]
    return [(node_coords[i] + F(h[i]) * s[i]).astype(FLOAT) for i in range(3)], s


def update_overlap_nodes_coords(lv, vcon, h, ratio):
    """cF:1727-1764."""
    s = [_shift(vcon[i], h[i]) for i in range(3)]
    lv["overlapNodes"] = [lv["orig_overlap_nodes"][i] + ratio[i] * s[i] for i in range(3)]
    lv["overlapCoords"] = [(lv["orig_overlap_coors"][i] + F(h[i]) * s[i]).astype(FLOAT)
                           for i in range(3)]
    return lv


def update_overlap_nodes_coords_L1L2(lv, vcon, h, layer):
    """cF:1767-1804."""
    s = [_shift(vcon[i], h[i]) for i in range(3)]
    sz_p = F(layer) * _shift(vcon[2], layer)
    lv["overlapNodes"] = [lv["orig_overlap_nodes"][i] + s[i] for i in range(3)]
    lv["overlapCoords"] = [(lv["orig_overlap_coors"][0] + F(h[0]) * s[0]).astype(FLOAT),
                           (lv["orig_overlap_coors"][1] + F(h[1]) * s[1]).astype(FLOAT),
                           (lv["orig_overlap_coors"][2] + sz_p).astype(FLOAT)]
    return lv


def update_overlap_nodes_coords_L2(lv, vcon, h, ratio):
    """cF:1807-1845."""
    s = [_shift(vcon[i], h[i]) for i in range(3)]
    lv["overlapNodes_L2"] = [lv["orig_overlap_nodes_L2"][i] + ratio[i] * s[i] for i in range(3)]
    lv["overlapCoords_L2"] = [(lv["orig_overlap_coors_L2"][i] + F(h[i]) * s[i]).astype(FLOAT)
                              for i in range(3)]
    return lv


def moveEverything(v, vstart, Levels, move_v, LInterp, L1L2Eratio, L2L3Eratio, height):
    """cF:2400-2510: move L3 then L2 with the laser and re-interpolate fields."""
    vtot = F(np.asarray(v, dtype=np.float64)[:3]) - F(np.asarray(vstart, dtype=np.float64)[:3])
    vtot = [vtot[i] for i in range(3)]
    L1, L2, L3, L0 = Levels[1], Levels[2], Levels[3], Levels[0]
    v3 = jit_constrain_v(vtot, L3)
    new3, _ = move_fine_mesh(L3["init_node_coors"], L2["h"], v3)
    update_overlap_nodes_coords(L3, v3, L2["h"], [1, 1, 1])
    T3 = interpolatePoints(L1, L1["T0"], new3)
    Tp2 = interpolatePoints(L2, L2["Tprime0"], new3)
    Tp3 = interpolatePoints(L3, L3["Tprime0"], new3)
    L3["T0"] = (T3 + (Tp2 + Tp3)).astype(FLOAT)
    L3["Tprime0"] = Tp3
    L3["node_coords"] = new3

    v2 = jit_constrain_v(vtot, L2)
    hL1 = [L1["h"][0], L1["h"][1], height]
    new2, mv = move_fine_mesh(L2["init_node_coors"], hL1, v2)
    move_v = [mv[i] * L1L2Eratio[i] for i in range(3)]
    update_overlap_nodes_coords_L1L2(L2, v2, L1["h"], height)
    T2 = interpolatePoints(L1, L1["T0"], new2)
    L2["Tprime0"] = interpolatePoints(L2, L2["Tprime0"], new2)
    L2["T0"] = (T2 + L2["Tprime0"]).astype(FLOAT)
    L2["node_coords"] = new2

    L2L1Shape = computeCoarseFineShapeFunctions(L1, L2)
    LInterp = list(LInterp)
    LInterp[0] = interpolatePointsMatrix(L1, new2)
    L3["overlapNodes"] = [L3["overlapNodes"][i] - move_v[i] for i in range(3)]

    update_overlap_nodes_coords(L0, v3, L2["h"], L2L3Eratio)
    update_overlap_nodes_coords_L2(
        L0, v2, [L1["h"][0], L1["h"][1], L2["h"][2]],
        [L1L2Eratio[0] * L2L3Eratio[0], L1L2Eratio[1] * L2L3Eratio[1], L2L3Eratio[2]])
    L0["overlapNodes"][2] = L0["overlapNodes"][2] - move_v[2] * L2L3Eratio[2]
    L0["overlapNodes_L2"][2] = L0["overlapNodes_L2"][2] - move_v[2] * L2L3Eratio[2]
    L0["idx"] = getOverlapRegion(L0["overlapNodes"], L0["nodes"][0], L0["nodes"][1])
    L0["idx_L2"] = getOverlapRegion(L0["overlapNodes_L2"], L0["nodes"][0], L0["nodes"][1])

    L2["S1"] = L0["S1"][L0["idx_L2"]].astype(FLOAT)
    L3["S1"] = L0["S1"][L0["idx"]].astype(FLOAT)
    L3["S2"] = L0["S2"][L0["idx"]].copy()

    L3L1Shape = computeCoarseFineShapeFunctions(L1, L3)
    L3L2Shape = computeCoarseFineShapeFunctions(L2, L3)
    LInterp[1] = interpolatePointsMatrix(L2, new3)
    return Levels, [L2L1Shape, L3L1Shape, L3L2Shape], LInterp, move_v


# ---------------------------------------------------------------------------
# Time integrators  (cF:2304-2397, 2617-2664, 3224-3632)
# ---------------------------------------------------------------------------

def stepGOMELT(Levels, ne_nn, tmp_ne_nn, Shapes, LInterp, v, p, dt, laserP, substrate):
    """cF:2304-2397: equal-dt predictor/corrector over all three levels."""
    preS2 = Levels[3]["S2"].copy()
    Levels, Lk, Lrc = updateStateProperties(Levels, p, substrate)
    L1, L2, L3 = Levels[1], Levels[2], Levels[3]
    Fc, Fm, Ff = computeSources(L3, v, Shapes, ne_nn, p, laserP)
    Fc = computeConvRadBC(L1, L1["T0"], tmp_ne_nn[0], ne_nn[2], p, Fc)
    Fm = computeConvRadBC(L2, L2["T0"], ne_nn[0], ne_nn[3], p, Fm)
    Ff = computeConvRadBC(L3, L3["T0"], ne_nn[1], ne_nn[4], p, Ff)
    Fs = [0, Fc, Fm, Ff]
    Vcu, Vmu = computeCoarseTprimeTerm(Levels, Lk[3], Lk[2], Shapes)
    L1T, L2T, L3T = computeSolutions(Levels, ne_nn, tmp_ne_nn, Fs, Vcu, LInterp, Lk, Lrc,
                                     Vmu, dt, p)
    L1T, L2T, L3T = _clamp(L1T, p), _clamp(L2T, p), _clamp(L3T, p)
    L3Tp, L2Tp, L2T, L1T = getBothNewTprimes(Levels, L3T, L2T, LInterp[1], L1T, LInterp[0])
    Vcu, Vmu = computeCoarseTprimeMassTerm(Levels, L3Tp, L2Tp, Lrc[3], Lrc[2], dt, Shapes,
                                           Vcu, Vmu)
    L1T, L2T, L3T = computeSolutions(Levels, ne_nn, tmp_ne_nn, Fs, Vcu, LInterp, Lk, Lrc,
                                     Vmu, dt, p)
    L1T, L2T = _clamp(L1T, p), _clamp(L2T, p)
    Levels[3]["T0"] = _clamp(L3T, p)
    (Levels[3]["Tprime0"], Levels[2]["Tprime0"], Levels[2]["T0"],
     Levels[1]["T0"]) = getBothNewTprimes(Levels, Levels[3]["T0"], L2T, LInterp[1], L1T,
                                          LInterp[0])
    # Level0 state (cF:2385-2392); updated in place (the reference's
    # functional .at[].set builds new arrays of the same values)
    L0 = Levels[0]
    L0["S1"][L0["idx"]] = Levels[3]["S1"]
    L0["S2"].fill(False)
    L0["S2"][L0["idx"]] = Levels[3]["S2"]
    resetmask = ((1 - 2 * preS2.astype(np.int64)) * Levels[3]["S2"]) == 1
    return Levels, resetmask


def stepGOMELTDwellTime(Levels, tmp_ne_nn, ne_nn, p, dt, substrate):
    """cF:2617-2664: L1-only step during dwell."""
    L1 = Levels[1]
    Fc = computeConvRadBC(L1, L1["T0"], tmp_ne_nn[0], ne_nn[2], p, 0)
    _, _, k, rc = computeStateProperties(L1["T0"], L1["S1"], p, substrate[1])
    T = solveMatrixFreeFE(L1, ne_nn[2], tmp_ne_nn[0], k, rc, dt, L1["T0"], Fc, 0)
    T = substitute_Tbar(T, tmp_ne_nn[1], p["T_amb"])
    Levels[1]["T0"] = assignBCs(T, Levels)
    return Levels


def subcycleGOMELT(Levels, ne_nn, Shapes, substrate, LInterp, tmp_ne_nn, laser_position,
                   p, laserP, subcycle, max_accum_L3, accum_L3):
    """cF:3224-3632: one L1 step with N2 L2 substeps of N3 L3 substeps each,
    predictor (Part1) then corrector (Part2) restarted from the start-of-step
    L2/L3 state."""
    N2, N3 = subcycle[0], subcycle[1]
    fN2, fN3 = subcycle[3], subcycle[4]
    lp = np.asarray(laser_position, dtype=np.float64)
    P = np.asarray(laserP, dtype=np.float64)
    L1, L2, L3 = Levels[1], Levels[2], Levels[3]
    dt1 = float(F(lp[:, 5]).sum())

    _, _, L3k_L1, L3rc_L1 = computeStateProperties(L3["T0"], L3["S1"], p, substrate[3])
    _, _, L2k_L1, L2rc_L1 = computeStateProperties(L2["T0"], L2["S1"], p, substrate[2])
    val = interpolatePoints(L2, L2["S1"], L2["overlapCoords"])
    idx = getOverlapRegion(L2["overlapNodes"], L1["nodes"][0], L1["nodes"][1])
    L1["S1"] = L1["S1"].copy()
    L1["S1"][idx] = val
    L1["S1"][:substrate[1]] = 1
    _, _, L1k, L1rc = computeStateProperties(L1["T0"], L1["S1"], p, substrate[1])

    L1F = computeLevelSource(Levels, ne_nn, lp, Shapes[1], p, P)
    L1F = computeConvRadBC(L1, L1["T0"], tmp_ne_nn[0], ne_nn[2], p, L1F)
    L1V = computeL1TprimeTerms_Part1(Levels, ne_nn, L3k_L1, Shapes, L2k_L1)
    L1T = _clamp(computeL1Temperature(Levels, ne_nn, tmp_ne_nn, L1F, L1V, L1k, L1rc, dt1, p), p)

    def l2_substep(s, c, part2, hist_tp=None):
        a2 = (s + 1) / fN2
        rows = s * N3 + np.arange(N3)
        dt2 = float(F(lp[rows, 5]).sum())
        L2T0c, L2S1c, L3Tc, L3Tpc, L3S1c = c["L2T"], c["L2S1"], c["L3T"], c["L3Tp"], c["L3S1"]
        _, _, L3k_L2, L3rc_L2 = computeStateProperties(L3Tc, L3S1c, p, substrate[3])
        L2S1n, _, L2k, L2rc = computeStateProperties(L2T0c, L2S1c, p, substrate[2])
        L2F = computeLevelSource(Levels, ne_nn, lp[rows], Shapes[2], p, P[rows])
        L2F = computeConvRadBC(L2, L2T0c, ne_nn[0], ne_nn[3], p, L2F)
        L2V = computeL2TprimeTerms_Part1(Levels, ne_nn, L3Tpc, L3k_L2, Shapes)
        if part2:
            L2V = computeL2TprimeTerms_Part2(Levels, ne_nn, hist_tp[s], L3Tpc, L3rc_L2, dt2,
                                             Shapes, L2V)
        BC = (F(a2) * L1T + F(1 - a2) * L1["T0"]).astype(FLOAT)
        L2Tn = _clamp(computeL2Temperature(BC, LInterp[0], Levels, ne_nn, L2T0c, L2F, L2V,
                                           L2k, L2rc, dt2), p)
        L3T, L3S1 = L3Tc, L3S1c
        L3S2, mx, acc = c.get("L3S2"), c.get("max"), c.get("acc")
        for j in range(N3):
            L3S1n, L3S2n, L3k, L3rc = computeStateProperties(L3T, L3S1, p, substrate[3])
            r = s * N3 + j
            L3F = computeSourcesL3(L3, lp[r], ne_nn, p, P[r])
            L3F = computeConvRadBC(L3, L3T, ne_nn[1], ne_nn[4], p, L3F)
            a3 = (j + 1) / fN3
            BC3 = (F(a3) * L2Tn + F(1 - a3) * L2T0c).astype(FLOAT)
            L3Tn = _clamp(computeSolutions_L3(BC3, LInterp[1], Levels, ne_nn, L3T, L3F, L3k,
                                              L3rc, lp[r, 5]), p)
            if part2:
                reset = ((1 - 2 * L3S2.astype(np.int64)) * L3S2n) == 1
                racc = acc * reset
                mx = np.maximum(racc, mx).astype(FLOAT)
                acc = (acc + F(lp[r, 5]) * L3S2n - racc).astype(FLOAT)
                L3S2 = L3S2n
            L3T, L3S1 = L3Tn, L3S1n
        L3Tp, L2Tn = getNewTprime(L3, L3T, L2Tn, L2, LInterp[1])
        out = {"L2T": L2Tn, "L2S1": L2S1n, "L3T": L3T, "L3Tp": L3Tp, "L3S1": L3S1}
        if part2:
            out.update({"L3S2": L3S2, "max": mx, "acc": acc})
        return out

    c = {"L2T": L2["T0"], "L2S1": L2["S1"], "L3T": L3["T0"], "L3Tp": L3["Tprime0"],
         "L3S1": L3["S1"]}
    hist_tp = []
    for s in range(N2):
        c = l2_substep(s, c, False)
        hist_tp.append(c["L3Tp"])
    L2Tp, L1T = getNewTprime(L2, c["L2T"], L1T, L1, LInterp[0])
    L1V = computeL1TprimeTerms_Part2(Levels, ne_nn, c["L3Tp"], L2Tp, L3rc_L1, L2rc_L1, dt1,
                                     Shapes, L1V)
    L1T = _clamp(computeL1Temperature(Levels, ne_nn, tmp_ne_nn, L1F, L1V, L1k, L1rc, dt1, p), p)

    c = {"L2T": L2["T0"], "L2S1": L2["S1"], "L3T": L3["T0"], "L3Tp": L3["Tprime0"],
         "L3S1": L3["S1"], "L3S2": L3["S2"], "max": np.asarray(max_accum_L3, dtype=FLOAT),
         "acc": np.asarray(accum_L3, dtype=FLOAT)}
    L2all, L3all, L3pall = [], [], []
    for s in range(N2):
        c = l2_substep(s, c, True, hist_tp)
        L2all.append(c["L2T"])
        L3all.append(c["L3T"])
        L3pall.append(c["L3Tp"])
    L2["T0"], L2["S1"] = c["L2T"], c["L2S1"]
    L3["T0"], L3["Tprime0"], L3["S1"], L3["S2"] = c["L3T"], c["L3Tp"], c["L3S1"], c["L3S2"]
    L2["Tprime0"], L1["T0"] = getNewTprime(L2, L2["T0"], L1T, L1, LInterp[0])
    # Level0 state (cF:3628-3630), updated in place
    L0 = Levels[0]
    L0["S1"][L0["idx"]] = L3["S1"]
    L0["S2"].fill(False)
    L0["S2"][L0["idx"]] = L3["S2"]
    return (Levels, np.stack(L2all), np.stack(L3all), np.stack(L3pall), c["max"], c["acc"])


def melting_temp(temps, dt, T_melt, accum_time, idx):
    """cF:3696-3712."""
    out = accum_time.copy()
    np.add.at(out, idx, (temps > T_melt) * F(dt))
    return out


def energy_balance(lv, ne, k, rhocp, dt, T, Fc, Corr, Tnew, bc_faces=0, T_clamp=-np.inf):
    """Energy audit of an explicit update T -> Tnew (FP64).

    No reference counterpart; it is the discrete identity of solveMatrixFreeFE
    (cF:582-642): with the lumped capacity C = sum_e Me*dt (element-loop
    scatter as in the solver) and the element-mean-k stiffness K, the
    unconstrained update Tt = T + dt (F - K T) / C satisfies
    sum C (Tt - T) = dt sum F, since K annihilates constants.  Nodes on the
    faces in ``bc_faces`` (bits: x0, x1, y0, y1, z0) or held at the floor
    (Tnew <= T_clamp while Tt < T_clamp) count with Tt, the others with Tnew.
    Returns the 8 sums of gm_energy_balance: free C dT, held C dTt, dt sum F,
    dt sum KT, held reaction, scale, residual, held count.
    """
    f64 = np.float64
    N, dNdx, wq = computeQuad3dFemShapeFunctions(getSampleCoords(lv))
    N, dNdx, wq = (np.asarray(a, f64) for a in (N, dNdx, wq))
    w = wq[0, 0]
    BTB = sum(dNdx[:, :, d].T @ dNdx[:, :, d] for d in range(3))
    Mrow = ((N.T @ N) * w).sum(axis=0)
    _, _, _, idx = _element_nodes(lv, ne)
    k, rhocp, T, Tnew = (np.asarray(a, f64) for a in (k, rhocp, T, Tnew))
    nn = len(T)
    kvec = k[idx].mean(axis=0)
    cvec = rhocp[idx].mean(axis=0)
    KTe = (T[idx].T @ (BTB * w).T) * kvec[:, None]
    KT = np.bincount(idx.T.reshape(-1), KTe.reshape(-1), minlength=nn)
    C = np.bincount(idx.T.reshape(-1), (cvec[:, None] * Mrow[None, :]).reshape(-1), minlength=nn)
    Fv = np.zeros(nn)
    for a in (Fc, Corr):
        if not np.isscalar(a):
            Fv = Fv + np.asarray(a, f64)
        elif a != 0:
            Fv = Fv + float(a)
    dt = float(dt)
    act = C > 0
    Tt = np.where(act, T + dt * (Fv - KT) / np.where(act, C, 1.0), T)
    nx, ny, nz = lv["nodes"]
    iz, iy, ix = np.unravel_index(np.arange(nn), (nz, ny, nx))
    face = np.zeros(nn, bool)
    for bit, m in enumerate((ix == 0, ix == nx - 1, iy == 0, iy == ny - 1, iz == 0)):
        if bc_faces >> bit & 1:
            face |= m
    held = act & (face | ((Tnew <= T_clamp) & (Tt < T_clamp)))
    free = act & ~held
    s0 = float((C * (Tnew - T))[free].sum())
    s1 = float((C * (Tt - T))[held].sum())
    s2 = float(dt * Fv[act].sum())
    s3 = float(dt * KT[act].sum())
    s4 = float((C * (Tnew - Tt))[held].sum())
    s5 = float(np.abs(dt * Fv[act]).sum() + np.abs(C * (Tnew - T))[free].sum()
               + np.abs(C * (Tt - T))[held].sum())
    s6 = (s0 + s1 - s2) / s5 if s5 > 0 else 0.0
    return np.array([s0, s1, s2, s3, s4, s5, s6, float(held.sum())])


def melt_pool_dims(T, axes, T_liq):
    """Melt-pool depth, width and length of a level's field (FP64).

    No reference counterpart: the reference only forms S2 = T >= T_liquidus
    (cF:2585).  The definition is the library's own: the molten set is the
    piecewise-linear interpolant of the nodal T along mesh edges at or above
    T_liq.  Depth is measured from the top node plane down to the lowest
    point of the molten set on any vertical grid line; width (length) is the
    y (x) extent of the molten set on the top plane.  Here each grid line is
    walked as a 1-D piecewise-linear function (a sorted list of the intervals
    where it is >= T_liq), which is then reduced; the GPU kernel instead
    evaluates edge crossings independently per node.
    Returns (depth, width, length, n_molten, ylo, yhi, xlo, xhi).
    """
    x, y, z = (np.asarray(a, np.float64) for a in axes)
    nx, ny, nz = len(x), len(y), len(z)
    A = np.asarray(T, np.float64).reshape(nz, ny, nx)
    Tl = float(T_liq)

    def molten_intervals(vals, coords):
        # maximal intervals [a, b] of coords where the linear interpolant is >= Tl
        out = []
        start = None
        for i, v in enumerate(vals):
            if v >= Tl:
                if start is None:
                    if i == 0:
                        start = coords[0]
                    else:
                        u = vals[i - 1]
                        start = coords[i - 1] + (coords[i] - coords[i - 1]) * (Tl - u) / (v - u)
            elif start is not None:
                u = vals[i - 1]
                out.append((start, coords[i - 1] + (coords[i] - coords[i - 1]) * (Tl - u) / (v - u)))
                start = None
        if start is not None:
            out.append((start, coords[-1]))
        return out

    ztop = z[-1]
    depth = 0.0
    for j in range(ny):
        for i in range(nx):
            col = A[:, j, i]
            if not (col >= Tl).any():
                continue
            # the column read from the top down is a function of depth below ztop
            for a, b in molten_intervals(col[::-1], ztop - z[::-1]):
                depth = max(depth, b)
    top = A[-1]
    ys = [iv for i in range(nx) for iv in molten_intervals(top[:, i], y)]
    xs = [iv for j in range(ny) for iv in molten_intervals(top[j, :], x)]
    ylo = min((a for a, _ in ys), default=np.nan)
    yhi = max((b for _, b in ys), default=np.nan)
    xlo = min((a for a, _ in xs), default=np.nan)
    xhi = max((b for _, b in xs), default=np.nan)
    width = (yhi - ylo) if ys else 0.0
    length = (xhi - xlo) if xs else 0.0
    return depth, width, length, int((A >= Tl).sum()), ylo, yhi, xlo, xhi


# ---------------------------------------------------------------------------
# Toolpath  (cP:6-201)
# ---------------------------------------------------------------------------

def format_fixed(val, width=15, precision=8):
    """cP:199-201."""
    return f"{val:.{precision}e}".rjust(width)


def parsingGcode(Nonmesh, Properties, L2h, gcode_text=None):
    """cP:6-188: G-code -> toolpath rows ``x,y,z,Ljump,Ldwell,dt,P``.

    Returns (row count, toolpath text).  The text is also written to
    ``Nonmesh['toolpath']`` when that key names a path.
    """
    dt3 = Nonmesh["timestep_L3"]
    coef = float(Nonmesh["dwell_time_multiplier"] * Nonmesh["subcycle_num_L2"]
                 * Nonmesh["subcycle_num_L3"])
    dwell_t = max(Nonmesh["dwell_time"] - Nonmesh["wait_time"] * dt3, Nonmesh["dwell_time"])
    if gcode_text is None:
        with open(Nonmesh["gcode"], "r") as fh:
            gcode_text = fh.read()
    pat = (r"(?:G(\d+)\s*X(-?\d+\.\d+|-?\d+)\s*Y(-?\d+\.\d+|-?\d+)"
           r"(?:\s*Z(-?\d+\.\d+|-?\d+))?)")
    pts = []
    cur_z = None
    for m in re.findall(pat, gcode_text):
        skip = 1.0 if m[0] != "1" else 0.0
        if m[3]:
            cur_z = float(m[3])
        pts.append((float(m[1]), float(m[2]), cur_z, skip))
    dx = Nonmesh["laser_velocity"] * dt3
    out = []
    n = 0

    def row(x, y, z, lj, ld, dt, P):
        out.append(f"{format_fixed(x)},{format_fixed(y)},{format_fixed(z)},"
                   f"{lj:d},{ld:d},{dt:.8e},{P:.8e}\n")

    def dwell_rows(x, y, z, lj):
        cnt = 0
        for k in range(int(Nonmesh["wait_time"])):
            if (k + 1) * dt3 > Nonmesh["dwell_time"]:
                break
            row(x, y, z, lj, 0, dt3, 0)
            cnt += 1
        if dwell_t > 0:
            rem = max(0, dwell_t - Nonmesh["wait_time"] * dt3)
            steps = int(rem / dt3 / coef)
            for _ in range(steps):
                row(x, y, z, lj, 0, dt3 * coef, 0)
                cnt += 1
            small = rem - steps * dt3 * coef
            if small > 0:
                row(x, y, z, lj, 0, small, 0)
                cnt += 1
        return cnt

    z = pts[0][2]
    nx_, ny_ = pts[0][0], pts[0][1]
    Ljump = 1
    for i in range(len(pts) - 1):
        Ljump, Ldwell = 1, 1
        if pts[i][2] != pts[i + 1][2]:
            nx_, ny_ = pts[i][0], pts[i][1]
            n += dwell_rows(nx_, ny_, z, Ljump)
            continue
        if pts[i + 1][3] == 1:
            Ljump = 0
        dist = float(np.linalg.norm(np.array([pts[i + 1][0] - pts[i][0],
                                              pts[i + 1][1] - pts[i][1]])))
        nseg = int(dist // dx)
        shortdt = (dist % dx) / dx * dt3
        nx_, ny_ = pts[i][0], pts[i][1]
        vx = Nonmesh["laser_velocity"] * (pts[i + 1][0] - pts[i][0]) / dist
        vy = Nonmesh["laser_velocity"] * (pts[i + 1][1] - pts[i][1]) / dist
        z = pts[i][2]
        for _ in range(nseg):
            nx_ += vx * dt3
            ny_ += vy * dt3
            row(nx_, ny_, z, Ljump, Ldwell, dt3, Ljump * Properties["laser_power"])
            n += 1
        if shortdt > 0:
            nx_ += vx * shortdt
            ny_ += vy * shortdt
            row(nx_, ny_, z, Ljump, Ldwell, shortdt, Ljump * Properties["laser_power"])
            n += 1
    # trailing dwell (cP:158-187): the per-row time test comes first here
    for k in range(int(Nonmesh["wait_time"])):
        if (k + 1) * dt3 > Nonmesh["dwell_time"]:
            break
        row(nx_, ny_, z, Ljump, 0, dt3, 0)
        n += 1
    rem = max(0, dwell_t - Nonmesh["wait_time"] * dt3)
    steps = int(rem / dt3 / coef)
    for _ in range(steps):
        row(nx_, ny_, z, Ljump, 0, dt3 * coef, 0)
        n += 1
    small = rem - steps * dt3 * coef
    if small > 0:
        row(nx_, ny_, z, Ljump, 0, small, 0)
        n += 1
    text = "".join(out)
    path = Nonmesh.get("toolpath")
    if path:
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        with open(path, "w") as fh:
            fh.write(text)
    return n, text


# ---------------------------------------------------------------------------
# Driver  (gm:16-530), without file output
# ---------------------------------------------------------------------------

def go_melt(solver_input, rows=None, max_rows=None, log=None, log_rows=False):
    """gm:16-530: the time loop (single-step / subcycle / dwell), no output files.

    ``rows`` may pass pre-parsed toolpath rows (n, 7); otherwise the G-code of
    ``solver_input['nonmesh']`` is parsed.  Returns (Levels, info dict).
    ``log(time_inc, Levels)`` is called after every chunk (and, with
    ``log_rows``, after every single-step row).
    Layer changes, dwell and the checkpoint-driven single-step rule follow
    the reference; checkpoint, VTK and npz files are not written.
    """
    p = SetupProperties(solver_input.get("properties", {}))
    Levels = SetupLevels(solver_input, p)
    nm = SetupNonmesh(solver_input.get("nonmesh", {}))
    ne_nn = getStaticNodesAndElements(Levels)
    sub = getStaticSubcycle(nm)
    r12 = [int(np.round(Levels[1]["h"][i] / Levels[2]["h"][i])) for i in range(2)] + \
          [int(np.round(p["layer_height"] / Levels[2]["h"][2]))]
    r23 = [int(np.round(Levels[2]["h"][i] / Levels[3]["h"][i])) for i in range(3)]
    if rows is None:
        _, text = parsingGcode(dict(nm, toolpath=None), p, Levels[2]["h"])
        rows = np.array([[float(v) for v in ln.split(",")] for ln in text.splitlines()])
    rows = np.asarray(rows, dtype=np.float64)
    if max_rows is not None:
        rows = rows[:max_rows]
    vstart = np.array(p["laser_center"], dtype=np.float64) if p["laser_center"] else rows[0]
    LInterp = [interpolatePointsMatrix(Levels[1], Levels[2]["node_coords"]),
               interpolatePointsMatrix(Levels[2], Levels[3]["node_coords"])]
    accum = np.zeros(Levels[0]["nn"], dtype=FLOAT)
    max_accum = np.zeros(Levels[0]["nn"], dtype=FLOAT)
    move_hist = [0, 0, 0]
    prev_z = float("inf")
    wait_inc = time_inc = 0
    load_chkpt = new_checkpoint = False
    layer_num = nm["layer_num"]
    layer_check = nm["layer_num"] + nm["restart_layer_num"]
    Shapes = tmp_ne_nn = substrate = None
    info = {"single": 0, "subcycle": 0, "dwell": 0, "chunks": []}
    N23 = sub[2]
    pos = 0
    while True:
        chunk = rows[pos:pos + N23]
        pos += chunk.shape[0]
        if chunk.shape[0] == 0:
            break
        ongoing = chunk.shape[0] == N23
        jump = False
        if chunk.shape[0] > 1:
            jump = bool((np.abs(np.diff(F(chunk), axis=0)[:, :2] / F(chunk[:-1, 5]).max())
                         > 100 * nm["laser_velocity"]).any())
        single = (bool(np.any(chunk[:, 2] != prev_z)) or load_chkpt
                  or wait_inc > max(0, nm["wait_time"] - sub[0] * sub[1] * 2)
                  or not ongoing or chunk.shape[0] == 1 or jump)
        info["chunks"].append("single" if single else "subcycle")
        if single:
            for lp in chunk:
                wait_inc = wait_inc + 1 if lp[4] == 0 else 0
                if lp[2] != prev_z and time_inc > 0 and not load_chkpt:
                    new_checkpoint = True
                move_vert = False
                if lp[2] != prev_z:
                    tmp = copy.deepcopy(Levels[1]["orig_node_coords"])
                    kidx = 0
                    while not np.isclose(tmp[2] - lp[2], 0, atol=1e-4).any():
                        tmp[2] = (tmp[2] + p["layer_height"]).astype(FLOAT)
                        kidx += 1
                    if not load_chkpt:
                        Levels[1]["T0"] = _clamp(
                            interpolatePoints(Levels[1], Levels[1]["T0"], tmp), p)
                        Levels[1]["S1_storage"][kidx - 1, :] = Levels[1]["S1"]
                        Levels[1]["S1"] = Levels[1]["S1_storage"][kidx, :].copy()
                        Levels[1]["node_coords"] = copy.deepcopy(tmp)
                    LInterp = [interpolatePointsMatrix(Levels[1], Levels[2]["node_coords"]),
                               interpolatePointsMatrix(Levels[2], Levels[3]["node_coords"])]
                    tmp_ne_nn = calcStaticTmpNodesAndElements(Levels, lp)
                    prev_z = lp[2]
                    wait_inc = 0
                    move_vert = True
                    if not load_chkpt:
                        L0 = Levels[0]
                        accum = np.maximum(accum, max_accum)
                        n1 = L0["nodes"][0] * L0["nodes"][1] * L0["layer_idx_delta"]
                        n2 = L0["nodes"][0] * L0["nodes"][1] * (L0["nodes"][2]
                                                                 - L0["layer_idx_delta"])
                        L0["S1"] = L0["S1"].copy()
                        L0["S1"][:n2] = L0["S1"][n1:].copy()
                        L0["S1"][n2:] = 0
                        L0["node_coords"][2] = (L0["orig_node_coords"][2] + lp[2]
                                                - L0["orig_node_coords"][2][-1]).astype(FLOAT)
                        max_accum = np.zeros(L0["nn"], dtype=FLOAT)
                        accum = accum.copy()
                        accum[:n2] = accum[n1:].copy()
                        accum[n2:] = 0
                Levels, Shapes, LInterp, move_hist = moveEverything(
                    lp, vstart, Levels, move_hist, LInterp, r12, r23, p["layer_height"])
                if move_vert:
                    substrate = getSubstrateNodes(Levels)
                    Levels[0]["S1"][:substrate[0]] = 1
                if wait_inc <= nm["wait_time"]:
                    Levels, reset = stepGOMELT(Levels, ne_nn, tmp_ne_nn, Shapes, LInterp, lp, p,
                                               lp[5], lp[6], substrate)
                    idx = Levels[0]["idx"]
                    racc = accum[idx] * (reset > 0)
                    max_accum[idx] = np.maximum(racc, max_accum[idx])
                    accum[idx] = accum[idx] - racc
                    # melting_temp (cF:3696-3712), in place on the driver's map
                    np.add.at(accum, idx,
                              (Levels[3]["T0"] > p["T_liquidus"]).astype(FLOAT) * F(lp[5]))
                    info["single"] += 1
                    if log_rows and log is not None:
                        log(time_inc + 1, Levels)
                else:
                    if (not (Levels[2]["Tprime0"] == 0).all()
                            and not (Levels[3]["Tprime0"] == 0).all()):
                        Levels[2]["Tprime0"] = np.zeros_like(Levels[2]["Tprime0"])
                        Levels[3]["Tprime0"] = np.zeros_like(Levels[3]["Tprime0"])
                    Levels = stepGOMELTDwellTime(Levels, tmp_ne_nn, ne_nn, p, lp[5], substrate)
                    info["dwell"] += 1
                time_inc += 1
            if new_checkpoint:
                layer_num += 1
                if layer_num == layer_check:
                    break
                new_checkpoint = False
                load_chkpt = True
            else:
                load_chkpt = False
        else:
            wait_inc = ((wait_inc + len(chunk) - chunk[:, 4].sum())
                        if (chunk[:, 4] == 0).any() else 0)
            Levels, Shapes, LInterp, move_hist = moveEverything(
                chunk[0], vstart, Levels, move_hist, LInterp, r12, r23, p["layer_height"])
            idx = Levels[0]["idx"]
            Levels, _, _, _, mx, ac = subcycleGOMELT(
                Levels, ne_nn, Shapes, substrate, LInterp, tmp_ne_nn, chunk, p, chunk[:, 6],
                sub, max_accum[idx], accum[idx])
            max_accum[idx] = mx
            accum[idx] = ac
            time_inc += chunk.shape[0]
            info["subcycle"] += 1
        if log is not None:
            log(time_inc, Levels)
        if not ongoing:
            break
    info["time_inc"] = time_inc
    info["accum"] = accum
    info["max_accum"] = max_accum
    return Levels, info


def load_input(path):
    with open(path) as fh:
        return json.load(fh)
