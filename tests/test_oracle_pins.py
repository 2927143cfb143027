"""Pins of the CPU oracle against closed forms, brute force and invariants.

The reference ships no golden outputs and cannot run here, so every oracle
function the GPU path is checked against is itself checked by something the
mathematics fixes (computeFunctions.py = cF):

* element operator (cF:582-642, 692-775): exact box-element stiffness and
  mass as tensor products of the 1-D linear element; dense assembly on a tiny
  mesh; patch test; constant invariance; the discrete Fourier symbol of the
  explicit lumped-mass update on cosine modes (exact eigenvectors of the
  natural-boundary problem on a uniform mesh);
* material state (cF:2567-2614): hand-evaluated phase and property values;
* laser source (cF:991-1025, 2960-3012): the Gaussian's half-space integral
  equals eta*P;
* surface flux (cF:2207-2301): uniform-temperature closed form per top node;
* interpolation (cF:1028-1210): exact on trilinear fields, zero outside;
* coarse/fine tabulation and restriction (cF:1213-1565): partition of unity,
  zero gradient sum, brute-force loops, total mass closed form;
* T' (cF:2060-2099): zero for a coarse-interpolant fine field;
* window shift (cF:1716-1718): truncation toward zero of v/h + 0.01.

Each test is built so that a plausible slip (a dropped term, a wrong sign, a
transposed axis, an off-by-one element) fails it.
"""

import math

import numpy as np
import pytest

import workloads as W


def _level(O, elements, bounds):
    """One level dict in the reference layout (cF:134-178)."""
    lv = {"elements": list(elements), "bounds": {k: list(v) for k, v in bounds.items()}}
    lv["length"], lv["h"] = O.calc_length_h(lv)
    lv["nodes"] = O.calcNumNodes(lv["elements"])
    lv["ne"] = int(np.prod(lv["elements"]))
    lv["nn"] = int(np.prod(lv["nodes"]))
    lv["BC"] = O.getBCindices(lv)
    b = lv["bounds"]
    lv["node_coords"], lv["connect"] = O.createMesh3D(
        (b["x"][0], b["x"][1], lv["nodes"][0]), (b["y"][0], b["y"][1], lv["nodes"][1]),
        (b["z"][0], b["z"][1], lv["nodes"][2]))
    return lv


def _xyz(lv):
    ax = lv["node_coords"]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return X.reshape(-1), Y.reshape(-1), Z.reshape(-1)


# local node order of the reference element: bits (x, y, z) of nodes 0..7
_BITS = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def _box_matrices(h):
    """Exact stiffness (unit k) and consistent mass of a box element."""
    K1 = [np.array([[1.0, -1.0], [-1.0, 1.0]]) / hd for hd in h]
    M1 = [np.array([[2.0, 1.0], [1.0, 2.0]]) * hd / 6.0 for hd in h]
    K = np.zeros((8, 8))
    M = np.zeros((8, 8))
    for a, (ax, ay, az) in enumerate(_BITS):
        for b, (bx, by, bz) in enumerate(_BITS):
            K[a, b] = (K1[0][ax, bx] * M1[1][ay, by] * M1[2][az, bz]
                       + M1[0][ax, bx] * K1[1][ay, by] * M1[2][az, bz]
                       + M1[0][ax, bx] * M1[1][ay, by] * K1[2][az, bz])
            M[a, b] = M1[0][ax, bx] * M1[1][ay, by] * M1[2][az, bz]
    return K, M


# ---------------------------------------------------------------------------
# element operator and the explicit update
# ---------------------------------------------------------------------------

def test_element_operator_is_the_exact_box_element(oracle):
    O = oracle
    lv = _level(O, (3, 2, 2), {"x": [0.0, 0.9], "y": [0.0, 0.4], "z": [-0.2, 0.0]})
    N, dNdx, wq = O.computeQuad3dFemShapeFunctions(O.getSampleCoords(lv))
    h = lv["h"]
    w = wq[:, 0]
    K = sum(dNdx[:, :, d].T @ (w[:, None] * dNdx[:, :, d]) for d in range(3))
    M = N.T @ (w[:, None] * N)
    Kx, Mx = _box_matrices(h)
    np.testing.assert_allclose(K, Kx, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(M, Mx, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(N.sum(axis=1), 1.0, rtol=1e-14)
    np.testing.assert_allclose(M.sum(axis=0), np.prod(h) / 8.0, rtol=1e-12)


def _dense_step(lv, k, rc, dt, T, Fc, Corr, ne=None):
    """Assembled K and lumped M with element-mean coefficients, one update."""
    nn = lv["nn"]
    ex, ey, ez = lv["elements"]
    nx, ny = lv["nodes"][0], lv["nodes"][1]
    Ke, Me = _box_matrices(lv["h"])
    K = np.zeros((nn, nn))
    m = np.zeros(nn)
    ne = ex * ey * ez if ne is None else ne
    for e in range(ne):
        i, j, l = e % ex, (e // ex) % ey, e // (ex * ey)
        ids = [(i + bx) + nx * (j + by) + nx * ny * (l + bz) for bx, by, bz in _BITS]
        kb = k[ids].mean()
        cb = rc[ids].mean()
        K[np.ix_(ids, ids)] += kb * Ke
        m[ids] += cb / dt * Me.sum(axis=1)
    return (m * T - K @ T + Fc + Corr) / m


def test_solve_matches_dense_assembly(oracle):
    O = oracle
    rng = np.random.default_rng(1)
    lv = _level(O, (4, 3, 3), {"x": [0.0, 0.2], "y": [0.0, 0.12], "z": [-0.09, 0.0]})
    nn = lv["nn"]
    k = rng.uniform(0.01, 0.05, nn)
    rc = rng.uniform(3e-3, 5e-3, nn)
    T = rng.uniform(300.0, 2000.0, nn)
    Fc = rng.normal(0.0, 1e-3, nn)
    Corr = rng.normal(0.0, 1e-3, nn)
    dt = 1e-5
    got = O.solveMatrixFreeFE(lv, nn, lv["ne"], k, rc, dt, T, Fc, Corr)
    ref = _dense_step(lv, k, rc, dt, T, Fc, Corr)
    np.testing.assert_allclose(got, ref, rtol=1e-12)


def test_solve_on_active_layers_only(oracle):
    """With ne below the full count only the lower element layers take part
    (the L1 active region, cF:2175); nodes above divide 0/0."""
    O = oracle
    rng = np.random.default_rng(2)
    lv = _level(O, (3, 3, 4), {"x": [0.0, 0.3], "y": [0.0, 0.3], "z": [-0.4, 0.0]})
    nn, ne = lv["nn"], 3 * 3 * 2
    k = rng.uniform(0.01, 0.05, nn)
    rc = rng.uniform(3e-3, 5e-3, nn)
    T = rng.uniform(300.0, 2000.0, nn)
    z = np.zeros(nn)
    got = O.solveMatrixFreeFE(lv, nn, ne, k, rc, 1e-4, T, z, z)
    nact = 16 * 3
    ref = _dense_step(lv, k, rc, 1e-4, T, z, z, ne=ne)
    np.testing.assert_allclose(got[:nact], ref[:nact], rtol=1e-12)
    assert np.all(np.isnan(got[nact:]))


def test_solve_patch_test_and_constant_invariance(oracle):
    O = oracle
    lv = _level(O, (5, 4, 3), {"x": [0.0, 0.5], "y": [0.0, 0.4], "z": [-0.3, 0.0]})
    nn = lv["nn"]
    X, Y, Z = _xyz(lv)
    k = np.full(nn, 0.03)
    rc = np.full(nn, 4e-3)
    z = np.zeros(nn)
    Tc = np.full(nn, 812.5)
    np.testing.assert_allclose(O.solveMatrixFreeFE(lv, nn, lv["ne"], k, rc, 1e-3, Tc, z, z),
                               Tc, rtol=1e-13)
    T = 400.0 + 300.0 * X - 200.0 * Y + 500.0 * Z
    got = O.solveMatrixFreeFE(lv, nn, lv["ne"], k, rc, 1e-3, T, z, z)
    nx, ny, nz = lv["nodes"]
    inner = np.zeros((nz, ny, nx), bool)
    inner[1:-1, 1:-1, 1:-1] = True
    np.testing.assert_allclose(got[inner.reshape(-1)], T[inner.reshape(-1)], rtol=1e-12)
    # a boundary node of the linear field is not at equilibrium (flux leaves)
    assert np.max(np.abs(got - T)) > 1.0


@pytest.mark.parametrize("modes", [(1, 0, 0), (0, 2, 0), (0, 0, 1), (2, 1, 3), (4, 3, 2)])
def test_solve_fourier_symbol(oracle, modes):
    """cos modes are eigenvectors: T+ = g(theta) T with
    g = 1 - (dt k / rho c) sum_d (2 - 2 cos th_d)/h_d^2 prod_{j != d} (2 + cos th_j)/3."""
    O = oracle
    el = (6, 5, 4)
    lv = _level(O, el, {"x": [0.0, 0.3], "y": [0.0, 0.2], "z": [-0.12, 0.0]})
    nn = lv["nn"]
    h = lv["h"]
    th = [math.pi * modes[d] / el[d] for d in range(3)]
    ax = [np.cos(th[d] * np.arange(el[d] + 1)) for d in range(3)]
    T = (ax[2][:, None, None] * ax[1][None, :, None] * ax[0][None, None, :]).reshape(-1)
    kk, rc, dt = 0.02, 4e-3, 2e-5
    z = np.zeros(nn)
    got = O.solveMatrixFreeFE(lv, nn, lv["ne"], np.full(nn, kk), np.full(nn, rc), dt, T, z, z)
    s = 0.0
    for d in range(3):
        p = 1.0
        for j in range(3):
            if j != d:
                p *= (2.0 + math.cos(th[j])) / 3.0
        s += (2.0 - 2.0 * math.cos(th[d])) / h[d] ** 2 * p
    g = 1.0 - dt * kk / rc * s
    np.testing.assert_allclose(got, g * T, atol=1e-12)


# ---------------------------------------------------------------------------
# material state
# ---------------------------------------------------------------------------

def test_state_properties_closed_form(oracle):
    O = oracle
    p = O.SetupProperties(W._props(
        thermal_conductivity_powder=5.0, thermal_conductivity_bulk_a0=10.0,
        thermal_conductivity_bulk_a1=0.01, thermal_conductivity_fluid_a0=30.0,
        heat_capacity_solid_a0=400.0, heat_capacity_solid_a1=0.1, heat_capacity_mushy=900.0,
        heat_capacity_fluid=700.0, density=8e-6))
    T = np.array([300.0, 1000.0, 1000.0, 1550.0, 1609.0, 2500.0, 400.0])
    S1 = np.array([0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0])
    S1o, S2, k, rc = O.computeStateProperties(T, S1, p, 1)
    # node 0 is substrate (n_sub = 1)
    np.testing.assert_array_equal(S1o, [1, 0, 1, 1, 1, 1, 0])
    np.testing.assert_array_equal(S2, [0, 0, 0, 0, 1, 1, 0])
    k_exp = np.array([10 + 3.0, 5.0, 10 + 10.0, 10 + 15.5, 30.0, 30.0, 5.0]) / 1000.0
    cp_exp = np.array([430.0, 500.0, 500.0, 900.0, 700.0, 700.0, 440.0])
    np.testing.assert_allclose(k, k_exp, rtol=1e-14)
    np.testing.assert_allclose(rc, 8e-6 * cp_exp, rtol=1e-14)


def test_property_table_reduces_to_linear_model(oracle):
    """A 2-knot table is the reference's linear bulk law (cF:2595-2612);
    beyond the end knots the table extends linearly."""
    O = oracle
    base = W._props(thermal_conductivity_bulk_a0=10.0, thermal_conductivity_bulk_a1=0.01,
                    heat_capacity_solid_a0=400.0, heat_capacity_solid_a1=0.1)
    p_lin = O.SetupProperties(base)
    p_tab = O.SetupProperties(dict(base,
                                   thermal_conductivity_bulk_table=[[300.0, 3000.0], [13.0, 40.0]],
                                   heat_capacity_solid_table=[[300.0, 3000.0], [430.0, 700.0]]))
    assert p_tab["k_table"] is not None and p_tab["cp_table"] is not None
    T = np.array([200.0, 300.0, 777.7, 1520.0, 3500.0])
    S1 = np.ones(5)
    a = O.computeStateProperties(T, S1, p_lin, 0)
    b = O.computeStateProperties(T, S1, p_tab, 0)
    for x, y in zip(a, b):
        np.testing.assert_allclose(np.asarray(x, float), np.asarray(y, float), rtol=1e-13)
    # an interior knot: value at the knot and linear on each side
    p3 = O.SetupProperties(dict(base, thermal_conductivity_bulk_table=[
        [300.0, 1000.0, 2000.0], [10.0, 20.0, 25.0]]))
    _, _, k, _ = O.computeStateProperties(np.array([650.0, 1000.0, 1500.0, 230.0]),
                                          np.ones(4), p3, 0)
    np.testing.assert_allclose(k * 1000.0, [15.0, 20.0, 22.5, 9.0], rtol=1e-13)


# ---------------------------------------------------------------------------
# laser source and surface flux
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("v", [(0.0, 0.0), (0.013, -0.021)])
def test_source_half_space_integral(oracle, v):
    """The Gaussian is normalised so that its integral over z < v_z is eta*P
    (cF:1018-1023); nodal loads sum to the integral (sum_a N_a = 1)."""
    O = oracle
    p = O.SetupProperties(W._props(laser_radius=0.05, laser_depth=0.04,
                                   laser_absorptivity=0.35))
    lv = _level(O, (40, 40, 20), {"x": [-0.3, 0.3], "y": [-0.3, 0.3], "z": [-0.3, 0.0]})
    P = 200.0
    Fv = O.computeSourcesL3(lv, [v[0], v[1], 0.0], (0, lv["ne"], 0, 0, lv["nn"]), p, P)
    np.testing.assert_allclose(Fv.sum(), 0.35 * P, rtol=2e-3)
    assert Fv.min() >= 0.0


def test_convrad_uniform_temperature(oracle):
    """Uniform T: an interior top node receives q(T) hx hy; an edge node half,
    a corner a quarter; nodes below the top receive nothing (cF:2254-2299)."""
    O = oracle
    lv = _level(O, (4, 3, 2), {"x": [0.0, 0.4], "y": [0.0, 0.6], "z": [-0.2, 0.0]})
    nn = lv["nn"]
    hx, hy = lv["h"][0], lv["h"][1]
    Tu = 1200.0
    T = np.full(nn, Tu)
    # convection only: h (T_amb - T) with h in W/mm^2 K (the 1e6 and 1e-6 of
    # cF:319 and cF:2290 cancel)
    p = O.SetupProperties(W._props(h_conv=2e-5, emissivity=0.0, evaporation_coefficient=0.0))
    Fv = O.computeConvRadBC(lv, T, lv["ne"], nn, p, np.zeros(nn))
    q = 2e-5 * (300.0 - Tu)
    nx, ny = lv["nodes"][0], lv["nodes"][1]
    top = Fv.reshape(3, ny, nx)[-1]
    np.testing.assert_allclose(top[1, 1], q * hx * hy, rtol=1e-12)
    np.testing.assert_allclose(top[0, 1], q * hx * hy / 2, rtol=1e-12)
    np.testing.assert_allclose(top[0, 0], q * hx * hy / 4, rtol=1e-12)
    assert np.all(Fv.reshape(3, ny, nx)[:-1] == 0.0)
    # radiation only: sigma eps (T_amb^4 - T^4), sigma in W/mm^2 K^4
    p = O.SetupProperties(W._props(h_conv=0.0, emissivity=0.4, evaporation_coefficient=0.0))
    Fv = O.computeConvRadBC(lv, T, lv["ne"], nn, p, np.zeros(nn))
    sig = 5.67e-8 * 1e-6
    q = sig * 0.4 * (300.0 ** 4 - Tu ** 4)
    np.testing.assert_allclose(Fv.reshape(3, ny, nx)[-1][1, 2], q * hx * hy, rtol=2e-3)


# ---------------------------------------------------------------------------
# interpolation
# ---------------------------------------------------------------------------

def _trilinear(X, Y, Z):
    return 3.0 + 2.0 * X - 5.0 * Y + 7.0 * Z + 1.5 * X * Y - 2.5 * Y * Z + 0.5 * X * Z + 4.0 * X * Y * Z


def test_interpolation_reproduces_trilinear_fields(oracle):
    O = oracle
    lv = _level(O, (5, 4, 3), {"x": [0.0, 1.0], "y": [0.0, 0.8], "z": [-0.6, 0.0]})
    X, Y, Z = _xyz(lv)
    u = _trilinear(X, Y, Z)
    tx = np.linspace(0.013, 0.987, 11)
    ty = np.linspace(0.0, 0.8, 7)          # includes coincident nodes and the ends
    tz = np.linspace(-0.59, -0.01, 5)
    got = O.interpolatePoints(lv, u, [tx, ty, tz])
    Zt, Yt, Xt = np.meshgrid(tz, ty, tx, indexing="ij")
    np.testing.assert_allclose(got, _trilinear(Xt, Yt, Zt).reshape(-1), rtol=1e-12, atol=1e-12)
    # coincident nodes: identity
    got = O.interpolatePoints(lv, u, lv["node_coords"])
    np.testing.assert_allclose(got, u, rtol=1e-12, atol=1e-12)


def test_interpolation_outside_is_zero(oracle):
    """Points more than 1% of an element outside get all-zero weights
    (cF:1094-1102); within 1% the weights are clipped to [0, 1]."""
    O = oracle
    lv = _level(O, (4, 4, 4), {"x": [0.0, 0.4], "y": [0.0, 0.4], "z": [-0.4, 0.0]})
    u = np.full(lv["nn"], 5.0)
    got = O.interpolatePoints(lv, u, [np.array([-0.05, -0.0005, 0.2, 0.4005, 0.45]),
                                      np.array([0.2]), np.array([-0.2])])
    np.testing.assert_allclose(got, [0.0, 5.0, 5.0, 5.0, 0.0], rtol=1e-12)


# ---------------------------------------------------------------------------
# coarse/fine tabulation, restriction and T'
# ---------------------------------------------------------------------------

def _pair(O, ratio=(2, 2, 2), offset=(1, 1, 1)):
    C = _level(O, (6, 6, 4), {"x": [0.0, 0.6], "y": [0.0, 0.6], "z": [-0.4, 0.0]})
    hc = C["h"]
    lo = [offset[d] * hc[d] + C["bounds"][a][0] for d, a in enumerate("xyz")]
    ncov = (3, 2, 2)
    Fb = {a: [lo[d], lo[d] + ncov[d] * hc[d]] for d, a in enumerate("xyz")}
    Fi = _level(O, tuple(ncov[d] * ratio[d] for d in range(3)), Fb)
    return C, Fi


def test_shapes_partition_of_unity(oracle):
    O = oracle
    C, Fi = _pair(O, ratio=(3, 2, 4))
    S = O.computeCoarseFineShapeFunctions(C, Fi)
    np.testing.assert_allclose(S[0].sum(axis=2), 1.0, rtol=1e-12)
    for d in range(3):
        np.testing.assert_allclose(S[1][d].sum(axis=2), 0.0, atol=1e-9)
    # gradient reproduction: sum_a dN_a/dx_d * f(node a) = df/dx_d for linear f
    Xc, Yc, Zc = _xyz(C)
    f = 2.0 * Xc - 3.0 * Yc + 5.0 * Zc
    vals = f[S[2]]                           # (nef, 8) at each element's coarse nodes
    for d, g in enumerate((2.0, -3.0, 5.0)):
        np.testing.assert_allclose((S[1][d][:, 0, :] * vals).sum(axis=1), g, rtol=1e-10)


def _brute_restrict(O, C, Fi, Tp, k, dTp, rc, dt):
    """Triple loop over fine elements and Gauss points in FP64; coarse
    values in the Gauss point's own clipped coarse element, added at the
    nodes of the coarse element holding Gauss point 0 (cF:1292-1357)."""
    g = 1.0 / math.sqrt(3.0)
    hf, hc = Fi["h"], C["h"]
    ef, ec = Fi["elements"], C["elements"]
    nfx, nfy = Fi["nodes"][0], Fi["nodes"][1]
    ncx, ncy = C["nodes"][0], C["nodes"][1]
    Vf = np.zeros(C["nn"])
    Vm = np.zeros(C["nn"])
    w = hf[0] * hf[1] * hf[2] / 8.0
    for e in range(int(np.prod(ef))):
        i, j, l = e % ef[0], (e // ef[0]) % ef[1], e // (ef[0] * ef[1])
        ids = [(i + bx) + nfx * (j + by) + nfx * nfy * (l + bz) for bx, by, bz in _BITS]
        x0 = [Fi["node_coords"][0][i], Fi["node_coords"][1][j], Fi["node_coords"][2][l]]
        kbar = k[ids].mean()
        cbar = rc[ids].mean()
        tgt = None
        for q, (qx, qy, qz) in enumerate(_BITS):
            # Gauss point q (reference order) in local coordinates (-1/+1 -> -g/+g)
            s = [(-g if b == 0 else g) for b in (qx, qy, qz)]
            xq = [x0[d] + (1 + s[d]) * hf[d] / 2 for d in range(3)]
            # fine trilinear basis and gradients at the point
            Nf = np.array([np.prod([(1 + (2 * bb - 1) * s[d]) / 2 for d, bb in enumerate(bits)])
                           for bits in _BITS])
            dNf = np.array([[(2 * bits[d] - 1) / hf[d] * np.prod(
                [(1 + (2 * bits[dd] - 1) * s[dd]) / 2 for dd in range(3) if dd != d])
                for d in range(3)] for bits in _BITS])
            gT = dNf.T @ Tp[ids]
            dT = Nf @ dTp[ids]
            ce = [min(max(int(math.floor((xq[d] - C["bounds"]["xyz"[d]][0]) / hc[d])), 0),
                      ec[d] - 1) for d in range(3)]
            t = [(xq[d] - (C["bounds"]["xyz"[d]][0] + ce[d] * hc[d])) / hc[d] for d in range(3)]
            if q == 0:
                tgt = [(ce[0] + bx) + ncx * (ce[1] + by) + ncx * ncy * (ce[2] + bz)
                       for bx, by, bz in _BITS]
            for a, bits in enumerate(_BITS):
                phi = [t[d] if bits[d] else 1 - t[d] for d in range(3)]
                dphi = [(1 if bits[d] else -1) / hc[d] for d in range(3)]
                Nc = phi[0] * phi[1] * phi[2]
                gNc = [dphi[0] * phi[1] * phi[2], phi[0] * dphi[1] * phi[2],
                       phi[0] * phi[1] * dphi[2]]
                Vf[tgt[a]] -= sum(gNc[d] * kbar * gT[d] for d in range(3)) * w
                Vm[tgt[a]] -= Nc * cbar * dT * w / dt
    return Vf, Vm


@pytest.mark.parametrize("ratio,offset", [((2, 2, 2), (1, 1, 1)), ((3, 2, 4), (2, 1, 0))])
def test_restriction_bruteforce(oracle, ratio, offset):
    O = oracle
    rng = np.random.default_rng(3)
    C, Fi = _pair(O, ratio, offset)
    nf = Fi["nn"]
    Tp = rng.normal(0.0, 50.0, nf)
    k = rng.uniform(0.01, 0.05, nf)
    dTp = rng.normal(0.0, 5.0, nf)
    rc = rng.uniform(3e-3, 5e-3, nf)
    S = O.computeCoarseFineShapeFunctions(C, Fi)
    Vf = O.restrictFlux(S, Fi, Tp, k)
    Vm = O.restrictMass(S, Fi, dTp, rc, 1e-5)
    bf, bm = _brute_restrict(O, C, Fi, Tp, k, dTp, rc, 1e-5)
    np.testing.assert_allclose(Vf, bf, rtol=1e-10, atol=1e-12 * np.abs(bf).max())
    np.testing.assert_allclose(Vm, bm, rtol=1e-10, atol=1e-12 * np.abs(bm).max())


def test_restriction_totals(oracle):
    """sum_A of the flux term is 0 (sum_A grad N_A = 0); the mass term of a
    uniform dT' totals -rho c dT' V_fine / dt (partition of unity)."""
    O = oracle
    C, Fi = _pair(O, (3, 2, 4), (1, 2, 1))
    nf = Fi["nn"]
    rng = np.random.default_rng(4)
    S = O.computeCoarseFineShapeFunctions(C, Fi)
    Vf = O.restrictFlux(S, Fi, rng.normal(0, 10, nf), rng.uniform(0.01, 0.05, nf))
    assert abs(Vf.sum()) < 1e-9 * np.abs(Vf).max()
    Vm = O.restrictMass(S, Fi, np.full(nf, 2.5), np.full(nf, 4e-3), 1e-5)
    vol = np.prod(Fi["length"])
    np.testing.assert_allclose(Vm.sum(), -4e-3 * 2.5 * vol / 1e-5, rtol=1e-12)


def test_new_tprime_zero_for_coarse_interpolant(oracle):
    """cF:2060-2099: a fine field that is the coarse interpolant of a trilinear
    coarse field gives T' = 0 and leaves the coarse overlap values unchanged."""
    O = oracle
    C, Fi = _pair(O, (2, 3, 2), (1, 1, 1))
    Xc, Yc, Zc = _xyz(C)
    Tc = _trilinear(Xc, Yc, Zc)
    Xf, Yf, Zf = _xyz(Fi)
    Tf = _trilinear(Xf, Yf, Zf)
    nodes = [O.getCoarseNodesInFineRegion(Fi["node_coords"][i], C["node_coords"][i])
             for i in range(3)]
    Fi["overlapNodes"] = nodes
    Fi["overlapCoords"] = [C["node_coords"][i][nodes[i]] for i in range(3)]
    C2F = O.interpolatePointsMatrix(C, Fi["node_coords"])
    Tp, Tc2 = O.getNewTprime(Fi, Tf, Tc, C, C2F)
    np.testing.assert_allclose(Tp, 0.0, atol=1e-10)
    np.testing.assert_allclose(Tc2, Tc, rtol=1e-12)
    # a fine-scale bump survives in T' and is injected into the coarse overlap
    Tf2 = Tf.copy()
    Tf2[Fi["nn"] // 2] += 100.0
    Tp, Tc2 = O.getNewTprime(Fi, Tf2, Tc, C, C2F)
    assert Tp.max() > 0.0


# ---------------------------------------------------------------------------
# window quantisation
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("v,h,s", [
    (0.0, 0.1, 0), (0.0995, 0.1, 1), (0.099, 0.1, 1), (0.0989, 0.1, 0), (0.25, 0.1, 2),
    (-0.05, 0.1, 0), (-0.1, 0.1, 0), (-0.1011, 0.1, -1), (-0.2, 0.1, -1), (-0.2101, 0.1, -2)])
def test_window_shift_truncates_toward_zero(oracle, v, h, s):
    """(v/h + 0.01) cast to int (cF:1716-1718): -1.0 h shifts by 0, not -1."""
    assert oracle._shift(v, h) == s


# ---------------------------------------------------------------------------
# melt-pool size (the library's own measure; the reference forms only
# S2 = T >= T_liquidus, cF:2585)
# ---------------------------------------------------------------------------

def _pool_mesh(n=(9, 11, 7), h=(0.1, 0.05, 0.025), x0=(0.3, -0.2, -0.15)):
    return [x0[d] + h[d] * np.arange(n[d]) for d in range(3)]


def _field(ax, f):
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return f(X, Y, Z).reshape(-1)


def test_melt_pool_linear_depth_closed_form(oracle):
    """T linear in z: the crossing, and so the depth, is exact off the nodes."""
    ax = _pool_mesh()
    ztop = ax[2][-1]
    Tl, T0, c = 1600.0, 2000.0, 4000.0  # depth (T0 - Tl) / c = 0.1 (4 elements)
    T = _field(ax, lambda X, Y, Z: T0 + c * (Z - ztop))
    d, w, l, n, ylo, yhi, xlo, xhi = oracle.melt_pool_dims(T, ax, Tl)
    assert d == pytest.approx(0.1, abs=1e-12)
    # the whole top plane is molten: extents reach the mesh boundary
    assert (ylo, yhi, xlo, xhi) == (ax[1][0], ax[1][-1], ax[0][0], ax[0][-1])
    assert w == pytest.approx(ax[1][-1] - ax[1][0]) and l == pytest.approx(ax[0][-1] - ax[0][0])
    assert n == 9 * 11 * 5  # planes within 0.1 of the top: 5 of 7
    # an off-node depth: 0.0625 = 2.5 elements
    T = _field(ax, lambda X, Y, Z: Tl + 400.0 - 6400.0 * (ztop - Z))
    assert oracle.melt_pool_dims(T, ax, Tl)[0] == pytest.approx(0.0625, abs=1e-12)


def test_melt_pool_tent_width_and_depth_closed_form(oracle):
    """A tent in y (kink on a node) and linear in z and x: width, length and
    depth of the linear interpolant are exact."""
    ax = _pool_mesh()
    ztop = ax[2][-1]
    yc, xc = ax[1][5], ax[0][4]
    Tl, Tpk = 1600.0, 1900.0
    T = _field(ax, lambda X, Y, Z: Tpk - 2000.0 * np.abs(Y - yc) - 1000.0 * np.abs(X - xc)
               - 6000.0 * (ztop - Z))
    d, w, l, n, ylo, yhi, xlo, xhi = oracle.melt_pool_dims(T, ax, Tl)
    assert w == pytest.approx(2 * 300.0 / 2000.0, abs=1e-12)
    assert ylo == pytest.approx(yc - 0.15, abs=1e-12) and yhi == pytest.approx(yc + 0.15, abs=1e-12)
    assert l == pytest.approx(2 * 300.0 / 1000.0, abs=1e-12)
    assert xlo == pytest.approx(xc - 0.3, abs=1e-12)
    assert d == pytest.approx(300.0 / 6000.0, abs=1e-12)


def test_melt_pool_rosenthal_field(oracle):
    """The measure on a realistic pool: Rosenthal's field sampled on a 5 um
    grid (source a third of a spacing off the nodes), against the isotherm's
    depth, width and length from root finding.  The linear interpolation
    along grid lines is second-order accurate: depth and width agree to
    1e-4 of the pool size here (measured 5e-5), the length to 1e-3 (4.6e-4;
    its ends are where T varies fastest along x)."""
    from rosenthal import ROSENTHAL, rosenthal_pool
    T, depth, width, length = rosenthal_pool()
    h = 0.005
    x = np.arange(-0.5, 0.05 + 1e-12, h) + h / 3
    y = np.arange(-0.15, 0.15 + 1e-12, h)
    z = np.arange(-0.15, 1e-12, h)
    F = np.minimum(_field((x, y, z), lambda X, Y, Z: T(X, np.sqrt(Y * Y + Z * Z))), 1e6)
    d, w, l = oracle.melt_pool_dims(F, (x, y, z), ROSENTHAL["Tl"])[:3]
    assert d == pytest.approx(depth, rel=1e-4)
    assert w == pytest.approx(width, rel=1e-4)
    assert l == pytest.approx(length, rel=1e-3)


def test_melt_pool_special_cases(oracle):
    ax = _pool_mesh()
    ztop = ax[2][-1]
    Tl = 1600.0
    # nothing molten
    d, w, l, n, ylo, yhi, xlo, xhi = oracle.melt_pool_dims(np.full(9 * 11 * 7, 300.0), ax, Tl)
    assert (d, w, l, n) == (0.0, 0.0, 0.0, 0) and np.isnan([ylo, yhi, xlo, xhi]).all()
    # molten to the bottom: the full mesh depth
    d = oracle.melt_pool_dims(np.full(9 * 11 * 7, 2000.0), ax, Tl)[0]
    assert d == pytest.approx(ztop - ax[2][0], abs=1e-12)
    # a buried hot spot: depth counts it, the top plane has no width
    T = np.full((7, 11, 9), 300.0)
    T[2, 4, 3] = 2000.0
    d, w, l, n = oracle.melt_pool_dims(T.reshape(-1), ax, Tl)[:4]
    assert w == 0.0 and l == 0.0 and n == 1
    # T_liquidus is crossed on the edge to the node below, 400/1700 of h
    # under the hot node (2000 K above, 300 K below)
    assert d == pytest.approx(ztop - ax[2][2] + 0.025 * 400.0 / 1700.0, abs=1e-12)
    # exactly T_liquidus counts as molten (T >= T_liquidus, cF:2585)
    T = np.full((7, 11, 9), 300.0)
    T[-1, 5, 4] = Tl
    d, w, l, n, ylo, yhi = oracle.melt_pool_dims(T.reshape(-1), ax, Tl)[:6]
    assert n == 1 and w == 0.0 and ylo == yhi == pytest.approx(ax[1][5])


def test_melt_pool_brute_force_along_grid_lines(oracle):
    """Random fields: the extents equal those of the linear interpolant
    sampled densely along every grid line (to the sampling step)."""
    rng = np.random.default_rng(3)
    ax = _pool_mesh(n=(6, 5, 4))
    Tl = 1600.0
    S = 2001
    t = np.linspace(0.0, 1.0, S)
    for trial in range(4):
        T = rng.uniform(1000.0, 2000.0, size=(4, 5, 6))
        T[-1, 0, :] = 500.0 if trial % 2 else T[-1, 0, :]
        d, w, l, n, ylo, yhi, xlo, xhi = oracle.melt_pool_dims(T.reshape(-1), ax, Tl)
        ztop = ax[2][-1]
        depth_bf = 0.0
        for j in range(5):
            for i in range(6):
                for k in range(3):
                    v = T[k, j, i] + (T[k + 1, j, i] - T[k, j, i]) * t
                    z = ax[2][k] + (ax[2][k + 1] - ax[2][k]) * t
                    if (v >= Tl).any():
                        depth_bf = max(depth_bf, ztop - z[v >= Tl].min())
        ys, xs = [], []
        for i in range(6):
            for j in range(4):
                v = T[-1, j, i] + (T[-1, j + 1, i] - T[-1, j, i]) * t
                ys += list((ax[1][j] + (ax[1][j + 1] - ax[1][j]) * t)[v >= Tl])
        for j in range(5):
            for i in range(5):
                v = T[-1, j, i] + (T[-1, j, i + 1] - T[-1, j, i]) * t
                xs += list((ax[0][i] + (ax[0][i + 1] - ax[0][i]) * t)[v >= Tl])
        step = max(ax[0][1] - ax[0][0], ax[1][1] - ax[1][0], ax[2][1] - ax[2][0]) / (S - 1)
        assert d == pytest.approx(depth_bf, abs=1.01 * step)
        assert ylo == pytest.approx(min(ys), abs=1.01 * step)
        assert yhi == pytest.approx(max(ys), abs=1.01 * step)
        assert xlo == pytest.approx(min(xs), abs=1.01 * step)
        assert xhi == pytest.approx(max(xs), abs=1.01 * step)
        assert w == pytest.approx(max(ys) - min(ys), abs=2.02 * step)
        assert n == int((T >= Tl).sum())


# ---------------------------------------------------------------------------
# energy balance of an explicit update (the library's audit; its identity is
# that of solveMatrixFreeFE, cF:582-642)
# ---------------------------------------------------------------------------

def _energy_case(O, seed=0, elements=(5, 4, 3)):
    rng = np.random.default_rng(seed)
    lv = _level(O, elements, {"x": [0.0, 0.25], "y": [0.0, 0.16], "z": [-0.09, 0.0]})
    nn, ne = lv["nn"], lv["ne"]
    T = rng.uniform(300.0, 2000.0, nn)
    k = rng.uniform(0.01, 0.03, nn)
    rc = rng.uniform(3e-3, 5e-3, nn)
    Fc = rng.uniform(-1.0, 5.0, nn)
    dt = 1e-5
    return lv, nn, ne, T, k, rc, Fc, dt


def test_energy_balance_of_the_solver_closes(oracle):
    """The FP64 solver's own update closes the balance to rounding, and the
    stiffness part sums to zero (K annihilates constants)."""
    O = oracle
    lv, nn, ne, T, k, rc, Fc, dt = _energy_case(O)
    Tn = O.solveMatrixFreeFE(lv, nn, ne, k, rc, dt, T, Fc, 0.0)
    s = O.energy_balance(lv, ne, k, rc, dt, T, Fc, 0.0, Tn)
    assert abs(s[6]) < 1e-12
    assert s[1] == 0.0 and s[7] == 0.0
    assert s[0] == pytest.approx(s[2], rel=1e-11)
    KTabs = np.abs(dt * (O.solveMatrixFreeFE(lv, nn, ne, k, rc, dt, T, 0.0, 0.0) - T)).sum()
    assert abs(s[3]) < 1e-12 * max(KTabs, 1.0)


def test_energy_balance_constraints_and_planted_error(oracle):
    """Dirichlet faces and the T_amb floor are held nodes: with them named
    the balance closes and their reaction equals C (Tn - Tt); an error of
    delta planted at a free interior node moves the sum by exactly
    rho*cp*h^3*delta (capacity of an interior node: 8 elements of V/8)."""
    O = oracle
    lv, nn, ne, T, k, rc, Fc, dt = _energy_case(O, seed=1)
    Tt = O.solveMatrixFreeFE(lv, nn, ne, k, rc, dt, T, Fc, 0.0)
    nx, ny, nz = lv["nodes"]
    iz, iy, ix = np.unravel_index(np.arange(nn), (nz, ny, nx))
    faces = (ix == 0) | (iy == ny - 1) | (iz == 0)
    mask = 1 | 8 | 16
    Tn = Tt.copy()
    Tn[faces] = 300.0
    floor = 1000.0
    Tn = np.where(faces, Tn, np.maximum(Tn, floor))
    held_floor = (~faces) & (Tt < floor)
    assert held_floor.any()
    s = O.energy_balance(lv, ne, k, rc, dt, T, Fc, 0.0, Tn, bc_faces=mask, T_clamp=floor)
    assert abs(s[6]) < 1e-12
    assert s[7] == faces.sum() + held_floor.sum()
    # the reaction, from the capacities of the solver's own mass sum
    C = dt * O.solveMatrixFreeFE(lv, nn, ne, np.zeros(nn), rc, dt, np.zeros(nn), 1.0, 0.0) ** -1
    held = faces | held_floor
    assert s[4] == pytest.approx(float((C * (Tn - Tt))[held].sum()), rel=1e-10)
    # without naming the faces, their reaction shows up in the residual (face
    # nodes at 300 K whose update fell below the floor still count as held)
    s_nf = O.energy_balance(lv, ne, k, rc, dt, T, Fc, 0.0, Tn, bc_faces=0, T_clamp=floor)
    react = float((C * (Tn - Tt))[faces & (Tt >= floor)].sum())
    assert abs(react) > 1e-3 * abs(s_nf[5])
    assert (s_nf[0] + s_nf[1] - s_nf[2]) == pytest.approx(react, rel=1e-9)
    # a planted error at an interior node
    i = 2 + nx * (2 + ny * 1)
    assert 0 < ix[i] < nx - 1 and 0 < iy[i] < ny - 1 and 0 < iz[i] < nz - 1
    rc_u = np.full(nn, 4e-3)
    Tt2 = O.solveMatrixFreeFE(lv, nn, ne, k, rc_u, dt, T, Fc, 0.0)
    base = O.energy_balance(lv, ne, k, rc_u, dt, T, Fc, 0.0, Tt2)
    Tp = Tt2.copy()
    Tp[i] += 7.0
    s = O.energy_balance(lv, ne, k, rc_u, dt, T, Fc, 0.0, Tp)
    h = lv["h"]
    assert s[0] - base[0] == pytest.approx(4e-3 * h[0] * h[1] * h[2] * 7.0, rel=1e-10)


def test_energy_balance_active_layers(oracle):
    """Only nodes of active elements take part (L1 before its top layers are
    built): the nodes above the active layers are skipped, and the balance
    of the solver's update still closes."""
    O = oracle
    lv, nn, ne, T, k, rc, Fc, dt = _energy_case(O, seed=2)
    nx, ny = lv["nodes"][0], lv["nodes"][1]
    ne_a = lv["elements"][0] * lv["elements"][1] * 2
    Tn = O.solveMatrixFreeFE(lv, nn, ne_a, k, rc, dt, T, Fc, 0.0)
    Tn = np.where(np.isfinite(Tn), Tn, 12345.0)  # 0/0 above the active layers
    s = O.energy_balance(lv, ne_a, k, rc, dt, T, Fc, 0.0, Tn)
    assert abs(s[6]) < 1e-12
    n_act = nx * ny * 3
    assert s[2] == pytest.approx(dt * Fc[:n_act].sum(), rel=1e-12)


# ---------------------------------------------------------------------------
# convolutions of the learned operator (oracle/conv_oracle.py)
# ---------------------------------------------------------------------------

def test_conv_identity_and_shift_kernels():
    from oracle import conv_oracle as CO
    rng = np.random.default_rng(0)
    X = rng.standard_normal((2, 4, 5, 6, 3))
    W = np.zeros((27, 3, 3))
    W[13] = np.eye(3)  # centre tap: identity
    np.testing.assert_allclose(CO.conv3d(X, W), X, atol=0, rtol=0)
    # tap (dz, dy, dx) = (1, 0, -1): Y[z, y, x] = X[z+1, y, x-1], zero outside
    W = np.zeros((27, 3, 3))
    W[2 * 9 + 1 * 3 + 0] = np.eye(3)
    Y = CO.conv3d(X, W)
    np.testing.assert_array_equal(Y[:, :-1, :, 1:], X[:, 1:, :, :-1])
    assert np.all(Y[:, -1] == 0) and np.all(Y[:, :, :, 0] == 0)
    # bias and channel mixing on one tap
    M = rng.standard_normal((2, 3))
    W = np.zeros((27, 2, 3))
    W[13] = M
    np.testing.assert_allclose(CO.conv3d(X, W, [1.0, -2.0]), X @ M.T + [1.0, -2.0], rtol=1e-12)


def test_conv_gradients_are_adjoints():
    """<dY, conv(X; W)> = <bwd_data(dY; W), X> = <bwd_weight(X, dY), W>."""
    from oracle import conv_oracle as CO
    rng = np.random.default_rng(1)
    X = rng.standard_normal((1, 3, 4, 5, 4))
    W = rng.standard_normal((27, 6, 4))
    dY = rng.standard_normal((1, 3, 4, 5, 6))
    lhs = float((dY * CO.conv3d(X, W)).sum())
    assert float((CO.conv3d_bwd_data(dY, W) * X).sum()) == pytest.approx(lhs, rel=1e-12)
    assert float((CO.conv3d_bwd_weight(X, dY) * W).sum()) == pytest.approx(lhs, rel=1e-12)


def test_conv_two_layer_gradients_finite_differences():
    from oracle import conv_oracle as CO
    rng = np.random.default_rng(2)
    X = rng.standard_normal((1, 3, 3, 4, 2))
    W1 = 0.3 * rng.standard_normal((27, 4, 2))
    b1 = 0.1 * rng.standard_normal(4)
    W2 = 0.3 * rng.standard_normal((27, 3, 4))
    b2 = 0.1 * rng.standard_normal(3)
    T = rng.standard_normal((1, 3, 3, 4, 3))
    L, (dW1, db1, dW2, db2), _ = CO.two_layer_loss(X, W1, b1, W2, b2, T)
    eps = 1e-6
    for P, G in ((W1, dW1), (b1, db1), (W2, dW2), (b2, db2)):
        for idx in [tuple(rng.integers(s) for s in P.shape) for _ in range(6)]:
            old = P[idx]
            P[idx] = old + eps
            Lp = CO.two_layer_loss(X, W1, b1, W2, b2, T)[0]
            P[idx] = old - eps
            Lm = CO.two_layer_loss(X, W1, b1, W2, b2, T)[0]
            P[idx] = old
            assert (Lp - Lm) / (2 * eps) == pytest.approx(G[idx], rel=1e-5, abs=1e-7)


def test_bf16_rounding():
    from oracle import conv_oracle as CO
    v = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -3.0, 0.0], np.float32)
    # 1 + 2^-8 ties to even (1.0); 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7
    np.testing.assert_array_equal(CO.bf16_round(v), [1.0, 1.0, 1.0078125, 1.0078125, -3.0, 0.0])
    assert CO.bf16_bits(np.float32(1.0)) == 0x3F80
