"""ctypes binding of the C++/OpenMP restatement (oracle/cpu_port.cpp).

Test and measurement infrastructure only: the benchmark's CPU baseline runs
the oracle's orchestration (oracle/gomelt_oracle.py: go_melt, subcycleGOMELT,
stepGOMELT, moveEverything) with the element loops below in place of the
oracle's NumPy ones, in FP32, on all host cores.  Nothing in the product path
imports this module.

``install(O)`` replaces, in the oracle module ``O``, the functions that carry
the element loops of the reference (go_melt/computeFunctions.py = cF):
solveMatrixFreeFE (cF:582-642), computeStateProperties (cF:2567-2614),
computeSourcesL3 (cF:2960-3012), computeSources (cF:928-988),
computeLevelSource (cF:2667-2730), computeConvRadBC (cF:2207-2301),
interpolatePoints / interpolatePointsMatrix / interpolate_w_matrix
(cF:1028-1210), computeCoarseFineShapeFunctions (cF:1213-1358, kept as a
handle of the two meshes: the products below evaluate the coarse shape
functions at the fine Gauss points in the loop) and restrictFlux /
restrictMass (cF:1396-1565).  ``uninstall(O)`` restores the NumPy ones.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgmcpu.so")

_lib = None
_saved = {}

_P = C.c_void_p
_I64 = C.c_int64
_F = C.c_float


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (make -C {HERE})")
        L = C.CDLL(LIB_PATH)
        sig = {
            "gmc_threads": (C.c_int, []),
            "gmc_solve": (C.c_int, [_P, _P, _P, _P, _I64, _P, _P, _F, _P, _P, _P, _P]),
            "gmc_props": (C.c_int, [_I64, _P, _P, _I64, _P, _P, C.c_int, _P, C.c_int, _P, _P,
                                    _P, _P]),
            "gmc_source_l3": (C.c_int, [_P, _P, _P, _P, _P, _P, _F, _P]),
            "gmc_level_source": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                           C.c_int, _P]),
            "gmc_restrict": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, C.c_int, _P, _P, _F,
                                       _P]),
            "gmc_conv_rad": (C.c_int, [_P, _P, _P, _P, _I64, _P, _P, _P, _P]),
            "gmc_interp": (C.c_int, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def threads():
    return int(lib().gmc_threads())


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))


def _p(a):
    return a.ctypes.data


class _Mesh:
    """Node axes and counts of a level (a snapshot: the window may move)."""

    def __init__(self, lv_or_axes):
        axes = lv_or_axes["node_coords"] if isinstance(lv_or_axes, dict) else lv_or_axes
        self.ax = [_f32(a) for a in axes]
        self.n = np.array([a.size for a in self.ax], np.int32)

    def args(self):
        return [_p(self.ax[0]), _p(self.ax[1]), _p(self.ax[2]), _p(self.n)]

    @property
    def nn(self):
        return int(np.prod(self.n.astype(np.int64)))


class _Pair:
    """computeCoarseFineShapeFunctions handle: the coarse and fine meshes."""

    def __init__(self, coarse, fine):
        self.coarse, self.fine = _Mesh(coarse), _Mesh(fine)


class _Interp:
    """interpolatePointsMatrix handle: source mesh and target axes."""

    def __init__(self, lv, coords_new):
        self.src, self.tgt = _Mesh(lv), _Mesh(coords_new)


def _src(p):
    return np.array([p["laser_eta"], p["laser_radius"], p["laser_depth"]], np.float32)


def install(O):
    """Put the C++ element loops in place of the oracle's NumPy ones (FP32)."""
    L = lib()
    O.set_precision("float32")
    names = ["solveMatrixFreeFE", "computeStateProperties", "computeSourcesL3",
             "computeSources", "computeLevelSource", "computeConvRadBC", "interpolatePoints",
             "interpolatePointsMatrix", "interpolate_w_matrix",
             "computeCoarseFineShapeFunctions", "restrictFlux", "restrictMass"]
    for n in names:
        _saved.setdefault(n, getattr(O, n))

    def solveMatrixFreeFE(lv, nn, ne, k, rhocp, dt, T, Fc, Corr):
        m = _Mesh(lv)
        k, rc, T = _f32(k), _f32(rhocp), _f32(T)
        Fc = _f32(np.broadcast_to(np.asarray(Fc, np.float32), (m.nn,)))
        Cr = None if np.isscalar(Corr) or np.asarray(Corr).ndim == 0 else _f32(Corr)
        if Cr is None and float(Corr) != 0.0:
            Cr = np.full(m.nn, float(Corr), np.float32)
        out = np.empty(m.nn, np.float32)
        L.gmc_solve(*m.args(), int(ne), _p(k), _p(rc), float(dt), _p(T), _p(Fc),
                    None if Cr is None else _p(Cr), _p(out))
        return out

    def computeStateProperties(T, S1, p, n_sub):
        T, S1 = _f32(T), _f32(S1)
        n = T.size
        mat = np.array([p["T_solidus"], p["T_liquidus"], p["k_powder"], p["k_bulk_coeff_a0"],
                        p["k_bulk_coeff_a1"], p["k_fluid_coeff_a0"], p["cp_solid_coeff_a0"],
                        p["cp_solid_coeff_a1"], p["cp_mushy"], p["cp_fluid"], p["rho"]],
                       np.float64)
        kt = np.asarray(p["k_table"], np.float64).reshape(-1) if p.get("k_table") else None
        ct = np.asarray(p["cp_table"], np.float64).reshape(-1) if p.get("cp_table") else None
        S1o = np.empty(n, np.float32)
        S2 = np.empty(n, np.uint8)
        k = np.empty(n, np.float32)
        rc = np.empty(n, np.float32)
        L.gmc_props(n, _p(T), _p(S1), int(n_sub), _p(mat),
                    None if kt is None else _p(kt), 0 if kt is None else kt.size // 2,
                    None if ct is None else _p(ct), 0 if ct is None else ct.size // 2,
                    _p(S1o), _p(S2), _p(k), _p(rc))
        return S1o, S2.astype(bool), k, rc

    def computeSourcesL3(L3, v, ne_nn, p, P):
        m = _Mesh(L3)
        out = np.empty(m.nn, np.float32)
        vv = _f32(np.asarray(v, np.float64)[:3])
        L.gmc_source_l3(*m.args(), _p(_src(p)), _p(vv), float(P), _p(out))
        return out

    def _level_source(pair, rows, P, p):
        rows = _f32(np.asarray(rows, np.float64).reshape(-1, 7))
        P = _f32(P)
        out = np.empty(pair.coarse.nn, np.float32)
        L.gmc_level_source(*pair.coarse.args(), *pair.fine.args(), _p(_src(p)), _p(rows),
                           _p(P), int(P.size), _p(out))
        return out

    def computeSources(L3, v, Shapes, ne_nn, p, P):
        Ff = computeSourcesL3(L3, v, ne_nn, p, P)
        row = np.zeros(7)
        row[:3] = np.asarray(v, np.float64)[:3]
        Fc = _level_source(Shapes[1], row, [P], p)
        Fm = _level_source(Shapes[2], row, [P], p)
        return Fc, Fm, Ff

    def computeLevelSource(Levels, ne_nn, laser_position, LevelShape, p, laserP):
        return _level_source(LevelShape, laser_position, laserP, p)

    def computeConvRadBC(lv, T, ne, nn, p, F_in):
        m = _Mesh(lv)
        mat = np.array([p["T_amb"], p["T_boiling"], p["h_conv"],
                        np.float32(p["sigma_sb"]) * np.float32(p["vareps"]), p["evc"],
                        p["CP_coeff"], p["CT_coeff"], p["CM_coeff"], p["Lev"], p["cp_fluid"]],
                       np.float32)
        Fin = _f32(np.broadcast_to(np.asarray(F_in, np.float32), (m.nn,)))
        out = np.empty(m.nn, np.float32)
        L.gmc_conv_rad(*m.args(), int(ne), _p(_f32(T)), _p(mat), _p(Fin), _p(out))
        return out

    def interpolatePointsMatrix(lv, coords_new):
        return _Interp(lv, coords_new)

    def interpolate_w_matrix(C2F, T):
        out = np.empty(C2F.tgt.nn, np.float32)
        u = _f32(T)
        L.gmc_interp(*C2F.src.args(), _p(u), *C2F.tgt.args()[:3], _p(C2F.tgt.n), _p(out))
        return out

    def interpolatePoints(lv, u, coords_new):
        return interpolate_w_matrix(_Interp(lv, coords_new), u)

    def computeCoarseFineShapeFunctions(Coarse, Fine):
        return _Pair(Coarse, Fine)

    def _restrict(pair, mode, u, coef, dt):
        V = np.zeros(pair.coarse.nn, np.float32)
        L.gmc_restrict(*pair.coarse.args(), *pair.fine.args(), int(mode), _p(_f32(u)),
                       _p(_f32(coef)), float(dt), _p(V))
        return V

    def restrictFlux(Shape, Fine, Tp, k):
        return _restrict(Shape, 0, Tp, k, 1.0)

    def restrictMass(Shape, Fine, dTp, rhocp, dt):
        return _restrict(Shape, 1, dTp, rhocp, dt)

    for n in names:
        setattr(O, n, locals()[n])


def uninstall(O):
    for n, f in _saved.items():
        setattr(O, n, f)
    _saved.clear()
