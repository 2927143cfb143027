"""ctypes binding of libgomelt.so (the C ABI declared in include/gomelt.h).

There is no fallback: if the library or a GPU is missing, calls raise
``GoMeltNativeError``.  Device memory is owned by :class:`DeviceArray`.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgomelt.so")
GM_MAX_KNOTS = 16

(GM_T0, GM_TPRIME0, GM_S1, GM_S2, GM_ACCUM, GM_MAX_ACCUM, GM_HIST_L2, GM_HIST_L3,
 GM_HIST_TP3, GM_ACCUM_L3, GM_MAX_ACCUM_L3, GM_RESET_L3) = range(12)
(GM_STAGE_L3_IN, GM_STAGE_L3_OUT, GM_STAGE_FACE_NEW, GM_STAGE_FACE_OLD,
 GM_STAGE_S1_IN, GM_STAGE_L3_LAST) = range(6)


class GoMeltNativeError(RuntimeError):
    pass


class gm_grid(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("x0", C.c_float * 3), ("h", C.c_float * 3)]

    @classmethod
    def make(cls, n, x0, h):
        g = cls()
        for i in range(3):
            g.n[i] = int(n[i])
            g.x0[i] = float(np.float32(x0[i]))
            g.h[i] = float(np.float32(h[i]))
        return g

    def nn(self):
        return int(self.n[0]) * int(self.n[1]) * int(self.n[2])

    def axes(self):
        """Node coordinate axes (FP32), node i at x0 + i*h."""
        return [np.float32(self.x0[d]) + np.arange(self.n[d], dtype=np.float32)
                * np.float32(self.h[d]) for d in range(3)]


class gm_material(C.Structure):
    _fields_ = [
        ("T_amb", C.c_float), ("T_solidus", C.c_float), ("T_liquidus", C.c_float),
        ("T_boiling", C.c_float),
        ("k_powder", C.c_float), ("k_bulk_a0", C.c_float), ("k_bulk_a1", C.c_float),
        ("k_fluid", C.c_float),
        ("cp_solid_a0", C.c_float), ("cp_solid_a1", C.c_float), ("cp_mushy", C.c_float),
        ("cp_fluid", C.c_float),
        ("rho", C.c_float),
        ("k_table_n", C.c_int32), ("cp_table_n", C.c_int32),
        ("k_table_T", C.c_float * GM_MAX_KNOTS), ("k_table_v", C.c_float * GM_MAX_KNOTS),
        ("cp_table_T", C.c_float * GM_MAX_KNOTS), ("cp_table_v", C.c_float * GM_MAX_KNOTS),
        ("h_conv", C.c_float), ("sigma_eps", C.c_float),
        ("evc_cp", C.c_float), ("CT", C.c_float), ("CM", C.c_float), ("Lev", C.c_float),
        ("laser_eta", C.c_float), ("laser_radius", C.c_float), ("laser_depth", C.c_float),
    ]


class gm_config(C.Structure):
    _fields_ = [
        ("level", gm_grid * 4),
        ("l1_bc", C.c_float * 5),
        ("layer_height", C.c_float),
        ("n2", C.c_int32), ("n3", C.c_int32),
        ("ratio12", C.c_int32 * 3), ("ratio23", C.c_int32 * 3),
        ("move_lo", (C.c_float * 3) * 2), ("move_hi", (C.c_float * 3) * 2),
        ("ov_lo", (C.c_int32 * 3) * 4), ("ov_n", (C.c_int32 * 3) * 4),
        ("ov_x0", (C.c_float * 3) * 4),
        ("l0_stride_l2", C.c_int32 * 3),
        ("l0_layer_delta", C.c_int32),
    ]


_P = C.c_void_p
_G = C.POINTER(gm_grid)
_M = C.POINTER(gm_material)
_I64 = C.c_int64
_F = C.c_float
_I = C.c_int
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int32)

_SIGS = {
    "gm_last_error": (C.c_char_p, []),
    "gm_device_init": (_I, [_I]),
    "gm_device_name": (_I, [C.c_char_p, _I]),
    "gm_malloc": (_I, [C.POINTER(_P), C.c_size_t]),
    "gm_free": (_I, [_P]),
    "gm_memcpy_h2d": (_I, [_P, _P, C.c_size_t, _P]),
    "gm_memcpy_d2h": (_I, [_P, _P, C.c_size_t, _P]),
    "gm_memcpy_d2d": (_I, [_P, _P, C.c_size_t, _P]),
    "gm_memset": (_I, [_P, _I, C.c_size_t, _P]),
    "gm_stream_create": (_I, [C.POINTER(_P)]),
    "gm_stream_destroy": (_I, [_P]),
    "gm_stream_sync": (_I, [_P]),
    "gm_event_create": (_I, [C.POINTER(_P)]),
    "gm_event_destroy": (_I, [_P]),
    "gm_event_record": (_I, [_P, _P]),
    "gm_stream_wait_event": (_I, [_P, _P]),
    "gm_event_elapsed_ms": (_I, [_P, _P, _FP]),
    "gm_state_properties": (_I, [_M, _P, _P, _I64, _I64, _P, _P, _P, _P, _P]),
    "gm_solve_matrix_free": (_I, [_G, _I64, _P, _P, _F, _P, _P, _P, _P, _P]),
    "gm_source_l3": (_I, [_G, _M, _FP, _F, _P, _P]),
    "gm_source_l3_mean": (_I, [_G, _M, _FP, _I, _P, _P]),
    "gm_level_source": (_I, [_G, _G, _M, _FP, _I, _P, _P]),
    "gm_conv_rad_bc": (_I, [_G, _M, _P, _I64, _P, _P]),
    "gm_melt_pool": (_I, [_G, _P, _F, _P, _P]),
    "gm_energy_balance": (_I, [_G, _I64, _P, _P, _F, _P, _P, _P, _P, _I, _F, _P, _P]),
    "gm_interpolate_points": (_I, [_G, _P, _G, _P, _P]),
    "gm_interpolate_matrix": (_I, [_G, _G, _P, _P, _P]),
    "gm_interpolate_w_matrix": (_I, [_P, _P, _I64, _P, _P, _P]),
    "gm_restrict": (_I, [_G, _G, _M, _P, _P, _I64, _P, _P, _P, _F, _P, _P]),
    "gm_restrict_noalloc": (_I, [_G, _G, _M, _P, _P, _I64, _P, _P, _P, _F, _P, _P]),
    "gm_new_tprime": (_I, [_G, _P, _G, _P, _IP, _IP, _P, _P]),
    "gm_engine_create": (_I, [C.POINTER(gm_config), _M, _P, C.POINTER(_P)]),
    "gm_engine_destroy": (_I, [_P]),
    "gm_engine_field": (_I, [_P, _I, _I, C.POINTER(_P), C.POINTER(_I64)]),
    "gm_engine_grid": (_I, [_P, _I, _G]),
    "gm_engine_set_layer": (_I, [_P, _I64, _I64, C.POINTER(_I64)]),
    "gm_engine_move": (_I, [_P, _FP, _FP, _IP]),
    "gm_engine_subcycle": (_I, [_P, _FP, _I]),
    "gm_engine_step": (_I, [_P, _FP]),
    "gm_engine_step_dwell": (_I, [_P, _F]),
    "gm_engine_layer_change": (_I, [_P, _I, _F, _F, _I]),
    "gm_engine_clear_tprime": (_I, [_P]),
    "gm_engine_level0_substrate": (_I, [_P, _I64]),
    "gm_engine_minmax": (_I, [_P, _FP]),
    "gm_engine_use_graph": (_I, [_P, _I]),
    "gm_engine_time_kernel": (_I, [_P, _I, _I, _FP]),
    "gm_engine_set_accum_mode": (_I, [_P, _I]),
    "gm_engine_overlap_box": (_I, [_P, _I, _IP, _IP, _IP]),
    "gm_engine_state_bytes": (_I, [_P, C.POINTER(_I64)]),
    "gm_engine_save_state": (_I, [_P, _P]),
    "gm_engine_load_state": (_I, [_P, _P, _I64]),
    "gm_batch_create": (_I, [_I, _G, _M, _P, C.POINTER(_P)]),
    "gm_batch_destroy": (_I, [_P]),
    "gm_batch_step": (_I, [_P, _FP, _I]),
    "gm_batch_set_tile": (_I, [_P, _I, _I]),
    "gm_batch_autotune": (_I, [_P, _I, _FP, _IP, _IP]),
    "gm_batch_time": (_I, [_P, _I, _I, _FP]),
    "gm_batch_field": (_I, [_P, _I, _I, C.POINTER(_P), C.POINTER(_I64)]),
    "gm_batch_refresh": (_I, [_P]),
    "gm_batch_jump_state": (_I, [_P, _I, _P]),
    "gm_conv3d_fwd": (_I, [_IP, _P, _P, _P, _I, _P, _P, _P]),
    "gm_conv3d_fwd_ws_bytes": (_I, [_IP, C.POINTER(_I64)]),
    "gm_conv3d_fwd_ws": (_I, [_IP, _P, _P, _P, _I, _P, _P, _P, _I64, _P]),
    "gm_conv3d_bwd_data": (_I, [_IP, _P, _P, _P, _P, _P, _P]),
    "gm_conv3d_bwd_weight": (_I, [_IP, _P, _P, _P, _P]),
    "gm_adam": (_I, [_P, _P, _P, _P, _P, _I64, _F, _F, _F, _F, _I, _P]),
    "gm_f32_to_bf16": (_I, [_P, _P, _I64, _P]),
    "gm_nn_pool2": (_I, [_I, _IP, _P, _P, _P]),
    "gm_nn_up2": (_I, [_I, _IP, _P, _P, _P]),
    "gm_nn_concat": (_I, [_I64, _I, _I, _P, _P, _P, _P]),
    "gm_nn_split": (_I, [_I64, _I, _I, _P, _P, _P, _P]),
    "gm_nn_relu_bwd": (_I, [_I64, _P, _P, _P, _P, _P, _P]),
    "gm_nn_bias_grad": (_I, [_I64, _I, _P, _P, _P]),
    "gm_nn_relu_bias_bwd": (_I, [_I64, _I, _P, _P, _P, _P, _P, _P, _P]),
    "gm_nn_mse": (_I, [_I64, _I, _P, _P, _P, _P, _P]),
    "gm_nn_window_input": (_I, [_IP, _FP, _P, _P, _FP, _P, _P]),
    "gm_nn_window_apply": (_I, [_IP, _P, _P, _P, _I, _F, _F, _P, _P]),
    "gm_nn_axpy": (_I, [_I64, _F, _P, _P, _P]),
    "gm_u8_to_f32": (_I, [_I64, _P, _P, _P]),
    "gm_nn_mse_energy": (_I, [_I64, _I, _I, _I, _I, _F, _P, _P, _P, _P, _P, _P]),
    "gm_engine_stage_begin": (_I, [_P, _P, _I]),
    "gm_engine_stage_l2": (_I, [_P, _P, _P]),
    "gm_engine_stage_l3": (_I, [_P, _I]),
    "gm_engine_stage_field": (_I, [_P, _I, _P, _P]),
    "gm_multi_create": (_I, [_P, _I, _P]),
    "gm_multi_l3": (_I, [_P]),
    "gm_multi_subcycle": (_I, [_P]),
    "gm_multi_set_tile": (_I, [_P, _I, _I]),
    "gm_multi_set_batch_l2": (_I, [_P, _I]),
    "gm_multi_set_batch_restrict": (_I, [_P, _I]),
    "gm_multi_destroy": (_I, [_P]),
}

SYMBOLS = tuple(_SIGS)

_lib = None
_device_ready = False


def lib():
    """Load libgomelt.so (once) and declare every entry point's signature."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GoMeltNativeError(
                f"{LIB_PATH} is missing: build it with `make -C go-melt_amd/csrc` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise GoMeltNativeError(lib().gm_last_error().decode())


def init_device(device=0):
    """Select the GPU; raises if HIP finds none."""
    global _device_ready
    if not _device_ready:
        check(lib().gm_device_init(int(device)))
        _device_ready = True


def device_name():
    init_device()
    buf = C.create_string_buffer(256)
    check(lib().gm_device_name(buf, 256))
    return buf.value.decode()


def sync(stream=None):
    check(lib().gm_stream_sync(stream))


class DeviceArray:
    """A flat device buffer (FP32 by default).  ``owner`` False marks a view of
    memory owned by an engine."""

    def __init__(self, n, dtype=np.float32, ptr=None, owner=True):
        init_device()
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.owner = owner and ptr is None
        if ptr is None:
            p = C.c_void_p()
            check(lib().gm_malloc(C.byref(p), max(1, self.n * self.dtype.itemsize)))
            ptr = p.value
        self.ptr = ptr

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @classmethod
    def from_numpy(cls, a, dtype=None, stream=None):
        a = np.ascontiguousarray(a, dtype=dtype or (a.dtype if a.dtype != np.float64 else np.float32))
        d = cls(a.size, a.dtype)
        check(lib().gm_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes, stream))
        sync(stream)
        return d

    def numpy(self, stream=None):
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().gm_memcpy_d2h(out.ctypes.data, self.ptr, self.nbytes, stream))
        return out

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(-1)
        if a.size != self.n:
            raise ValueError("size mismatch")
        check(lib().gm_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream))
        sync(stream)

    def zero(self, stream=None):
        check(lib().gm_memset(self.ptr, 0, self.nbytes, stream))

    def free(self):
        if self.owner and self.ptr:
            lib().gm_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self):
        return self.n

    def __array__(self, dtype=None, copy=None):
        a = self.numpy()
        return a.astype(dtype) if dtype is not None else a

    # Host reads in the style of a jax.Array, for driver code written against
    # the reference (go_melt.py): comparisons, arithmetic and indexing
    # download the buffer and return host arrays, and ``.at[i].set(v)`` /
    # ``.at[i].add(v)`` return an updated host copy (the buffer itself is
    # not changed, as with the functional update of jax.Array).
    __array_priority__ = 100
    __hash__ = object.__hash__

    def __getitem__(self, i):
        return self.numpy()[i]

    def _cmp(self, o, op):
        return op(self.numpy(), np.asarray(o))

    def __gt__(self, o):
        return self._cmp(o, np.greater)

    def __ge__(self, o):
        return self._cmp(o, np.greater_equal)

    def __lt__(self, o):
        return self._cmp(o, np.less)

    def __le__(self, o):
        return self._cmp(o, np.less_equal)

    def __eq__(self, o):
        return self._cmp(o, np.equal)

    def __ne__(self, o):
        return self._cmp(o, np.not_equal)

    def __mul__(self, o):
        return self.numpy() * np.asarray(o)

    __rmul__ = __mul__

    def __add__(self, o):
        return self.numpy() + np.asarray(o)

    __radd__ = __add__

    def __sub__(self, o):
        return self.numpy() - np.asarray(o)

    def __rsub__(self, o):
        return np.asarray(o) - self.numpy()

    @property
    def shape(self):
        return (self.n,)

    @property
    def at(self):
        return _At(self.numpy())


class _At:
    """``a.at[idx].set(v)`` / ``.add(v)`` / ``.get()`` on a host copy."""

    def __init__(self, a):
        self._a = a

    def __getitem__(self, idx):
        return _AtIndex(self._a, idx)


class _AtIndex:
    def __init__(self, a, idx):
        self._a, self._i = a, idx

    def set(self, v):
        b = self._a.copy()
        b[self._i] = v
        return b

    def add(self, v):
        b = self._a.copy()
        np.add.at(b, self._i, v)
        return b

    def get(self):
        return self._a[self._i]


def as_device(x, dtype=np.float32):
    """DeviceArray for x (uploaded if it is host data); keeps a DeviceArray;
    a GPU tensor (PyTorch) of the same element type is viewed, not copied."""
    if isinstance(x, DeviceArray):
        return x
    if getattr(x, "is_cuda", False) and hasattr(x, "data_ptr"):
        if x.element_size() != np.dtype(dtype).itemsize or not x.is_contiguous():
            raise GoMeltNativeError("a GPU tensor argument must be contiguous and of the "
                                    f"expected element type ({np.dtype(dtype)})")
        return DeviceArray(x.numel(), dtype, ptr=x.data_ptr(), owner=False)
    return DeviceArray.from_numpy(np.asarray(x), dtype=dtype)


class Stream:
    def __init__(self):
        init_device()
        p = C.c_void_p()
        check(lib().gm_stream_create(C.byref(p)))
        self.ptr = p.value

    def sync(self):
        check(lib().gm_stream_sync(self.ptr))

    def __del__(self):
        try:
            if self.ptr:
                lib().gm_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().gm_event_create(C.byref(p)))
        self.ptr = p.value

    def record(self, stream):
        check(lib().gm_event_record(self.ptr, stream.ptr if isinstance(stream, Stream) else stream))

    def wait_on(self, stream):
        """Later work on ``stream`` waits for this event."""
        check(lib().gm_stream_wait_event(stream.ptr if isinstance(stream, Stream) else stream,
                                         self.ptr))

    def elapsed_ms(self, later):
        ms = C.c_float()
        check(lib().gm_event_elapsed_ms(self.ptr, later.ptr, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().gm_event_destroy(self.ptr)
        except Exception:
            pass


def fptr(a):
    """ctypes float* to a host float32 array."""
    return a.ctypes.data_as(_FP)
