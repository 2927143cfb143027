"""Batched process-parameter instances.

* Batch: one kernel launch per step for many independent L3 windows of their
  own mesh size and laser (gm_batch_* in include/gomelt.h).
* Stacks: many independent three-level runs advanced in lockstep, each
  substep stage of all of them one launch (gm_multi_*).

There is no CPU fallback: without libgomelt.so or a GPU every call raises."""

from __future__ import annotations

import ctypes as C

import numpy as np

import _native as N
from computeFunctions import SetupProperties, _material


class Batch:
    """n instances; ``grids`` are gm_grid, ``props`` reference-schema
    property dicts (one per instance)."""

    def __init__(self, grids, props, stream=None):
        N.init_device()
        self.n = len(grids)
        self.stream = stream or N.Stream()
        self.grids = (N.gm_grid * self.n)(*grids)
        self.props = [SetupProperties(p) for p in props]
        mats = (N.gm_material * self.n)(*[_material(p) for p in self.props])
        h = C.c_void_p()
        N.check(N.lib().gm_batch_create(self.n, self.grids, mats, self.stream.ptr, C.byref(h)))
        self.h = h.value

    @classmethod
    def from_instances(cls, instances, stream=None):
        """From workloads.batch_instances(): the mesh of each instance spans
        [0, nx h] x [0, ny h] x [-nz h, 0]."""
        grids = []
        for b in instances:
            nel, h = b["elements"], b["h"]
            grids.append(N.gm_grid.make([e + 1 for e in nel], [0.0, 0.0, -nel[2] * h], [h] * 3))
        return cls(grids, [b["properties"] for b in instances], stream)

    def __del__(self):
        try:
            if self.h:
                N.lib().gm_batch_destroy(self.h)
        except Exception:
            pass

    def step(self, rows, batched=True):
        r = np.ascontiguousarray(rows, np.float32).reshape(self.n, 7)
        N.check(N.lib().gm_batch_step(self.h, N.fptr(r), 1 if batched else 0))

    TILES = ((1, 2), (1, 3), (1, 4), (2, 2), (2, 3))

    def set_tile(self, yb, zc):
        """Brick shape of the batched launch: yb rows x zc planes (TILES)."""
        N.check(N.lib().gm_batch_set_tile(self.h, int(yb), int(zc)))

    def autotune(self, iters=20):
        """Times every brick shape, keeps the fastest; returns ((yb, zc), ms
        per shape in TILES order)."""
        ms = np.zeros(len(self.TILES), np.float32)
        yb, zc = C.c_int32(), C.c_int32()
        N.check(N.lib().gm_batch_autotune(self.h, int(iters), N.fptr(ms), C.byref(yb),
                                          C.byref(zc)))
        return (yb.value, zc.value), ms

    def time(self, batched=True, iters=20):
        ms = C.c_float()
        N.check(N.lib().gm_batch_time(self.h, 1 if batched else 0, int(iters), C.byref(ms)))
        return ms.value

    def _field(self, i, f):
        p = C.c_void_p()
        n = C.c_int64()
        N.check(N.lib().gm_batch_field(self.h, int(i), int(f), C.byref(p), C.byref(n)))
        return N.DeviceArray(n.value, np.uint8 if f == 1 else np.float32, ptr=p.value,
                             owner=False)

    def T(self, i):
        self.stream.sync()
        return self._field(i, 0).numpy()

    def S1(self, i):
        self.stream.sync()
        return self._field(i, 1).numpy()

    def set_fields(self, i, T=None, S1=None, faces=None):
        """Write T, S1 and/or the face values of instance i (host arrays)."""
        self.stream.sync()
        for f, v, dt in ((0, T, np.float32), (1, S1, np.uint8), (2, faces, np.float32)):
            if v is not None:
                d = self._field(i, f)
                a = np.ascontiguousarray(v, dt).reshape(-1)
                assert a.size == d.n
                N.check(N.lib().gm_memcpy_h2d(d.ptr, a.ctypes.data, a.nbytes, self.stream.ptr))
                self.stream.sync()  # the host array must outlive the copy
        N.check(N.lib().gm_batch_refresh(self.h))


class _Deferred:
    """Operator hook of one run in a Stacks: its subcycleGOMELT only takes the
    rows and runs the prologue; Stacks.step runs the L2 and L3 stages of all
    runs together."""

    def __init__(self, owner):
        import weakref
        self.owner = weakref.ref(owner)  # no cycle through the runs' engines

    def subcycle(self, e, rows):
        rows = np.ascontiguousarray(rows, dtype=np.float32)
        N.check(N.lib().gm_engine_stage_begin(e.h, N.fptr(rows), rows.shape[0]))
        self.owner()._pending.append(e)


class Stacks:
    """Independent three-level runs (each its own L1/L2/L3 stack, the
    reference's time loop gm:134-493 per run) advanced together.

    One TimeLoop per deck, the engines dealt over `streams` streams.  Each step gives every
    run its next chunk of N2*N3 toolpath rows; a single-step chunk runs as in
    the plain driver, a subcycled one (subcycleGOMELT cF:3224-3632) stage by
    stage in lockstep with the others: per L2 substep one batched launch each
    for the restrictions of the L3 T' terms, the L2 solves, the L3 face
    values, the N3 L3 solves and T' + injection of all subcycling runs
    (gm_multi_subcycle; batch_l2=False batches the L3 solves only,
    batch_restrict=False leaves each run's restriction on its own stream).
    Every run ends where its own unbatched run ends."""

    def __init__(self, decks, streams=4, tile="auto", batch_l2=True, batch_restrict=True):
        import computeFunctions as cf
        import go_melt as G
        N.init_device()
        self.decks = decks
        # runs dealt round-robin over a few streams, so the small per-run
        # kernels (L2 stages, moves, single steps) of different runs overlap
        self.streams = [N.Stream() for _ in range(max(1, min(int(streams), len(decks))))]
        self.runs = []
        for i, d in enumerate(decks):
            with cf.engines_on_stream(self.streams[i % len(self.streams)]):
                self.runs.append(G.TimeLoop(d, quiet=True, write_outputs=False))
        self.n = len(self.runs)
        self.batch_l2 = bool(batch_l2)
        self.batch_restrict = bool(batch_restrict)
        self._pending = []
        for r in self.runs:
            cf.attach_l3_operator(r.Levels, _Deferred(self))
        self._iters = [r.chunks() for r in self.runs]
        self._multi = None
        self._ev = [N.Event(), N.Event()]
        # brick shape of the batched L3 solves: the autotuner's pick on the
        # runs' L3 windows ("auto"), a given (rows, planes), or None (1 x 3)
        self.tile_ms = None
        self.tile = self.autotune_tile() if tile == "auto" else tile

    def autotune_tile(self, iters=20):
        """gm_batch_autotune on a batch of the runs' L3 windows (their meshes,
        materials and first laser position, a few steps in): returns the
        fastest (rows, planes); the times per shape go to self.tile_ms."""
        import computeFunctions as cf
        grids = [cf._grid_of(r.Levels[3]) for r in self.runs]
        bt = Batch(grids, [d["properties"] for d in self.decks])
        rows = np.zeros((self.n, 7), np.float32)
        for i, d in enumerate(self.decks):
            p = d["properties"]
            rows[i, :3] = np.asarray(p.get("laser_center", [0.0, 0.0, 0.0])[:3], np.float32)
            rows[i, 5] = float(d["nonmesh"].get("timestep_L3", 1e-6))
            rows[i, 6] = float(p.get("laser_power", 0.0))
        for _ in range(5):
            bt.step(rows)
        tile, ms = bt.autotune(iters)
        self.tile_ms = [float(v) for v in ms]
        del bt
        return tile

    def _multi_of(self, engines):
        key = tuple(e.h for e in engines)
        if self._multi is None or self._multi[0] != key:
            if self._multi is not None:
                N.lib().gm_multi_destroy(self._multi[1])
            arr = (C.c_void_p * len(key))(*key)
            h = C.c_void_p()
            N.check(N.lib().gm_multi_create(arr, len(key), C.byref(h)))
            self._multi = (key, h.value)
            if self.tile is not None:
                N.check(N.lib().gm_multi_set_tile(h.value, int(self.tile[0]), int(self.tile[1])))
            N.check(N.lib().gm_multi_set_batch_l2(h.value, 1 if self.batch_l2 else 0))
            N.check(N.lib().gm_multi_set_batch_restrict(h.value, 1 if self.batch_restrict else 0))
        return self._multi[1]

    def step(self):
        """One chunk of every run; False when a run's toolpath has ended."""
        lib = N.lib()
        self._pending = []
        going = True
        for r, it in zip(self.runs, self._iters):
            rows = next(it, None)
            if not rows:
                going = False
                continue
            going = r.step_chunk(rows) and going
        pend = self._pending
        if pend:
            N.check(lib.gm_multi_subcycle(self._multi_of(pend)))
        self._pending = []
        return going

    def step_each(self):
        """One chunk of every run, one run after another through its own
        recorded subcycle (no batching): the sequential baseline of step.
        Each run's work waits on the GPU for the previous run's (events
        across the streams), the host does not wait."""
        import computeFunctions as cf
        going = True
        prev = None
        for i, (r, it) in enumerate(zip(self.runs, self._iters)):
            rows = next(it, None)
            if not rows:
                going = False
                continue
            s = r.eng.stream
            if prev is not None and prev is not s:
                self._ev[i % 2].wait_on(s)
            cf.attach_l3_operator(r.Levels, None)
            try:
                going = r.step_chunk(rows) and going
            finally:
                cf.attach_l3_operator(r.Levels, _Deferred(self))
            self._ev[(i + 1) % 2].record(s)
            prev = s
        return going

    def sync(self):
        for s in self.streams:
            s.sync()
        for r in self.runs:
            r.eng.sync_coords()

    def close(self):
        """Release the batched-solve object and detach the runs (their
        engines are released with them)."""
        import computeFunctions as cf
        if self._multi is not None:
            N.lib().gm_multi_destroy(self._multi[1])
            self._multi = None
        for r in self.runs:
            cf.attach_l3_operator(r.Levels, None)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
