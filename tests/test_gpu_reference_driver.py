"""The reference driver's own melt-time lines on the facade's return values.

The reference driver keeps the Level0 melt-time maps itself and feeds the L3
window of them through stepGOMELT's reset mask and subcycleGOMELT's window
arrays (go_melt.py:339-357 and 436-456).  Here those lines run as written
there, on arrays with the jax.Array ``.at`` update (a NumPy stand-in for
``jnp``), inside the facade's time loop; the maps they build are compared
with the oracle driver's, and with the maps the engine keeps itself when the
same deck runs through the plain facade loop.
"""
import json

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


class _JArr(np.ndarray):
    """NumPy array with jax.Array's functional ``.at[idx].set/add``."""

    @property
    def at(self):
        arr = self

        class _Idx:
            def __getitem__(self, idx):
                class _Op:
                    def set(self, v):
                        b = np.array(arr, copy=True).view(_JArr)
                        b[idx] = v
                        return b

                    def add(self, v):
                        b = np.array(arr, copy=True).view(_JArr)
                        np.add.at(b, idx, v)
                        return b
                return _Op()
        return _Idx()


class jnp:  # noqa: N801  (the name the reference's lines use)
    @staticmethod
    def zeros(n):
        return np.zeros(n, np.float32).view(_JArr)

    @staticmethod
    def maximum(a, b):
        return np.maximum(np.asarray(a), np.asarray(b)).view(_JArr)


def _reference_loop_class():
    import go_melt as G
    from computeFunctions import melting_temp, subcycleGOMELT

    class ReferenceAccumLoop(G.TimeLoop):
        """TimeLoop whose melt-time maps are kept by the driver, with the
        reference's lines (go_melt.py:93-95, 339-357, 436-456)."""

        def __init__(self, deck):
            super().__init__(deck, quiet=True)
            Levels = self.Levels
            self.accum_time = jnp.zeros(Levels[0]["nn"])
            self.max_accum_time = jnp.zeros(Levels[0]["nn"])

        def accum_single(self, laser_pos, all_reset):
            Levels, Properties = self.Levels, self.Properties
            accum_time, max_accum_time = self.accum_time, self.max_accum_time
            # go_melt.py:339-357
            _resetaccumtime = accum_time[Levels[0]["idx"]] * (all_reset > 0)
            _max_check = jnp.maximum(
                _resetaccumtime, max_accum_time[Levels[0]["idx"]]
            )
            max_accum_time = max_accum_time.at[Levels[0]["idx"]].set(
                _max_check
            )
            accum_time = accum_time.at[Levels[0]["idx"]].add(
                -_resetaccumtime
            )

            accum_time = melting_temp(
                Levels[3]["T0"],
                laser_pos[5],  # Time step size
                Properties["T_liquidus"],
                accum_time,
                Levels[0]["idx"],
            )
            self.accum_time, self.max_accum_time = accum_time, max_accum_time

        def run_subcycle(self, laser_all):
            Levels, Properties = self.Levels, self.Properties
            ne_nn, Shapes, substrate = self.ne_nn, self.Shapes, self.substrate
            LInterp, tmp_ne_nn, subcycle = self.LInterp, self.tmp_ne_nn, self.subcycle
            accum_time, max_accum_time = self.accum_time, self.max_accum_time
            _P = laser_all[:, 6]
            # go_melt.py:437-456
            Levels, L2all, L3all, L3pall, _max_accum, _accum = subcycleGOMELT(
                Levels,
                ne_nn,
                Shapes,
                substrate,
                LInterp,
                tmp_ne_nn,
                laser_all,
                Properties,
                _P,
                subcycle,
                max_accum_time[Levels[0]["idx"]],
                accum_time[Levels[0]["idx"]],
            )
            max_accum_time = max_accum_time.at[Levels[0]["idx"]].set(_max_accum)
            accum_time = accum_time.at[Levels[0]["idx"]].set(_accum)
            self.accum_time, self.max_accum_time = accum_time, max_accum_time

    return ReferenceAccumLoop


def _deck(tmp_path, sub):
    inp, g = W.tiny(n=(8, 8, 4), n2=2, n3=3, track=1.6, props={
        "laser_power": 600.0, "emissivity": 0.4, "evaporation_coefficient": 0.82,
        "thermal_conductivity_bulk_a1": 0.016, "heat_capacity_solid_a1": 0.174,
        "heat_capacity_mushy": 3235.0, "heat_capacity_fluid": 769.0,
        "thermal_conductivity_fluid_a0": 29.0})
    with open(W.write_deck(inp, g, str(tmp_path / sub))) as fh:
        return json.load(fh)


def test_reference_melt_time_lines(oracle, tmp_path):
    from createPath import read_rows
    import go_melt as G
    Loop = _reference_loop_class()
    deck = _deck(tmp_path, "ref")
    run = Loop(deck)
    for rows in run.chunks():
        if not run.step_chunk(rows):
            break
    run.eng.sync()
    assert run.stats["single"] >= 1 and run.stats["subcycle"] >= 1
    acc_ref, macc_ref = np.asarray(run.accum_time), np.asarray(run.max_accum_time)

    # the oracle driver on the same rows
    Lo, info = oracle.go_melt(deck, rows=read_rows(deck["nonmesh"]["toolpath"]))
    assert run.stats["chunks"] == info["chunks"]
    assert np.count_nonzero(info["accum"]) > 0  # the deck melts
    dtmax = float(deck["nonmesh"]["timestep_L3"])
    for a, b in ((acc_ref, info["accum"]), (macc_ref, info["max_accum"])):
        d = np.abs(a.astype(np.float64) - b)
        assert float(np.max(d)) <= 2.5 * dtmax
        assert np.count_nonzero(d > 1e-5 * np.abs(b) + 1e-12) <= 3
    for k in (1, 2, 3):
        a = np.asarray(run.Levels[k]["T0"], np.float64)
        b = Lo[k]["T0"]
        assert float(np.max(np.abs(a - b)) / np.max(b - 300.0)) < 5e-4

    # the engine's own maps (plain facade loop) hold the same values
    deck2 = _deck(tmp_path, "eng")
    Lg = G.go_melt(deck2, quiet=True)
    for a, key in ((acc_ref, "accum_time"), (macc_ref, "max_accum_time")):
        b = np.asarray(Lg[0][key], np.float64)
        assert np.count_nonzero(np.abs(a - b) > 1e-6 * np.abs(b) + 1e-12) <= 3, key


def test_step_returns_reset_mask_and_overlap_indices(tmp_path):
    """stepGOMELT's second return is the reference's _resetmask
    (cF:2394-2397): nodes whose S2 went 0 -> 1 in that step, so it implies S2
    after the step; Levels[0]["idx"] holds the flat Level0 indices of the L3
    window (getOverlapRegion of overlapNodes, cF:2477-2497)."""
    import go_melt as G
    from computeFunctions import getOverlapRegion
    deck = _deck(tmp_path, "mask")
    run = G.TimeLoop(deck, quiet=True)
    L = run.Levels
    seen = []

    def spy(pos, all_reset):
        seen.append((np.asarray(all_reset > 0).copy(), np.asarray(L[3]["S2"]).copy()))

    run.accum_single = spy
    for rows in run.chunks():
        if not run.step_chunk(rows):
            break
    assert len(seen) >= 2
    n_new = 0
    for m, s2 in seen:
        assert m.dtype == bool and m.size == L[3]["nn"]
        assert not np.any(m & (s2 == 0))
        n_new += int(np.count_nonzero(m))
    assert n_new > 0  # the deck melts during its single steps
    box = L[0]["overlapNodes"]
    assert all(len(box[d]) == L[3]["nodes"][d] for d in range(3))
    idx = L[0]["idx"]
    assert np.array_equal(idx, getOverlapRegion(box, L[0]["nodes"][0], L[0]["nodes"][1]))
    assert idx.min() >= 0 and idx.max() < L[0]["nn"]
    # Level0 S1 at idx equals L3 S1 after the last step (cF:2390, 3628)
    assert np.array_equal(np.asarray(L[0]["S1"])[idx], np.asarray(L[3]["S1"]))
