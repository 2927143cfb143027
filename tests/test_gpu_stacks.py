"""The batched configuration as three-level stacks (batch.Stacks): each
instance is its own L1/L2/L3 run (the reference time loop gm:134-493 with
subcycleGOMELT cF:3224-3632 per instance); the runs advance together and the
L3 solves of all of them go out as one batched launch per solve.

Every instance is compared with its own unbatched run (one engine, its own
stream, the recorded subcycle graph) over the same chunks: the batched step
runs the same per-node arithmetic on the same inputs, so the fields agree to
FP32 rounding (the restriction sums add FP64 atomics in no fixed order, and
the autotuned brick shape of the batched solves groups a node's sums
as its own kernel's shape does only for 1 x 3 bricks)."""
import json

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _decks(tmp_path):
    """Eight instances: two of each L3 size of the batched configuration
    (32^3, 48x48x24, 64x64x32, 96x96x48 elements; L2 and L1 spanning 4x and
    16x the L3 extent), P, speed and radius drawn from their seeds."""
    out = []
    for i, size in enumerate(W.BATCH_SIZES):
        for j in range(2):
            inp, g = W.stack_decks(n=1, seed=10 * i + j, sizes=[size])[0]
            d = tmp_path / f"s{i}_{j}"
            with open(W.write_deck(inp, g, str(d))) as fh:
                out.append(json.load(fh))
    return out


def _alone(deck, chunks):
    import go_melt as G
    run = G.TimeLoop(deck, quiet=True, write_outputs=False)
    it = run.chunks()
    for _ in range(chunks):
        run.step_chunk(next(it))
    run.eng.sync()
    run.eng.sync_coords()
    return run


@pytest.mark.parametrize("batch_l2,batch_restrict", [(True, True), (True, False), (False, True)])
def test_stacks_equal_their_own_runs(tmp_path, monkeypatch, batch_l2, batch_restrict):
    """batch_l2: the restrictions, L2 solves, face values and T' + injection
    batched across the stacks too, or only the L3 solves; batch_restrict
    False: each stack restricts on its own stream inside the batched
    substep (the path of a stack whose restriction the batched launch does
    not run)."""
    import batch as B
    monkeypatch.delenv("GOMELT_NO_GRAPH", raising=False)
    decks = _decks(tmp_path)
    chunks = 4  # one single-step chunk, then three subcycled L1 steps
    st = B.Stacks(decks, batch_l2=batch_l2, batch_restrict=batch_restrict)
    for _ in range(chunks):
        assert st.step()
    st.sync()
    worst = 0.0
    bitwise = 0
    for i, deck in enumerate(decks):
        run = st.runs[i]
        ref = _alone(deck, chunks)
        assert run.stats["chunks"] == ref.stats["chunks"], (i, run.stats, ref.stats)
        assert run.stats["chunks"].count("subcycle") >= 2, (i, run.stats)
        for k in (1, 2, 3):
            a = np.asarray(run.Levels[k]["T0"], np.float64)
            b = np.asarray(ref.Levels[k]["T0"], np.float64)
            assert np.all(np.isfinite(a)) and float(np.max(b)) > 300.5, (i, k)
            err = float(np.max(np.abs(a - b)) / (float(np.max(b)) - 300.0))
            worst = max(worst, err)
            bitwise += int(np.array_equal(a, b))
            assert err < 1e-6, (i, k, err)
        for key in ("accum_time", "max_accum_time"):
            a = np.asarray(run.Levels[0][key], np.float64)
            b = np.asarray(ref.Levels[0][key], np.float64)
            assert float(np.max(np.abs(a - b))) <= 1e-9 + 1e-6 * float(np.max(b)), (i, key)
        del ref
    print(f"worst relative difference {worst:.3g}; {bitwise} of {3 * len(decks)} fields bitwise")


def test_stacks_one_after_another_across_streams(tmp_path):
    """Stacks.step_each (the sequential baseline: each run's chunk waits on
    the GPU for the run before it, across streams) ends every run where its
    own run ends."""
    import batch as B
    decks = []
    for j in range(3):
        inp, g = W.stack_decks(n=1, seed=40 + j, sizes=[W.BATCH_SIZES[0]])[0]
        with open(W.write_deck(inp, g, str(tmp_path / f"e{j}"))) as fh:
            decks.append(json.load(fh))
    st = B.Stacks(decks, streams=2)
    assert len({id(r.eng.stream) for r in st.runs}) == 2
    st.step()  # the single-step chunk
    for _ in range(2):
        assert st.step_each()
    st.sync()
    for i, deck in enumerate(decks):
        run = st.runs[i]
        ref = _alone(deck, 3)
        assert run.stats["chunks"] == ref.stats["chunks"], (i, run.stats, ref.stats)
        for k in (1, 2, 3):
            a = np.asarray(run.Levels[k]["T0"], np.float64)
            b = np.asarray(ref.Levels[k]["T0"], np.float64)
            assert float(np.max(np.abs(a - b))) <= 1e-6 * (float(np.max(b)) - 300.0), (i, k)
