"""The subcycled L1 step stage by stage, and the learned operator inside the
engine's moving L3 window (subcycleGOMELT cF:3224-3632 with the N3 L3 solves
of each L2 substep, cF:3367-3412 and cF:3530-3590, replaced by one jump).

* With the plain solves at every stage, the staged step equals the recorded
  graph (same kernels, same order).
* An untrained operator fails the energy audit on every jump: every substep
  falls back to the plain solves and the run ends on the plain fields.
* With no audit limit the jumps are kept: the fields stay finite and above
  T_amb, and the judge compares jumps with the plain solves at its interval.
* The recorder (no network) sees every substep's plain input and output.
"""
import json
import os

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _deck(tmp_path, track=1.6, n3=3):
    """L1 16 x 16 x 8 elements (0.5 mm), L2 16 x 16 x 8 (0.25 mm) and an L3 of
    16 x 16 x 8 nodes (15 x 15 x 7 elements, 0.125 mm: even node counts for
    the network's pooling), every window on the lattice of the next coarser
    level; dt3 = 0.25 ms keeps the explicit step stable."""
    inp, g = W.tiny(n=(16, 16, 8), n2=2, n3=n3, track=track, dt3=2.5e-4,
                    props={"laser_power": 600.0, "emissivity": 0.4})
    inp["Level3"] = {"elements": [15, 15, 7],
                     "bounds": {"x": [3.125, 5.0], "y": [3.125, 5.0], "z": [-0.875, 0.0]}}
    with open(W.write_deck(inp, g, str(tmp_path))) as fh:
        return json.load(fh)


def _run(deck, operator=None, graph=True, chunks=None):
    import go_melt as G
    import computeFunctions as cf
    os.environ["GOMELT_NO_GRAPH"] = "" if graph else "1"
    run = G.TimeLoop(deck, quiet=True, write_outputs=False)
    cf.attach_l3_operator(run.Levels, operator)
    done = 0
    for rows in run.chunks():
        if chunks is not None and done >= chunks:
            break
        if not run.step_chunk(rows):
            break
        done += 1
    run.eng.sync()
    return run


def _fields(run):
    out = {k: np.asarray(run.Levels[k]["T0"], np.float64) for k in (1, 2, 3)}
    out["Tp3"] = np.asarray(run.Levels[3]["Tprime0"], np.float64)
    out["acc"] = np.asarray(run.Levels[0]["accum_time"], np.float64)
    return out


def test_staged_plain_equals_graph(tmp_path):
    import surrogate as S
    deck = _deck(tmp_path)
    a = _fields(_run(deck))
    op = S.EngineOperator(None, None)
    run = _run(deck, operator=op)
    assert run.stats["subcycle"] >= 2
    b = _fields(run)
    print("plain field minima", {k: float(np.min(a[k])) for k in (1, 2, 3)},
          "where L1 < T_amb:", np.flatnonzero(a[1] < 299.0)[:20])
    for key in a:
        scale = max(float(np.max(np.abs(a[key]))), 1e-30)
        assert float(np.max(np.abs(a[key] - b[key]))) <= 1e-6 * scale, key


def _net(seed=0):
    import surrogate as S
    return S.UNet(channels=(16, 32), seed=seed)


# N3 = 4: the Part-1 solves of a substep end in the buffer they started from,
# so the operator's field goes to a third buffer and the fallback still starts
# from the substep's start field
@pytest.mark.parametrize("n3", [3, 4])
def test_untrained_operator_falls_back_everywhere(tmp_path, n3):
    import surrogate as S
    import computeFunctions as cf
    deck = _deck(tmp_path, n3=n3)
    plain = _fields(_run(deck))
    p = cf.SetupProperties(deck["properties"])
    op = S.EngineOperator(_net(), p, residual_limit=0.01, judge_every=0)
    run = _run(deck, operator=op)
    st = op.stats
    print("jumps", st["jumps"], "fallbacks", st["fallbacks"], "residuals",
          np.round(np.abs(st["residuals"]), 3))
    assert st["jumps"] == 2 * 2 * run.stats["subcycle"] and st["jumps"] > 0
    assert st["fallbacks"] == st["jumps"]
    # the untrained jumps fail the audit by far; the plain solves pass it
    assert min(abs(r) for r in st["residuals"]) > 0.05
    got = _fields(run)
    for key in plain:
        scale = max(float(np.max(np.abs(plain[key]))), 1e-30)
        assert float(np.max(np.abs(got[key] - plain[key]))) <= 1e-6 * scale, key


def test_operator_jumps_kept_without_limit(tmp_path):
    import surrogate as S
    import computeFunctions as cf
    deck = _deck(tmp_path)
    p = cf.SetupProperties(deck["properties"])
    op = S.EngineOperator(_net(1), p, residual_limit=float("inf"), judge_every=3)
    run = _run(deck, operator=op)
    st = op.stats
    assert st["jumps"] > 0 and st["fallbacks"] == 0
    assert len(st["judged"]) == st["jumps"] // 3
    for j in st["judged"]:
        assert np.isfinite(j["residual"]) and np.isfinite(j["peak_T_rel_err"])
    f = _fields(run)
    print("operator field minima", {k: float(np.min(f[k])) for k in (1, 2, 3)},
          "where L1 < T_amb:", np.flatnonzero(f[1] < 299.0)[:20])
    for k in (1, 2, 3):
        assert np.all(np.isfinite(f[k])) and float(np.min(f[k])) >= 300.0 - 1e-3, k


def test_forward_graph_replays_equal_eager(tmp_path):
    """The operator's forward pass recorded as a graph and replayed on every
    jump (with the engine's stage work and the audits between the replays)
    gives the run the eager forward gives: every jump kept, the residuals and
    the final fields equal up to the order of the FP64 atomic sums in the
    restrictions and the audit."""
    import surrogate as S
    import computeFunctions as cf
    deck = _deck(tmp_path)
    p = cf.SetupProperties(deck["properties"])
    out = []
    for graph in (False, True):
        op = S.EngineOperator(_net(1), p, residual_limit=float("inf"), judge_every=0,
                              graph=graph)
        run = _run(deck, operator=op)
        assert op.stats["jumps"] >= 4 and op.stats["fallbacks"] == 0
        assert (op._g is not None) == graph
        out.append((_fields(run), list(op.stats["residuals"])))
    (fa, ra), (fb, rb) = out
    np.testing.assert_allclose(ra, rb, rtol=1e-9, atol=1e-12)
    for key in fa:
        scale = max(float(np.max(np.abs(fa[key]))), 1e-30)
        assert float(np.max(np.abs(fa[key] - fb[key]))) <= 1e-6 * scale, key
    print("bitwise:", {k: bool(np.array_equal(fa[k], fb[k])) for k in fa})


def test_recorder_sees_every_substep(tmp_path):
    import surrogate as S
    deck = _deck(tmp_path)
    seen = []

    def rec(T0, face, rows, T1):
        seen.append((T0.numel(), face.numel(), rows.shape, T1.numel()))

    op = S.EngineOperator(None, None, record=rec)
    run = _run(deck, operator=op)
    n3 = int(run.Levels[3]["nn"])
    assert len(seen) == 2 * 2 * run.stats["subcycle"]
    assert all(s == (n3, n3, (3, 7), n3) for s in seen)


def test_jump_audit_of_plain_solves(tmp_path):
    """The audit of every judged substep's plain solves (N3 = 4, both parts):
    the energy identity holds for them to well under the 1% limit, so the
    trigger separates jumps by their own error."""
    import surrogate as S
    import computeFunctions as cf
    deck = _deck(tmp_path, n3=4)
    p = cf.SetupProperties(deck["properties"])
    op = S.EngineOperator(_net(2), p, residual_limit=float("inf"), judge_every=1)
    _run(deck, operator=op)
    plain = [abs(j["residual_plain"]) for j in op.stats["judged"]]
    print("plain-solve residuals", np.round(plain, 5))
    assert len(plain) == op.stats["jumps"] > 0
    assert max(plain) < 0.01


def test_engine_windows_train_and_run_in_the_loop(tmp_path):
    """Training windows from the plain engine (64 x 64 x 32-node L3 windows of
    seeded runs), a short training that beats the persistence predictor on
    windows of a held-out run, and the trained operator inside the engine on
    a held-out deck: every jump is audited; accepted jumps keep the fields
    finite and above T_amb."""
    import computeFunctions as cf
    import surrogate as S
    X, Y, seeds = S.engine_windows(64, seed0=0, n_l1=(60, 60, 30), length=1.0, per_run=16)
    assert tuple(X.shape[1:]) == (32, 64, 64, 16) and Y.shape[1] == 64 * 64 * 32
    seeds = np.asarray(seeds)
    assert X.shape[0] == 64 and len(set(seeds.tolist())) == 4
    held = seeds == seeds[-1]
    ntr = int(np.count_nonzero(~held))
    net = S.UNet(channels=(16, 32, 32), seed=0)
    e0 = S.rel_error(net, X[ntr:], Y[ntr:])
    hist = S.train(net, X[:ntr], Y[:ntr], iters=300, batch=4, lr=2e-3)
    err = S.rel_error(net, X[ntr:], Y[ntr:])
    print("held-out relative error of dT: untrained", e0, "trained", err, "loss", hist[0],
          "->", hist[-1])
    assert hist[-1] < 0.5 * hist[0]
    assert err < 1.0
    inp, g = W.window_deck(101, n_l1=(60, 60, 30), length=1.0)
    with open(W.write_deck(inp, g, str(tmp_path))) as fh:
        deck = json.load(fh)
    p = cf.SetupProperties(deck["properties"])
    op = S.EngineOperator(net, p, residual_limit=0.05, judge_every=5)
    run = _run(deck, operator=op, chunks=4)
    st = op.stats
    print("jumps", st["jumps"], "fallbacks", st["fallbacks"], "judged", st["judged"])
    assert st["jumps"] > 0 and len(st["residuals"]) == st["jumps"]
    f = _fields(run)
    for k in (1, 2, 3):
        assert np.all(np.isfinite(f[k])) and float(np.min(f[k])) >= 300.0 - 1e-3, k


def test_staged_and_batched_calls_refuse_misuse(tmp_path):
    """Host-side checks of the staged subcycle and the batched solves: stages
    out of order, the recorded subcycle during an open staged one, and
    batched solves over engines that are not at the same substep, or listed
    twice, are refused with an error (no kernel runs for them).  The checks
    run inside a subcycled L1 step of the driver, whose stages then complete
    as the plain ones."""
    import ctypes as C
    import _native as N
    import go_melt as G
    import computeFunctions as cf
    deck = _deck(tmp_path)
    other = G.TimeLoop(deck, quiet=True, write_outputs=False)
    e1 = other.eng
    seen = []

    class Probe:
        def subcycle(self, e, rows):
            lib = N.lib()
            rows = np.ascontiguousarray(rows, dtype=np.float32)
            with pytest.raises(N.GoMeltNativeError, match="out of order"):
                N.check(lib.gm_engine_stage_l2(e.h, None, None))
            N.check(lib.gm_engine_stage_begin(e.h, N.fptr(rows), rows.shape[0]))
            with pytest.raises(N.GoMeltNativeError):
                N.check(lib.gm_engine_subcycle(e.h, N.fptr(rows), rows.shape[0]))
            h = C.c_void_p()
            arr = (C.c_void_p * 2)(e.h, e.h)
            with pytest.raises(N.GoMeltNativeError, match="twice"):
                N.check(lib.gm_multi_create(arr, 2, C.byref(h)))
            arr = (C.c_void_p * 2)(e.h, e1.h)
            N.check(lib.gm_multi_create(arr, 2, C.byref(h)))
            try:
                # the other engine has not begun a staged subcycle
                with pytest.raises(N.GoMeltNativeError, match="just begun"):
                    N.check(lib.gm_multi_subcycle(h.value))
                with pytest.raises(N.GoMeltNativeError, match="same substep"):
                    N.check(lib.gm_multi_l3(h.value))
            finally:
                lib.gm_multi_destroy(h.value)
            for _ in range(2 * e.N2):
                N.check(lib.gm_engine_stage_l2(e.h, None, None))
                N.check(lib.gm_engine_stage_l3(e.h, 0))
            seen.append(rows.shape[0])

    run = G.TimeLoop(deck, quiet=True, write_outputs=False)
    cf.attach_l3_operator(run.Levels, Probe())
    it = run.chunks()
    for _ in range(4):  # until one subcycled L1 step has run
        assert run.step_chunk(next(it))
        if seen:
            break
    run.eng.sync()
    assert seen == [run.eng.N2 * run.eng.N3]
    for k in (1, 2, 3):
        assert np.all(np.isfinite(np.asarray(run.Levels[k]["T0"]))), k
