"""Shot batches (born_batch_*): shots with their own windowed domains in one
launch per step, the time loop captured in a HIP graph.

Every shot of a batch must give the traces of the same shot run alone
(born_rtm_forward on its own context); graph replays must repeat the first
run exactly, with other GPU work in between."""
import numpy as np
import pytest

from cases import hip_prop, make_case

pytestmark = pytest.mark.gpu

SHAPES = [(64, 48, 40), (80, 40, 44), (48, 56, 36)]


def _cases(shapes, nt=90):
    return [make_case(*sh, nt=nt, nbound=10, seed=i, src_z=8 + i) for i, sh in enumerate(shapes)]


def _shot(c):
    return {"nx": c.nx, "ny": c.ny, "nz": c.nz, "vel": c.V2, "model": c.model,
            "src_locs": c.src_locs, "src_vals": c.src_vals, "rec_nx": c.rec_nx,
            "rec_ny": c.rec_ny, "rec_locs": c.rec_locs}


def _single(pyborn, c, storage, scatter=0):
    H = hip_prop(pyborn, c)
    H.set_storage(storage)
    rec = np.zeros((c.nrec, c.ntb), np.float32)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, rec)
    H.forward_scatter(c.jt, c.model, rec, c.nt, scatter)
    H.close()
    return rec


def _batch(pyborn, cases, storage):
    H = hip_prop(pyborn, cases[0])  # sinc tables and the 10-cell sponge
    H.set_storage(storage)
    B = pyborn.Batch(H, [_shot(c) for c in cases], cases[0].d)
    return H, B


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("scatter", [0, 1])
def test_batch_matches_single_shots(pyborn, storage, scatter):
    cases = _cases(SHAPES)
    H, B = _batch(pyborn, cases, storage)
    B.forward(1, cases[0].nt, scatter)
    got = [B.traces(s) for s in range(len(cases))]
    print(storage, scatter, B.tile(scatter))
    B.close()
    H.close()
    for c, g in zip(cases, got):
        want = _single(pyborn, c, storage, scatter)
        assert np.abs(want).max() > 0
        err = np.linalg.norm(g - want) / np.linalg.norm(want)
        print("shot %dx%dx%d rel %.2e equal %s" % (c.nx, c.ny, c.nz, err, np.array_equal(g, want)))
        assert err <= 1e-6


def test_batch_with_odd_nx_uses_single_cell_lanes(pyborn):
    cases = _cases([(37, 29, 23), (64, 48, 40)], nt=60)
    H, B = _batch(pyborn, cases, "f32")
    B.forward(1, cases[0].nt, 0)
    assert B.tile(0)["tile"][0] == 64
    got = [B.traces(s) for s in range(len(cases))]
    B.close()
    H.close()
    for c, g in zip(cases, got):
        want = _single(pyborn, c, "f32")
        assert np.linalg.norm(g - want) <= 1e-6 * np.linalg.norm(want)


def test_graph_replays_repeat_the_first_run(pyborn):
    torch = pytest.importorskip("torch")
    cases = _cases(SHAPES, nt=70)
    H, B = _batch(pyborn, cases, "f16")
    runs = []
    for rep in range(3):
        B.forward(1, cases[0].nt, 0)
        runs.append([B.traces(s) for s in range(len(cases))])
        # unrelated eager work between the replays
        x = torch.ones(1 << 20, device="cuda:0")
        torch.cuda.synchronize()
        assert float(x.sum()) == float(1 << 20)
    B.close()
    H.close()
    for rep in (1, 2):
        for a, b in zip(runs[0], runs[rep]):
            assert np.array_equal(a, b)


def test_batch_argument_errors(pyborn):
    cases = _cases(SHAPES[:2], nt=40)
    H = hip_prop(pyborn, cases[0])
    shots = [_shot(c) for c in cases]
    bad = dict(shots[1])
    bad["src_vals"] = np.zeros((9, cases[1].ntb + 1), np.float32)  # another trace length
    with pytest.raises(pyborn.BornError, match="ARG"):
        pyborn.Batch(H, [shots[0], bad], cases[0].d)
    bad = dict(shots[1])
    bad["rec_locs"] = np.zeros_like(cases[1].rec_locs)  # cell 0 is in the shell
    with pytest.raises(pyborn.BornError, match="UNSUPPORTED"):
        pyborn.Batch(H, [shots[0], bad], cases[0].d)
    H.close()


def _single_pair(pyborn, c, storage, scatter):
    """single-shot forward, then the exact adjoint of its own traces"""
    H = hip_prop(pyborn, c)
    H.set_storage(storage)
    rec = np.zeros((c.nrec, c.ntb), np.float32)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, rec)
    H.forward_scatter(c.jt, c.model, rec, c.nt, scatter)
    img = np.zeros(c.nx * c.ny * c.nz, np.float32)
    H.adjoint_exact(c.jt, rec, img, c.nt, scatter)
    H.close()
    return rec, img


@pytest.mark.parametrize("storage", ["f32", "f16", "f24"])
@pytest.mark.parametrize("scatter", [0, 1])
def test_batch_pair_matches_single_shot_pairs(pyborn, storage, scatter):
    """born_batch_forward then born_batch_adjoint_exact: every shot's image
    is the single-shot exact adjoint of its traces, bit for bit"""
    cases = _cases(SHAPES)
    H, B = _batch(pyborn, cases, storage)
    B.forward(1, cases[0].nt, scatter)
    B.adjoint_exact(1, cases[0].nt, scatter)
    imgs = [B.image(s) for s in range(len(cases))]
    assert B.last_checkpoint_plan() == (1, 0)
    B.close()
    H.close()
    for c, g in zip(cases, imgs):
        _, want = _single_pair(pyborn, c, storage, scatter)
        assert np.abs(want).max() > 0
        print("%s scatter %d shot %dx%dx%d equal %s" % (storage, scatter, c.nx, c.ny, c.nz,
                                                        np.array_equal(g, want)))
        assert np.array_equal(g, want)


@pytest.mark.parametrize("storage", ["f32", "f24"])
def test_batch_adjoint_checkpoint_plans_agree(pyborn, storage):
    cases = _cases(SHAPES, nt=70)
    H, B = _batch(pyborn, cases, storage)
    rng = np.random.default_rng(8)
    for s, c in enumerate(cases):
        B.set_traces(s, rng.standard_normal((c.nrec, c.ntb)).astype(np.float32))
    B.adjoint_exact(1, 70, 0)
    full = [B.image(s) for s in range(len(cases))]
    assert B.last_checkpoint_plan() == (1, 0)
    esz = {"f32": 4, "f24": 3}[storage]
    # the batch's level: every shot's cells, each shot padded to 64 cells
    level = sum(-(-c.nx * c.ny * c.nz // 64) * 64 for c in cases) * esz
    for budget_levels, expect in [(40, 2), (25, 4)]:
        H.set_checkpoint_budget(budget_levels * level)
        B.adjoint_exact(1, 70, 0)
        segs, rec = B.last_checkpoint_plan()
        L = -(-70 // segs)
        assert segs == expect and rec == (segs - 1) * L
        for s in range(len(cases)):
            assert np.array_equal(B.image(s), full[s])
    B.close()
    H.close()
    # the set traces give the single-shot exact adjoint of those traces
    rng = np.random.default_rng(8)
    for s, c in enumerate(cases):
        w = rng.standard_normal((c.nrec, c.ntb)).astype(np.float32)
        S = hip_prop(pyborn, c)
        S.set_storage(storage)
        S.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, np.zeros_like(w))
        img = np.zeros(c.nx * c.ny * c.nz, np.float32)
        S.adjoint_exact(1, w, img, 70, 0)
        S.close()
        assert np.array_equal(full[s], img)
