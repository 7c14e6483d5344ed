"""Placement of the work pair (r of the Born forward, the adjoint state of
the exact adjoint) in the gaps of an FP32 history.

Where the two levels sit changes only the speed: traces and image must be
bitwise those of the plain layout (no gaps, levels in the work buffers),
with the places taken in turn window after window (every move exercised),
with the timed choice, with a kept history and with a checkpoint plan.  The
windows here are a few levels long so that a small grid has many of them.
"""
import numpy as np
import pytest

from cases import hip_prop, make_case

pytestmark = pytest.mark.gpu


def _pair(H, c, torch, model, w, scatter, keep=True):
    dev = torch.device("cuda:0")
    tr = torch.zeros((c.ntb, c.nrec), dtype=torch.float32, device=dev)
    H.dev_forward(c.jt, c.nt, model, tr, scatter=scatter, keep_background=keep)
    img = torch.zeros(c.nx * c.ny * c.nz, dtype=torch.float32, device=dev)
    H.dev_adjoint_exact(c.jt, c.nt, w, img, scatter=scatter)
    H.synchronize()
    return tr.cpu().numpy(), img.cpu().numpy()


@pytest.mark.parametrize("scatter", [0, 1])
def test_pair_placement_leaves_results_unchanged(pyborn, scatter):
    torch = pytest.importorskip("torch")
    c = make_case(48, 40, 32, nt=70, nbound=8, seed=11)
    dev = torch.device("cuda:0")
    model = torch.from_numpy(c.model.ravel().copy()).to(dev)
    w = torch.from_numpy(np.random.default_rng(5).standard_normal((c.ntb, c.nrec))
                         .astype(np.float32)).to(dev)
    H = hip_prop(pyborn, c)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, np.zeros((c.nrec, c.ntb), np.float32))
    F = c.nx * c.ny * c.nz * 4
    H.set_placement("off")
    want_tr, want = _pair(H, c, torch, model, w, scatter)
    assert H.last_placement() == (0, 0)
    assert np.abs(want).max() > 0 and np.abs(want_tr).max() > 0
    # windows of 8 fields: a gap of three (two levels and the adjoint's image
    # accumulator), then 5 levels; 14 windows.  Window w takes place w - 1
    # (the work buffers first), so the pair moves at every window but the
    # forward's first
    H.set_placement("cycle", 8 * F)
    tr, img = _pair(H, c, torch, model, w, scatter)
    nwin = -(-c.nt // 5)
    assert H.last_placement() == (nwin - 1, nwin)
    assert np.array_equal(tr, want_tr) and np.array_equal(img, want)
    # the timed choice (whatever it picks)
    H.set_placement("timed", 8 * F)
    tr, img = _pair(H, c, torch, model, w, scatter)
    assert np.array_equal(tr, want_tr) and np.array_equal(img, want)
    # a checkpoint plan over a gapped segment history: the forward's last
    # segment and the adjoint's recomputed ones
    H.set_placement("cycle", 6 * F)
    H.set_checkpoint_budget(40 * F)
    tr, img = _pair(H, c, torch, model, w, scatter)
    segs, _ = H.last_checkpoint_plan()
    fm, am = H.last_placement()
    assert segs >= 2 and fm >= 2 and am >= 2
    assert np.array_equal(tr, want_tr) and np.array_equal(img, want)
    # an adjoint that sweeps the background itself (the first one used the
    # forward's plan up)
    img = torch.zeros(c.nx * c.ny * c.nz, dtype=torch.float32, device=dev)
    H.dev_adjoint_exact(c.jt, c.nt, w, img, scatter=scatter)
    H.synchronize()
    assert H.last_background_sweep() == c.nt
    assert np.array_equal(img.cpu().numpy(), want)
    # back to the plain layout after gapped ones
    H.set_checkpoint_budget(0)
    H.set_placement("off")
    tr, img = _pair(H, c, torch, model, w, scatter)
    assert H.last_placement() == (0, 0)
    assert np.array_equal(tr, want_tr) and np.array_equal(img, want)
    H.close()


@pytest.mark.parametrize("storage", ["f16", "f24"])
def test_narrow_pair_placement_leaves_results_unchanged(pyborn, storage):
    # narrow storage keeps s and r together in the work pair and writes the
    # scatter operand apart: that pair moves through the gaps instead of r
    torch = pytest.importorskip("torch")
    c = make_case(48, 40, 32, nt=70, nbound=8, seed=12)
    dev = torch.device("cuda:0")
    model = torch.from_numpy(c.model.ravel().copy()).to(dev)
    w = torch.from_numpy(np.random.default_rng(6).standard_normal((c.ntb, c.nrec))
                         .astype(np.float32)).to(dev)
    H = hip_prop(pyborn, c)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, np.zeros((c.nrec, c.ntb), np.float32))
    H.set_storage(storage)
    F = c.nx * c.ny * c.nz * 4
    H.set_placement("off")
    want_tr, want = _pair(H, c, torch, model, w, 0)
    assert H.last_placement() == (0, 0)
    assert np.abs(want).max() > 0 and np.abs(want_tr).max() > 0
    for mode in ("cycle", "timed"):
        H.set_placement(mode, 8 * F)
        tr, img = _pair(H, c, torch, model, w, 0)
        fm, am = H.last_placement()
        if mode == "cycle":
            assert fm >= 2 and am >= 2
        assert np.array_equal(tr, want_tr) and np.array_equal(img, want)
    # an adjoint that rebuilds the background into a gapped history itself
    H.set_checkpoint_budget(40 * F)
    H.set_placement("cycle", 8 * F)
    tr, img = _pair(H, c, torch, model, w, 0)
    segs, _ = H.last_checkpoint_plan()
    assert segs >= 2 and H.last_placement()[1] >= 2
    assert np.array_equal(img, want)
    H.close()


def test_placement_arguments(pyborn):
    c = make_case(24, 24, 24, nt=10, nbound=4)
    H = hip_prop(pyborn, c)
    assert H.lib.born_set_placement(H.h, 3, 0) != 0
    assert H.lib.born_set_placement(H.h, 1, -1) != 0
    H.set_placement("timed", 0)
    H.close()
