"""24-bit wavefield storage (born_set_storage(BORN_STORAGE_F24)) on the GPU:
every stored level holds its FP32 values rounded to 24 bits (16 significant
bits, the FP32 exponent, no scale); the arithmetic stays FP32.

  * data and exact-adjoint image against the FP64 oracle: rel-L2 <= 1e-3,
    the accuracy bar for the 16-bit-based storage, where FP16 levels alone
    miss it (their rounding noise grows with the step count);
  * checkpointed background recompute from 24-bit checkpoints: every plan
    gives bitwise the same image;
  * the history kept by the forward gives the image of the recomputed one;
  * dot-product mismatch of the pair with white-noise inputs;
  * every shot of a batch gives the traces of the same shot run alone.
"""
import numpy as np
import pytest

from cases import hip_prop, make_case, oracle_prop, rel_l2

pytestmark = pytest.mark.gpu

F24_ERR = 1e-3


def _hip(pyborn, c, storage="f24"):
    H = hip_prop(pyborn, c)
    H.set_storage(storage)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs,
                             np.zeros((c.nrec, c.ntb), np.float32))
    return H


@pytest.mark.parametrize("scatter", [0, 1])
def test_f24_data_and_image_against_fp64(pyborn, scatter):
    c = make_case(64, 56, 48, nt=200, nbound=12)
    rng = np.random.default_rng(21)
    w = rng.standard_normal((c.nrec, c.ntb)).astype(np.float32)
    P = oracle_prop(c)
    P.transfer_receiver_func(c.rec_locs, np.zeros((c.nrec, c.ntb), np.float32))
    d64 = P.born_forward_d(c.jt, c.model, c.nt, scatter)
    g64 = P.exact_adjoint_d(c.jt, w.astype(np.float64), c.nt, scatter=scatter).ravel()
    err = {}
    for storage in ("f16", "f24"):
        H = _hip(pyborn, c, storage)
        d = np.zeros((c.nrec, c.ntb), np.float32)
        H.forward_scatter(c.jt, c.model, d, c.nt, scatter)
        g = np.zeros(c.nx * c.ny * c.nz, np.float32)
        H.adjoint_exact(c.jt, w, g, c.nt, scatter=scatter)
        H.close()
        err[storage] = (rel_l2(d, d64), rel_l2(g, g64))
    print("scatter %d: f16 data %.2e image %.2e | f24 data %.2e image %.2e"
          % (scatter, *err["f16"], *err["f24"]))
    assert max(err["f24"]) <= F24_ERR
    # 16 significant bits take the error well below FP16's 11
    assert max(err["f24"]) < 0.1 * max(err["f16"])


def test_f24_checkpoint_plans_agree(pyborn):
    c = make_case(48, 40, 32, nt=70, nbound=8)
    rng = np.random.default_rng(3)
    w = rng.standard_normal((c.nrec, c.ntb)).astype(np.float32)
    H = _hip(pyborn, c)
    full = np.zeros(c.nx * c.ny * c.nz, np.float32)
    H.adjoint_exact(c.jt, w, full, c.nt)
    assert H.last_checkpoint_plan() == (1, 0)
    F = c.nx * c.ny * c.nz * 3  # 24-bit levels
    for budget_fields, expect_segments in [(40, 2), (25, 4)]:
        H.set_checkpoint_budget(budget_fields * F)
        g = np.zeros_like(full)
        H.adjoint_exact(c.jt, w, g, c.nt)
        segs, recomputed = H.last_checkpoint_plan()
        L = -(-c.nt // segs)
        assert segs == expect_segments and recomputed == (segs - 1) * L
        assert np.array_equal(g, full)
    H.close()


def test_f24_history_from_forward_equals_recomputed(pyborn):
    torch = pytest.importorskip("torch")
    c = make_case(48, 40, 32, nt=80, nbound=8)
    dev = torch.device("cuda:0")
    model = torch.from_numpy(c.model.ravel().copy()).to(dev)
    rng = np.random.default_rng(5)
    w = torch.from_numpy(rng.standard_normal((c.ntb, c.nrec)).astype(np.float32)).to(dev)
    imgs = []
    for keep in (True, False):
        H = _hip(pyborn, c)
        tr = torch.zeros((c.ntb, c.nrec), dtype=torch.float32, device=dev)
        img = torch.zeros(c.nx * c.ny * c.nz, dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        H.dev_forward(c.jt, c.nt, model, tr, scatter=0, keep_background=keep)
        H.dev_adjoint_exact(c.jt, c.nt, w, img, scatter=0)
        H.synchronize()
        imgs.append(img.cpu().numpy())
        H.close()
    assert np.abs(imgs[0]).max() > 0
    assert np.array_equal(imgs[0], imgs[1])


def test_f24_dot_product(pyborn):
    c = make_case(64, 56, 48, nt=150, nbound=12)
    rng = np.random.default_rng(11)
    m = (rng.random(c.model.shape, dtype=np.float32) - 0.5).astype(np.float32)
    d = (rng.random((c.nrec, c.ntb), dtype=np.float32) - 0.5).astype(np.float32)
    H = _hip(pyborn, c)
    fm = np.zeros((c.nrec, c.ntb), np.float32)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, fm)
    H.forward_scatter(c.jt, m, fm, c.nt, 1)
    ftd = np.zeros(c.nx * c.ny * c.nz, np.float32)
    H.adjoint_exact(c.jt, d, ftd, c.nt, scatter=1)
    H.close()
    a = float(np.dot(fm.ravel().astype(np.float64), d.ravel().astype(np.float64)))
    b = float(np.dot(m.ravel().astype(np.float64), ftd.astype(np.float64)))
    mis = abs(a - b) / abs(a)
    print("f24 dot mismatch %.2e" % mis)
    assert mis <= 1e-3


def _shot(c):
    return {"nx": c.nx, "ny": c.ny, "nz": c.nz, "vel": c.V2, "model": c.model,
            "src_locs": c.src_locs, "src_vals": c.src_vals, "rec_nx": c.rec_nx,
            "rec_ny": c.rec_ny, "rec_locs": c.rec_locs}


@pytest.mark.parametrize("scatter", [0, 1])
def test_f24_batch_matches_single_shots(pyborn, scatter):
    # the middle shot's cell count is not a multiple of 64 (odd nx)
    shapes = [(64, 48, 40), (81, 40, 44), (48, 56, 36)]
    cases = [make_case(*sh, nt=90, nbound=10, seed=i, src_z=8 + i) for i, sh in enumerate(shapes)]
    H = hip_prop(pyborn, cases[0])
    H.set_storage("f24")
    B = pyborn.Batch(H, [_shot(c) for c in cases], cases[0].d)
    B.forward(1, cases[0].nt, scatter)
    got = [B.traces(s) for s in range(len(cases))]
    B.close()
    H.close()
    for c, g in zip(cases, got):
        S = _hip(pyborn, c)
        want = np.zeros((c.nrec, c.ntb), np.float32)
        S.forward_scatter(c.jt, c.model, want, c.nt, scatter)
        S.close()
        assert np.abs(want).max() > 0
        assert np.array_equal(g, want)
