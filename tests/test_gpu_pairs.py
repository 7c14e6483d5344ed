"""Two-step forward launches (pair_kernel, born_pair.h) against single steps.

The pair kernel computes level k+1 on a tile widened by the stencil radius
and level k+2 on the tile, from the same per-cell arithmetic as the
single-step kernel (every rounding fixed, no compiler-chosen fusion), so
the Born data (and the background history it keeps for the exact adjoint)
must equal those of single-step launches bit for bit.  The cases
put sources and receivers across tile and z-chunk boundaries, use domains
that are not multiples of the tile, odd and even step counts, a time-sinc
receiver sampling (jt = 3) and every pair tile shape; the first test also
pins the pair path to the CPU oracle."""
import numpy as np
import pytest

from cases import hip_prop, make_case, oracle_prop, rel_l2

pytestmark = pytest.mark.gpu

SHAPES = [(1, 8), (2, 8), (2, 4), (4, 4)]  # rows per thread, wavefronts


def _forward(pyborn, c, scatter, pairs, shape=None, zc=None):
    H = hip_prop(pyborn, c)
    H.set_step_pairs(pairs)
    if shape is not None:
        kern = "fwd_pair" if scatter == 0 else "fwd_dV_pair"
        H.set_tuning(kern, shape[0], shape[1], 1, 1, zc)
    elif zc is not None:
        H.set_zchunk(zc)
    rec = np.zeros((c.nrec, c.ntb), np.float32)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, rec)
    H.forward_scatter(c.jt, c.model, rec, c.nt, scatter)
    used = H.last_step_pairs
    H.close()
    return rec, used


def _check(got, want, what):
    assert np.abs(want).max() > 0
    err = rel_l2(got, want)
    print("%s: rel %.2e, bitwise equal %s" % (what, err, np.array_equal(got, want)))
    assert np.array_equal(got, want), what


@pytest.mark.parametrize("scatter", [0, 1])
def test_pairs_match_single_steps_and_oracle(pyborn, scatter):
    c = make_case(120, 40, 44, nt=91, nbound=10, src_z=12)
    got, used = _forward(pyborn, c, scatter, True)
    assert used
    want, used1 = _forward(pyborn, c, scatter, False)
    assert not used1
    _check(got, want, "120x40x44 scatter %d" % scatter)
    P = oracle_prop(c)
    P.transfer_receiver_func(c.rec_locs, np.zeros((c.nrec, c.ntb), np.float32))
    ref = P.born_forward_d(c.jt, c.model, c.nt, scatter=scatter)
    assert rel_l2(got, ref) <= 1e-4


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("zc", [8, 13, 200])
def test_every_pair_shape_and_z_chunk(pyborn, shape, zc):
    # source stamp across the x tile boundary at 60 and the y boundaries;
    # chunks of 8 and 13 planes put chunk edges next to source and receivers
    c = make_case(121, 43, 37, nt=40, nbound=8, src_z=9, rec_z=7, seed=3)
    want, _ = _forward(pyborn, c, 0, False)
    got, used = _forward(pyborn, c, 0, True, shape=shape, zc=zc)
    assert used
    _check(got, want, "shape %s zc %d" % (shape, zc))


@pytest.mark.parametrize("nt", [5, 6, 50])
def test_step_counts_and_time_sinc(pyborn, nt):
    # receivers on the source plane and a model without the quiet top layer,
    # so that five steps already give data (odd counts end on a single step)
    c = make_case(70, 36, 30, nt=nt, nbound=6, jt=3, src_z=12, rec_z=12, rec_step=1)
    c.model = np.random.default_rng(2).standard_normal(c.model.shape).astype(np.float32)
    want, _ = _forward(pyborn, c, 1, False)
    got, used = _forward(pyborn, c, 1, True)
    assert used
    _check(got, want, "nt %d jt 3" % nt)


def test_small_domain_smaller_than_a_tile(pyborn):
    c = make_case(20, 19, 18, nt=30, nbound=3, src_z=8, rec_z=6, rec_step=1)
    want, _ = _forward(pyborn, c, 0, False)
    got, _ = _forward(pyborn, c, 0, True)
    _check(got, want, "20x19x18")


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("scatter", [0, 1])
def test_kept_history_leaves_the_data_unchanged(pyborn, storage, scatter):
    # with scatter 0 and a kept history, the s levels are written into the
    # history instead of a separate s buffer
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    c = make_case(70, 44, 40, nt=61, nbound=8)
    model = torch.from_numpy(c.model.ravel()).to(dev)
    out = []
    for keep in (False, True):
        H = hip_prop(pyborn, c)
        H.set_storage(storage)
        H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs,
                                 np.zeros((c.nrec, c.ntb), np.float32))
        tr = torch.zeros((c.ntb, c.nrec), dtype=torch.float32, device=dev)
        H.dev_forward(c.jt, c.nt, model, tr, scatter=scatter, keep_background=keep)
        H.synchronize()
        out.append(tr.cpu().numpy())
        H.close()
    _check(out[1], out[0], "kept history, %s scatter %d" % (storage, scatter))


@pytest.mark.parametrize("scatter", [0, 1])
def test_kept_history_gives_the_same_exact_adjoint(pyborn, scatter):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    c = make_case(90, 44, 40, nt=61, nbound=8)
    rng = np.random.default_rng(5)
    w = torch.from_numpy(rng.standard_normal((c.ntb, c.nrec)).astype(np.float32)).to(dev)
    model = torch.from_numpy(c.model.ravel()).to(dev)
    imgs = []
    for pairs in (True, False):
        H = hip_prop(pyborn, c)
        H.set_step_pairs(pairs)
        H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs,
                                 np.zeros((c.nrec, c.ntb), np.float32))
        tr = torch.zeros((c.ntb, c.nrec), dtype=torch.float32, device=dev)
        H.dev_forward(c.jt, c.nt, model, tr, scatter=scatter, keep_background=True)
        assert H.last_step_pairs == pairs
        img = torch.zeros(c.nx * c.ny * c.nz, dtype=torch.float32, device=dev)
        H.dev_adjoint_exact(c.jt, c.nt, w, img, scatter=scatter)
        H.synchronize()
        segs, rec = H.last_checkpoint_plan()
        assert segs == 1 and rec == 0  # the adjoint used the kept history
        imgs.append(img.cpu().numpy())
        H.close()
    _check(imgs[0], imgs[1], "image from kept history, scatter %d" % scatter)


# ---- adjoint-state sweep of the exact adjoint in two-step launches
# (lam_pair_kernel, born_pair_adj.h)

def _adjoint(pyborn, c, w, scatter, pairs, shape=None, zc=None, budget=None):
    H = hip_prop(pyborn, c)
    H.set_step_pairs(pairs)
    if shape is not None:
        H.set_tuning("adjoint_state_pair", shape[0], shape[1], 1, 1, zc)
    if budget is not None:
        H.set_checkpoint_budget(budget)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, np.zeros((c.nrec, c.ntb), np.float32))
    g = np.zeros(c.nx * c.ny * c.nz, np.float32)
    H.adjoint_exact(c.jt, w, g, c.nt, scatter=scatter)
    used = H.last_adjoint_pairs
    plan = H.last_checkpoint_plan()
    H.close()
    return g, used, plan


def _residual(c, seed=7):
    return np.random.default_rng(seed).standard_normal((c.nrec, c.ntb)).astype(np.float32)


@pytest.mark.parametrize("scatter", [0, 1])
def test_adjoint_pairs_match_single_steps_and_oracle(pyborn, scatter):
    c = make_case(120, 40, 44, nt=91, nbound=10, src_z=12)
    w = _residual(c)
    got, used, _ = _adjoint(pyborn, c, w, scatter, True)
    assert used
    want, used1, _ = _adjoint(pyborn, c, w, scatter, False)
    assert not used1
    _check(got, want, "adjoint 120x40x44 scatter %d" % scatter)
    P = oracle_prop(c)
    P.transfer_receiver_func(c.rec_locs, np.zeros((c.nrec, c.ntb), np.float32))
    g64 = P.exact_adjoint_d(c.jt, w.astype(np.float64), c.nt, scatter=scatter)
    assert rel_l2(got, g64.ravel()) <= 1e-4


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("zc", [8, 13, 200])
def test_every_adjoint_pair_shape_and_z_chunk(pyborn, shape, zc):
    # receivers inject on a plane next to chunk edges; jt = 3 sinc rows
    c = make_case(121, 43, 37, nt=40, nbound=8, src_z=9, rec_z=7, seed=3, jt=3)
    w = _residual(c, 11)
    want, _, _ = _adjoint(pyborn, c, w, 0, False)
    got, used, _ = _adjoint(pyborn, c, w, 0, True, shape=shape, zc=zc)
    assert used
    _check(got, want, "adjoint shape %s zc %d" % (shape, zc))


def test_adjoint_pairs_across_checkpoint_segments(pyborn):
    # a small budget: several segments of odd length, each swept in pairs
    # plus a single step
    c = make_case(64, 40, 36, nt=67, nbound=8)
    w = _residual(c, 5)
    n = c.nx * c.ny * c.nz * 4
    want, _, plan0 = _adjoint(pyborn, c, w, 1, False, budget=26 * n)
    got, used, plan = _adjoint(pyborn, c, w, 1, True, budget=26 * n)
    assert used and plan == plan0 and plan[0] > 2
    _check(got, want, "adjoint, %d checkpoint segments" % plan[0])


@pytest.mark.parametrize("storage", ["f32", "f16"])
@pytest.mark.parametrize("scatter", [0, 1])
def test_adjoint_recompute_matches_kept_history(pyborn, storage, scatter):
    # the background history the forward keeps and the one the exact adjoint
    # recomputes (with scatter 0 both write s straight into the history) give
    # the same image bit for bit
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    c = make_case(70, 44, 40, nt=53, nbound=8)
    rng = np.random.default_rng(9)
    w = torch.from_numpy(rng.standard_normal((c.ntb, c.nrec)).astype(np.float32)).to(dev)
    model = torch.from_numpy(c.model.ravel()).to(dev)
    imgs = []
    for keep in (True, False):
        H = hip_prop(pyborn, c)
        H.set_storage(storage)
        H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs,
                                 np.zeros((c.nrec, c.ntb), np.float32))
        if keep:
            tr = torch.zeros((c.ntb, c.nrec), dtype=torch.float32, device=dev)
            H.dev_forward(c.jt, c.nt, model, tr, scatter=scatter, keep_background=True)
        img = torch.zeros(c.nx * c.ny * c.nz, dtype=torch.float32, device=dev)
        H.dev_adjoint_exact(c.jt, c.nt, w, img, scatter=scatter)
        H.synchronize()
        imgs.append(img.cpu().numpy())
        H.close()
    _check(imgs[1], imgs[0], "recomputed vs kept history, %s scatter %d" % (storage, scatter))
