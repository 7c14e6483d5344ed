"""Every launch shape of the single-step kernels gives the same bits.

The step kernels evaluate each cell with its rounding fixed in the source, so
the tile shape (rows per thread, wavefronts per block, cells per lane), the
load lookahead (the next plane or the one after) and the z chunk must not
change one bit of the Born data or of the exact-adjoint image.  Each case
pins the forward, or the background recompute and the adjoint-state kernel,
to one shape on a domain that is not a multiple of any tile, with z chunks of
8 and 13 planes (chunk edges next to the source and receiver planes), and
compares with one fixed shape, for FP32, FP16 and 24-bit levels.  That
shape's results are pinned to the CPU oracle by test_gpu_parity.py,
test_gpu_fp16.py and test_gpu_f24.py."""
import numpy as np
import pytest

from cases import hip_prop, make_case, rel_l2

pytestmark = pytest.mark.gpu

# (rows per thread, wavefronts per block, cells per lane)
TILES = [(2, 4, 1), (4, 4, 1), (2, 8, 1), (2, 16, 1), (1, 8, 1), (1, 16, 1),
         (1, 4, 2), (2, 4, 2), (1, 8, 2), (2, 8, 2), (1, 12, 2), (2, 2, 2)]
REF = (1, 8, 1, 1)


def _case():
    # even nx for the two-cell lanes; sources and receivers near chunk edges
    return make_case(122, 43, 37, nt=40, nbound=8, src_z=9, rec_z=7, seed=3, jt=3)


def _run(pyborn, c, storage, kinds, shape, zc, w):
    H = hip_prop(pyborn, c)
    H.set_storage(storage)
    for k in kinds:
        H.set_tuning(k, *shape, zc)
    if w is None:
        out = np.zeros((c.nrec, c.ntb), np.float32)
        H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, out)
        H.forward_scatter(c.jt, c.model, out, c.nt, 0)
    else:
        H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs,
                                 np.zeros((c.nrec, c.ntb), np.float32))
        out = np.zeros(c.nx * c.ny * c.nz, np.float32)
        H.adjoint_exact(c.jt, w, out, c.nt, scatter=0)
    tiles = {k: H.tuning()[k] for k in kinds}
    H.close()
    for k in kinds:
        t = tiles[k]
        assert (t["rows_per_thread"], t["waves"], t["cells_per_lane"], t["lookahead"]) == shape
    return out


def _check(got, want, what):
    assert np.abs(want).max() > 0
    print("%s: rel %.2e, bitwise equal %s" % (what, rel_l2(got, want), np.array_equal(got, want)))
    assert np.array_equal(got, want), what


@pytest.mark.parametrize("storage", ["f32", "f16", "f24"])
@pytest.mark.parametrize("lookahead", [1, 2])
@pytest.mark.parametrize("tile", TILES)
def test_forward_every_shape(pyborn, tile, lookahead, storage):
    c = _case()
    kind = ["fwd" if storage == "f32" else "fwd_" + storage]
    want = _run(pyborn, c, storage, kind, REF, 64, None)
    for zc in (8, 13):
        got = _run(pyborn, c, storage, kind, tile + (lookahead,), zc, None)
        _check(got, want, "forward %s %s lookahead %d zc %d" % (storage, tile, lookahead, zc))


@pytest.mark.parametrize("storage", ["f32", "f16", "f24"])
@pytest.mark.parametrize("lookahead", [1, 2])
@pytest.mark.parametrize("tile", TILES)
def test_adjoint_every_shape(pyborn, tile, lookahead, storage):
    c = _case()
    w = np.random.default_rng(11).standard_normal((c.nrec, c.ntb)).astype(np.float32)
    sfx = "" if storage == "f32" else "_" + storage
    kinds = ["background" + sfx, "adjoint_state" + sfx]
    want = _run(pyborn, c, storage, kinds, REF, 64, w)
    for zc in (8, 13):
        got = _run(pyborn, c, storage, kinds, tile + (lookahead,), zc, w)
        _check(got, want, "adjoint %s %s lookahead %d zc %d" % (storage, tile, lookahead, zc))


@pytest.mark.parametrize("storage", ["f32", "f16", "f24"])
def test_two_cell_shape_on_odd_nx_runs_one_cell_tile(pyborn, storage):
    # a two-cell shape fixed on a grid of odd nx runs the one-cell shape of
    # the same tile (two cells per lane need an even nx): same bits, no error
    c = make_case(121, 43, 37, nt=30, nbound=8, src_z=9, rec_z=7, seed=4, jt=3)
    w = np.random.default_rng(12).standard_normal((c.nrec, c.ntb)).astype(np.float32)
    sfx = "" if storage == "f32" else "_" + storage
    kinds = ["fwd" + sfx, "background" + sfx, "adjoint_state" + sfx]
    outs = []
    for shape in ((1, 4, 2, 1), REF):
        H = hip_prop(pyborn, c)
        H.set_storage(storage)
        for k in kinds:
            H.set_tuning(k, *shape, 13)
        d = np.zeros((c.nrec, c.ntb), np.float32)
        H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, d)
        H.forward_scatter(c.jt, c.model, d, c.nt, 0)
        g = np.zeros(c.nx * c.ny * c.nz, np.float32)
        H.adjoint_exact(c.jt, w, g, c.nt, scatter=0)
        H.close()
        outs.append((d, g))
    for got, want, what in zip(outs[0], outs[1], ("data", "image")):
        _check(got, want, "%s %s odd nx" % (what, storage))
