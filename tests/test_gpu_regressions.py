"""Regression tests of fixed defects of the propagator library.

- A context resized to a larger grid after two-step launches must not keep
  the old, smaller two-step level buffers: its data and image equal those of
  a fresh context bit for bit.
- A shot batch's captured time loop holds the device address of the
  receiver sinc table; after a new table is transferred the next call must
  capture the loop again and use the new table.
- The exact adjoint divides by V B; a zero velocity or sponge factor in the
  interior is refused with BORN_ERR_UNSUPPORTED instead of giving inf/NaN.
"""
import numpy as np
import pytest

from cases import hip_prop, make_case

pytestmark = pytest.mark.gpu


def _load(H, c):
    """hip_prop's transfers, on an existing context"""
    H.create_space(*c.d, c.nx, c.ny, c.nz)
    if c.nbound is not None:
        H.set_sponge(c.nbound)
    H.transfer_sinc_table_s(c.tab)
    H.transfer_sinc_table_d(c.tab)
    H.transfer_source_func(c.src_locs, c.src_vals)
    H.transfer_vel_func1(c.V1)
    H.transfer_vel_func2(c.V2)


def _pair_run(H, c):
    rec = np.zeros((c.nrec, c.ntb), np.float32)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, rec)
    H.forward_scatter(c.jt, c.model, rec, c.nt, 0)
    img = np.zeros(c.nx * c.ny * c.nz, np.float32)
    H.adjoint_exact(c.jt, rec, img, c.nt, 0)
    return rec, img


def test_resize_after_step_pairs_matches_fresh_context(pyborn):
    small = make_case(48, 40, 36, nt=30, nbound=8)
    large = make_case(96, 72, 64, nt=40, nbound=8, seed=3)
    H = pyborn.Propagator(0)
    H.set_step_pairs(2)
    _load(H, small)
    _pair_run(H, small)
    assert H.last_step_pairs
    _load(H, large)
    got_rec, got_img = _pair_run(H, large)
    assert H.last_step_pairs
    H.close()
    F = hip_prop(pyborn, large)
    F.set_step_pairs(2)
    want_rec, want_img = _pair_run(F, large)
    F.close()
    assert np.abs(want_rec).max() > 0 and np.abs(want_img).max() > 0
    assert np.array_equal(got_rec, want_rec)
    assert np.array_equal(got_img, want_img)


def _shot(c):
    return {"nx": c.nx, "ny": c.ny, "nz": c.nz, "vel": c.V2, "model": c.model,
            "src_locs": c.src_locs, "src_vals": c.src_vals, "rec_nx": c.rec_nx,
            "rec_ny": c.rec_ny, "rec_locs": c.rec_locs}


def test_batch_recaptures_loop_after_new_receiver_table(pyborn):
    cases = [make_case(64, 48, 40, nt=60, nbound=10, seed=i, src_z=8 + i) for i in range(2)]
    H = hip_prop(pyborn, cases[0])
    # an exact one-tap receiver table: twice the table gives twice the traces
    # exactly (the sinc table of jt = 1 is a delta only to within 5e-6, whose
    # tiny taps give subnormal products that do not scale exactly); the
    # source table (injection) is unchanged
    e3 = np.zeros((1, 8), np.float32)
    e3[0, 3] = 1.0
    H.transfer_sinc_table_d(e3)
    B = pyborn.Batch(H, [_shot(c) for c in cases], cases[0].d)
    B.forward(1, cases[0].nt, 0)
    first = [B.traces(s) for s in range(2)]
    H.transfer_sinc_table_d(2 * e3)
    B.forward(1, cases[0].nt, 0)
    second = [B.traces(s) for s in range(2)]
    B.close()
    H.close()
    for a, b in zip(first, second):
        assert np.abs(a).max() > 0
        assert np.array_equal(b, 2 * a)


def test_exact_adjoint_refuses_zero_velocity(pyborn):
    c = make_case(48, 40, 36, nt=20, nbound=8)
    c.V2 = c.V2.copy()
    c.V2[18, 20, 24] = 0.0
    H = hip_prop(pyborn, c)
    rec = np.ones((c.nrec, c.ntb), np.float32)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, rec)
    img = np.zeros(c.nx * c.ny * c.nz, np.float32)
    with pytest.raises(pyborn.BornError, match="UNSUPPORTED"):
        H.adjoint_exact(c.jt, rec, img, c.nt, 0)
    # a zero outside the stencil interior is never divided by
    c.V2[18, 20, 24] = c.V2[18, 20, 23]
    c.V2[0, 0, 0] = 0.0
    H.transfer_vel_func2(c.V2)
    H.adjoint_exact(c.jt, rec, img, c.nt, 0)
    assert np.all(np.isfinite(img))
    H.close()


def test_exact_adjoint_refuses_zero_sponge_factor(pyborn):
    c = make_case(48, 40, 36, nt=20, nbound=8)
    H = hip_prop(pyborn, c)
    H.set_sponge(8, 1e4)  # exp(-1e4 * (8 - e)) underflows to 0
    rec = np.ones((c.nrec, c.ntb), np.float32)
    H.transfer_receiver_func(c.rec_nx, c.rec_ny, c.rec_locs, rec)
    img = np.zeros(c.nx * c.ny * c.nz, np.float32)
    with pytest.raises(pyborn.BornError, match="UNSUPPORTED"):
        H.adjoint_exact(c.jt, rec, img, c.nt, 0)
    H.close()


def _fresh_pair(pyborn, c, storage):
    F = hip_prop(pyborn, c)
    F.set_storage(storage)
    out = _pair_run(F, c)
    F.close()
    return out


def test_history_of_another_grid_leaves_no_values_in_level_shells(pyborn):
    # The step kernels write level interiors only.  The history buffer kept
    # from a larger grid holds interior values where the smaller grid's
    # level shells lie; they must be cleared before the levels are read.
    large = make_case(96, 72, 64, nt=40, nbound=8, seed=3)
    small = make_case(60, 50, 44, nt=40, nbound=8, seed=4)
    H = pyborn.Propagator(0)
    _load(H, large)
    _pair_run(H, large)
    _load(H, small)
    got = _pair_run(H, small)
    H.close()
    want = _fresh_pair(pyborn, small, "f32")
    for g, w in zip(got, want):
        assert np.abs(w).max() > 0
        assert np.array_equal(g, w)


def test_history_of_another_storage_leaves_no_values_in_level_shells(pyborn):
    # 16-bit levels followed by FP32 levels in the same history buffer
    c = make_case(64, 48, 40, nt=40, nbound=8, seed=5)
    H = hip_prop(pyborn, c)
    for storage in ("f16", "f24", "f32"):
        H.set_storage(storage)
        got = _pair_run(H, c)
        want = _fresh_pair(pyborn, c, storage)
        for g, w in zip(got, want):
            assert np.abs(w).max() > 0
            assert np.array_equal(g, w), storage
    H.close()


def test_batch_history_of_another_storage_leaves_no_values_in_level_shells(pyborn):
    cases = [make_case(64, 48, 40, nt=50, nbound=10, seed=i, src_z=8 + i) for i in range(2)]
    nt = cases[0].nt

    def run(storages):
        H = hip_prop(pyborn, cases[0])
        B = pyborn.Batch(H, [_shot(c) for c in cases], cases[0].d)
        for storage in storages:
            H.set_storage(storage)
            B.forward(1, nt, 0)
            B.adjoint_exact(1, nt, 0)
        out = [B.image(k) for k in range(len(cases))]
        B.close()
        H.close()
        return out

    for g, w in zip(run(["f16", "f32"]), run(["f32"])):
        assert np.abs(w).max() > 0
        assert np.array_equal(g, w)
