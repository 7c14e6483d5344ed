"""born_dev_source_prop_noalloc, the entry point that allocates nothing (its
scratch, the per-step source amplitudes, sits in a fixed device array),
against born_dev_source_prop on the same inputs: bit for bit, on two calls
in a row through one context, the second on a smaller grid with fewer
(and an odd number of) steps, and on a fresh context whose first call is the
allocation-free one (no launch shape tuned yet)."""
import numpy as np
import pytest

from cases import hip_prop, make_case

pytestmark = pytest.mark.gpu


def _levels(H, fn, c, nt):
    import torch

    dev = torch.device("cuda:0")
    p0 = torch.zeros(c.nx * c.ny * c.nz, dtype=torch.float32, device=dev)
    p1 = torch.zeros_like(p0)
    fn(1, nt, p0, p1)
    H.synchronize()
    return p0.cpu().numpy(), p1.cpu().numpy()


def _load(H, c):
    H.create_space(*c.d, c.nx, c.ny, c.nz)
    H.set_sponge(c.nbound)
    H.transfer_sinc_table_s(c.tab)
    H.transfer_sinc_table_d(c.tab)
    H.transfer_source_func(c.src_locs, c.src_vals)
    H.transfer_vel_func1(c.V1)
    H.transfer_vel_func2(c.V2)


def _want(pyborn, c, nt):
    H = hip_prop(pyborn, c)
    out = _levels(H, H.dev_source_prop, c, nt)
    H.close()
    return out


def test_noalloc_source_prop_equals_the_allocating_entry(pyborn):
    big = make_case(64, 48, 40, nt=90, rand_band=8, nbound=10)
    small = make_case(40, 36, 32, nt=41, rand_band=6, nbound=8, seed=4)
    want_big = _want(pyborn, big, 90)
    want_small = _want(pyborn, small, 41)
    H = pyborn.Propagator(0)
    _load(H, big)
    got_big = _levels(H, H.dev_source_prop_noalloc, big, 90)
    _load(H, small)
    got_small = _levels(H, H.dev_source_prop_noalloc, small, 41)
    H.close()
    for got, want in ((got_big, want_big), (got_small, want_small)):
        assert np.abs(want[0]).max() > 0
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    assert np.array_equal(got_small[0], got_small[1])  # odd nt (cpu_prop.cc:123-128)


def test_noalloc_source_prop_after_tuning_and_its_size_limit(pyborn):
    c = make_case(48, 40, 36, nt=60, rand_band=6, nbound=8, seed=2)
    H = hip_prop(pyborn, c)
    want = _levels(H, H.dev_source_prop, c, 60)  # tunes the launch shape
    got = _levels(H, H.dev_source_prop_noalloc, c, 60)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
    # 9 source points: (nt + 1) * 9 > 2^20 amplitudes is refused before any work
    import torch

    p = torch.zeros(c.nx * c.ny * c.nz, dtype=torch.float32, device="cuda:0")
    with pytest.raises(pyborn.BornError, match="exceed the scratch"):
        H.dev_source_prop_noalloc(1, 120000, p, p.clone())
    H.close()
