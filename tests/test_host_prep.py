"""The shipped host preparation (libborn_host.so, C++) against the oracle's
restatement of the same reference code: bit-exact, since both keep the
reference's float rounding (CPU only, no GPU needed)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import load_pkg_module

L = oracle.lib


@pytest.fixture(scope="module")
def H():
    return load_pkg_module("configs").host()


@pytest.mark.parametrize("jt", [1, 2, 3, 5, 8])
def test_sinc_table_bit_exact(H, jt):
    a = np.zeros((jt, 8), np.float32)
    H.born_host_sinc_table(jt, 8, a)
    assert np.array_equal(a, oracle.sinc_table(jt))


def test_padded_size_and_stability(H):
    for n, nb in [(100, 20), (256, 20), (512, 20), (400, 20), (640, 20), (500, 20), (37, 7)]:
        assert H.born_host_padded_size(n, nb, 4) == L().orc_padded_size(n, nb, 4)
    for vmax, dmin in [(1974.3, 10.0), (6000.0, 10.0), (3000.0, 7.5)]:
        a_dt, a_nt, b_dt, b_nt = C.c_float(), C.c_int(), C.c_float(), C.c_int()
        H.born_host_calc_stability(vmax, dmin, 0.001, 800, C.byref(a_dt), C.byref(a_nt))
        L().orc_calc_stability(vmax, dmin, np.array([0.001, 0.001], np.float32), 2, 800,
                               C.byref(b_dt), C.byref(b_nt))
        assert a_dt.value == b_dt.value and a_nt.value == b_nt.value


def test_source_stamp_bit_exact(H):
    rng = np.random.default_rng(0)
    w = rng.standard_normal(60).astype(np.float32)
    wd_a, wd_b = np.zeros_like(w), np.zeros_like(w)
    H.born_host_wavelet_deriv(60, 0.004, w, wd_a)
    L().orc_wavelet_deriv(60, 0.004, w, wd_b)
    assert np.array_equal(wd_a, wd_b)
    for sx, sy, sz, nts in [(123.4, 234.5, 20.0, 80), (55.0, 61.0, 37.3, 40)]:
        la, lb = np.zeros(9, np.int32), np.zeros(9, np.int32)
        va, vb = np.zeros(9 * nts, np.float32), np.zeros(9 * nts, np.float32)
        H.born_host_source_func(40, 36, -30.0, -30.0, -30.0, 10.0, 10.0, 10.0, sx, sy, sz,
                                wd_a, 60, nts, la, va)
        L().orc_source_func(40, 36, -30.0, -30.0, -30.0, 10.0, 10.0, 10.0, sx, sy, sz,
                            wd_b, 60, nts, lb, vb)
        assert np.array_equal(la, lb) and np.array_equal(va, vb)


@pytest.mark.parametrize("nrand", [0, 6])
def test_rand_subspace_bit_exact(H, nrand):
    rng = np.random.default_rng(1)
    vn = (12, 11, 10)
    v = (1500 + 1500 * rng.random(np.prod(vn))).astype(np.float32)
    n = (24, 22, 20)
    a, b = np.zeros(np.prod(n), np.float32), np.zeros(np.prod(n), np.float32)
    args = (99, nrand, nrand, nrand, np.float32(0.0008), *vn, 0.0, 0.0, 0.0, 10.0, 10.0, 10.0,
            v, *n, -60.0, -55.0, -50.0, 10.0, 10.0, 10.0)
    assert H.born_host_rand_subspace(*args, a) == L().orc_rand_subspace(*args, b)
    assert np.array_equal(a, b)


# the large case splits every loop over host threads (traces, planes and
# 2^16-element blocks); outputs start non-zero since every op accumulates
@pytest.mark.parametrize("big", [False, True])
def test_receivers_and_linear_trace_ops_bit_exact(H, big):
    rng = np.random.default_rng(2)
    la, lb = np.zeros(35, np.int32), np.zeros(35, np.int32)
    H.born_host_receiver_locs(7, 5, 10.0, 15.0, 20.0, 20.0, 10.0, 40, 36, -240.0, -240.0,
                              -240.0, 10.0, 10.0, 10.0, la)
    L().orc_receiver_locs(7, 5, 10.0, 15.0, 20.0, 20.0, 10.0, 40, 36, -240.0, -240.0, -240.0,
                          10.0, 10.0, 10.0, lb)
    assert np.array_equal(la, lb)
    nt, n2, n3, ntb = (50, 40, 30, 58) if big else (20, 4, 3, 28)
    lap = (40, 30, 80) if big else (5, 4, 6)
    nl = int(np.prod(lap))
    m = rng.standard_normal(ntb * n2 * n3).astype(np.float32)
    d = rng.standard_normal(nt * n2 * n3).astype(np.float32)
    z = (lambda a: rng.standard_normal(a.size).astype(np.float32)) if big else np.zeros_like
    q = rng.standard_normal(nl).astype(np.float32)
    pairs = [
        ("born_host_map_forward", "orc_map_data_forward", (nt, n2, n3, ntb), m, z(d)),
        ("born_host_map_adjoint", "orc_map_data_adjoint", (nt, n2, n3, ntb), z(m), d),
        ("born_host_deriv_forward", "orc_deriv_forward", (nt, n2, n3, 0.004), d, z(d)),
        ("born_host_deriv_adjoint", "orc_deriv_adjoint", (nt, n2, n3, 0.004), z(d), d),
        ("born_host_laplac_forward", "orc_laplac_forward", lap, q, z(q)),
        ("born_host_laplac_adjoint", "orc_laplac_adjoint", lap, z(q), q),
    ]
    for hname, oname, dims, x, y in pairs:
        fa = getattr(H, hname)
        fa.argtypes = [C.c_int] * 3 + ([C.c_int] if len(dims) == 4 and isinstance(dims[3], int)
                                       else [C.c_float] if len(dims) == 4 else []) + [
            np.ctypeslib.ndpointer(np.float32, flags="C")] * 2
        xa, ya, xb, yb = x.copy(), y.copy(), x.copy(), y.copy()
        fa(*dims, xa, ya)
        getattr(L(), oname)(*dims, xb, yb)
        assert np.array_equal(xa, xb) and np.array_equal(ya, yb), hname


def test_image_windows_bit_exact(H):
    rng = np.random.default_rng(3)
    image = rng.standard_normal(10 * 9 * 8).astype(np.float32)
    a, b = np.zeros(16 * 15 * 14, np.float32), np.zeros(16 * 15 * 14, np.float32)
    H.born_host_extract_sub(10, 9, 8, 0, 0, 0, 10, 10, 10, image, 16, 15, 14, -30, -30, -30, a)
    L().orc_extract_sub(10, 9, 8, 0, 0, 0, 10, 10, 10, image, 16, 15, 14, -30, -30, -30, b)
    assert np.array_equal(a, b)
    ia, ib = image.copy(), image.copy()
    H.born_host_add_image(10, 9, 8, 0, 0, 0, 10, 10, 10, ia, 16, 15, 14, -30, -30, -30, a)
    L().orc_add_image(10, 9, 8, 0, 0, 0, 10, 10, 10, ib, 16, 15, 14, -30, -30, -30, b)
    assert np.array_equal(ia, ib)
