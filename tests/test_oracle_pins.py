"""Pins of the CPU oracle (oracle/born_oracle.c) by checks that do not retype
its formulas: the compiled reference, closed forms, invariants and brute
force on tiny inputs.  Each test names the reference code it pins."""
import numpy as np
import pytest

import oracle
from cases import make_case, oracle_prop, rel_l2

L = oracle.lib


def space(d1, d2, d3, nx, ny, nz, width=None):
    P = oracle.Prop(nthreads=4)
    P.create_space(d1, d2, d3, nx, ny, nz)
    if width is not None:
        P.set_sponge(width)
    return P


# ---------------------------------------------------------------- sinc_bob
@pytest.mark.parametrize("jt", [1, 2, 3, 4, 5, 7, 8])
def test_sinc_table_bit_exact_against_compiled_reference(jt):
    """The reference's base/sinc_bob.cc, compiled from its source into oracle/_ref."""
    ref = oracle.ref_sinc_table(jt)
    if ref is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    assert np.array_equal(oracle.sinc_table(jt), ref)
    ofast = oracle.ref_sinc_table(jt, variant="_ofast")
    if ofast is not None:  # the reference's own -Ofast build
        assert np.abs(oracle.sinc_table(jt) - ofast).max() < 1e-6


@pytest.mark.parametrize("jt", range(1, 9))
def test_sinc_table_matches_golden_vectors(jt):
    """Golden vectors written by tests/golden/make_sinc_golden.py from the
    compiled reference (bit patterns): the oracle equals the -O2 build bit for
    bit and the -Ofast build within 1e-6."""
    import json
    import os

    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sinc_tables.json")))
    ours = oracle.sinc_table(jt).ravel().view(np.uint32)
    ref = np.array([int(h, 16) for h in g["tables"]["O2"][str(jt)]], np.uint32)
    assert np.array_equal(ours, ref)
    fast = np.array([int(h, 16) for h in g["tables"]["Ofast"][str(jt)]], np.uint32).view(np.float32)
    assert np.abs(oracle.sinc_table(jt).ravel() - fast).max() < 1e-6


def test_sinc_table_single_phase_is_a_delta():
    """jt=1: d=0, the Toeplitz right-hand side equals column 3, so the exact
    solution is e_3 (sinc_bob.cc:36-64)."""
    t = oracle.sinc_table(1)[0]
    e3 = np.zeros(8, np.float32)
    e3[3] = 1
    assert np.abs(t - e3).max() < 5e-6


def test_sinc_table_interpolates_low_frequencies():
    # a tapered sinc reproduces a slow cosine at fractional positions
    jt = 4
    tab = oracle.sinc_table(jt)
    n = np.arange(-3, 5)
    for it in range(jt):
        frac = it / jt
        y = np.cos(0.2 * n)
        est = (tab[it] * y).sum()
        assert abs(est - np.cos(0.2 * frac)) < 2e-3


# ------------------------------------------------------------ coefficients
def test_coefficients_are_textbook_eighth_order_weights():
    """cpu_prop.cc:296-326: C00(d)=8/(5d^2), taps C00*{1,-1/8,1/63,-1/896} and
    c0=-1025/576 sum C00 equal Fornberg's [-205/72, 8/5, -1/5, 8/315, -1/560]/d^2."""
    d1, d2, d3 = 8.0, 10.0, 12.5
    P = space(d1, d2, d3, 16, 16, 16)
    c = np.array(P.s.coeffs, np.float64)
    w = [8 / 5, -1 / 5, 8 / 315, -1 / 560]
    for k in range(4):
        # index 1+3k is the slow axis (d1), 2+3k fast (d2), 3+3k middle (d3)
        for a, d in enumerate((d1, d2, d3)):
            assert c[1 + 3 * k + a] == pytest.approx(w[k] / d ** 2, rel=1e-6)
    assert c[0] == pytest.approx(-205 / 72 * sum(1 / d ** 2 for d in (d1, d2, d3)), rel=1e-6)
    # the taps annihilate constants: c0 + 2 sum = 0
    assert abs(c[0] + 2 * c[1:].sum()) < 1e-7


def _poly_case(deg_x, deg_y, deg_z, d=(8.0, 10.0, 12.5), n=(24, 22, 20)):
    nx, ny, nz = n
    P = space(*d, nx, ny, nz)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y, Z = (x - nx / 2) / 8.0, (y - ny / 2) / 8.0, (z - nz / 2) / 8.0
    f = X ** deg_x + Y ** deg_y + Z ** deg_z
    # second derivatives in index units (X = (i - c)/8)
    def dd(U, k):
        return k * (k - 1) * U ** max(k - 2, 0) / 64.0 if k >= 2 else 0 * U

    d1, d2, d3 = d
    # fast axis uses d2, middle d3, slow d1 (cpu_prop.cc:6-18 with 298-326)
    lap = dd(X, deg_x) / d2 ** 2 + dd(Y, deg_y) / d3 ** 2 + dd(Z, deg_z) / d1 ** 2
    return P, f.astype(np.float32), lap


@pytest.mark.parametrize("degs", [(2, 0, 0), (0, 2, 0), (0, 0, 2), (4, 6, 8), (9, 3, 5)])
def test_stencil_is_exact_on_polynomials(degs):
    """The 8th-order stencil differentiates polynomials of degree <= 9
    exactly; separate degrees per axis pin the reference's assignment of the
    spacings to the axes."""
    P, f, lap = _poly_case(*degs)
    vel = np.ones(f.size, np.float32)
    p0 = np.zeros(f.size, np.float32)
    P.prop_step(p0, f.ravel(), vel)
    got = p0.reshape(f.shape) - 2 * f.astype(np.float64)
    sl = (slice(4, -4),) * 3
    scale = np.abs(f[sl]).max() * 2 / min(P.s.nx, 1) * 0 + np.abs(f[sl]).max()
    err = np.abs(got[sl] - lap[sl]).max()
    assert err <= 1e-6 * max(scale, 1.0) * 0.02 + 2e-7 * scale + 1e-9


def _stencil_numpy(p1, vel, p0, c):
    """Brute-force float64 stencil with shifted views."""
    p1 = p1.astype(np.float64)
    out = p0.astype(np.float64).copy()
    I = (slice(4, -4),) * 3
    def sh(dz, dy, dx):
        nz, ny, nx = p1.shape
        return p1[4 + dz:nz - 4 + dz, 4 + dy:ny - 4 + dy, 4 + dx:nx - 4 + dx]
    acc = c[0] * sh(0, 0, 0)
    for k in range(1, 5):
        acc = acc + c[2 + 3 * (k - 1)] * (sh(0, 0, -k) + sh(0, 0, k))
        acc = acc + c[3 + 3 * (k - 1)] * (sh(0, -k, 0) + sh(0, k, 0))
        acc = acc + c[1 + 3 * (k - 1)] * (sh(-k, 0, 0) + sh(k, 0, 0))
    out[I] = vel[I] * acc + 2 * p1[I] - p0[I]
    return out


def test_stencil_on_the_reference_test_recipe():
    """test.cu:50-55 inputs (36^3, p1 = i%4, vel = i%20, p0 = 7.0009,
    coefficients for spacing 36) against a brute-force float64 stencil."""
    n = 36
    P = space(36.0, 36.0, 36.0, n, n, n)
    i = np.arange(n ** 3)
    p1 = (i % 4).astype(np.float32)
    vel = (i % 20).astype(np.float32)
    p0 = np.full(n ** 3, 7.0009, np.float32)
    want = _stencil_numpy(p1.reshape(n, n, n), vel.reshape(n, n, n),
                          p0.reshape(n, n, n), np.array(P.s.coeffs, np.float64))
    P.prop_step(p0, p1, vel)
    assert np.abs(p0.reshape(n, n, n) - want).max() < 1e-4
    # the 4-cell shell is never written (cpu_prop.cc:226-230)
    shell = np.ones((n, n, n), bool)
    shell[4:-4, 4:-4, 4:-4] = False
    assert np.all(p0.reshape(n, n, n)[shell] == np.float32(7.0009))


def test_interior_operator_is_symmetric():
    """L restricted to the interior is symmetric: the basis of the exact
    adjoint's change of variables."""
    nx, ny, nz = 13, 12, 11
    P = space(7.0, 9.0, 11.0, nx, ny, nz)
    cells = [(z, y, x) for z in range(4, nz - 4) for y in range(4, ny - 4) for x in range(4, nx - 4)]
    idx = [z * nx * ny + y * nx + x for z, y, x in cells]
    Lm = np.zeros((len(idx), len(idx)))
    vel = np.ones(nx * ny * nz, np.float32)
    for j, c in enumerate(idx):
        e = np.zeros(nx * ny * nz, np.float32)
        e[c] = 1
        p0 = np.zeros_like(e)
        P.prop_step(p0, e, vel)
        Lm[:, j] = p0[idx] - 2 * e[idx]
    assert np.abs(Lm - Lm.T).max() < 1e-9


# ------------------------------------------------------------------ sponge
def test_sponge_special_cells():
    """cpu_prop.cc:141-165 with the hard-coded table (:329-332): edge 0 at the
    near face, edge 1 at the far face (n-4-i), untouched beyond 40 cells."""
    n = 100
    P = space(10.0, 10.0, 10.0, n, n, n)
    a = np.ones(n ** 3, np.float32)
    b = np.ones(n ** 3, np.float32)
    P.damp(a, b)
    A = a.reshape(n, n, n)
    bound = lambda e: np.float32(np.exp(np.float32(-0.0005) * np.float32(40 - e)))
    assert A[4, 50, 50] == pytest.approx(bound(0), rel=1e-7)
    assert A[50, 50, n - 5] == pytest.approx(bound(1), rel=1e-7)
    assert A[50, 4 + 7, 50] == pytest.approx(bound(7), rel=1e-7)
    assert A[50, 50, 50] == 1.0  # edge 46 >= 40
    assert A[0, 0, 0] == 1.0 and A[n - 1, 50, 50] == 1.0  # shell untouched
    assert np.array_equal(a, b)  # both levels damped alike
    # along one axis the sponge has 1 cell at edge 0 and 2 at each edge 1..39
    line = A[50, 50, 4:n - 4]
    table = np.array(P.s.bound[:40], np.float32)
    for e in range(0, 40):
        cnt = np.count_nonzero(line == table[e])
        assert cnt == (1 if e == 0 else 2)


# -------------------------------------------------- injection / extraction
def test_injection_and_extraction_brute_force():
    """cpu_prop.cc:166-215 on random tables, jt=3, including the id+7 guard."""
    rng = np.random.default_rng(0)
    nx = ny = nz = 16
    P = space(10.0, 10.0, 10.0, nx, ny, nz)
    jt, ntb, nrec = 3, 20, 5
    tab = rng.standard_normal((jt, 8)).astype(np.float32)
    P.transfer_sinc_table_s(tab)
    P.transfer_sinc_table_d(tab)
    locs = rng.choice(np.arange(nx * ny * nz), 4, replace=False).astype(np.int32)
    src = rng.standard_normal((4, ntb)).astype(np.float32)
    P.transfer_source_func(locs, src)
    rlocs = rng.choice(np.arange(nx * ny * nz), nrec, replace=False).astype(np.int32)
    rec = rng.standard_normal((nrec, ntb)).astype(np.float32)
    P.transfer_receiver_func(rlocs, rec)
    P.s.jt = jt
    for id_, ii in [(0, 0), (5, 2), (12, 1), (13, 0)]:
        f = rng.standard_normal(nx * ny * nz).astype(np.float32)
        for dirn in (1, -1):
            P.s.dir = dirn
            g = f.copy()
            L().orc_inject_source(P.ptr, id_, ii, g)
            want = f.astype(np.float64).copy()
            if id_ + 7 < ntb:
                for i in range(4):
                    want[locs[i]] += dirn / jt * np.dot(tab[ii].astype(np.float64),
                                                        src[i, id_:id_ + 8].astype(np.float64))
            assert np.abs(g - want).max() < 1e-5
            g = f.copy()
            L().orc_inject_receivers(P.ptr, id_, ii, g)
            want = f.astype(np.float64).copy()
            if id_ + 7 < ntb:
                for i in range(nrec):
                    want[rlocs[i]] += dirn / jt * np.dot(tab[ii].astype(np.float64),
                                                         rec[i, id_:id_ + 8].astype(np.float64))
            assert np.abs(g - want).max() < 1e-5
        before = rec.copy()
        L().orc_data_extract(P.ptr, id_, ii, f)
        want = before.astype(np.float64)
        if id_ + 7 < ntb:
            for i in range(nrec):
                want[i, id_:id_ + 8] += f[rlocs[i]] * tab[ii]
        assert np.abs(rec - want).max() < 1e-5


def test_born_scatter_is_rounded_once_in_double():
    """cpu_prop.cc:220: r += .000001 * s * m evaluated in double."""
    rng = np.random.default_rng(1)
    n = 12
    P = space(10.0, 10.0, 10.0, n, n, n)
    r = rng.standard_normal(n ** 3).astype(np.float32)
    s = (rng.standard_normal(n ** 3) * 1e3).astype(np.float32)
    m = rng.standard_normal(n ** 3).astype(np.float32)
    want = (r.astype(np.float64) + 1e-6 * s.astype(np.float64) * m.astype(np.float64)).astype(np.float32)
    L().orc_image_add(P.ptr, m, r, s)
    assert np.array_equal(r, want)


# ---------------------------------------------------------- trace / image ops
def _dot(a, b):
    return float(np.dot(np.ravel(a).astype(np.float64), np.ravel(b).astype(np.float64)))


def test_trace_and_image_operators_pass_dot_tests():
    rng = np.random.default_rng(2)
    nt, n2, n3, ntb = 30, 5, 4, 38
    m = rng.standard_normal(ntb * n2 * n3).astype(np.float32)
    d = rng.standard_normal(nt * n2 * n3).astype(np.float32)
    fm = np.zeros_like(d)
    L().orc_map_data_forward(nt, n2, n3, ntb, m, fm)
    am = np.zeros_like(m)
    L().orc_map_data_adjoint(nt, n2, n3, ntb, am, d)
    assert abs(_dot(fm, d) - _dot(m, am)) <= 1e-5 * abs(_dot(fm, d))
    # map_data: data sample t is buffer sample t+4 (map_data_3d.cc:71)
    assert np.array_equal(fm.reshape(n3, n2, nt), m.reshape(n3, n2, ntb)[:, :, 4:4 + nt])

    m = rng.standard_normal(nt * n2 * n3).astype(np.float32)
    fm = np.zeros_like(m)
    L().orc_deriv_forward(nt, n2, n3, 0.004, m, fm)
    am = np.zeros_like(m)
    L().orc_deriv_adjoint(nt, n2, n3, 0.004, am, d)
    assert abs(_dot(fm, d) - _dot(m, am)) <= 1e-5 * abs(_dot(fm, d))

    nx, ny, nz = 6, 5, 7
    m = rng.standard_normal(nx * ny * nz).astype(np.float32)
    d = rng.standard_normal(nx * ny * nz).astype(np.float32)
    fm = np.zeros_like(m)
    L().orc_laplac_forward(nx, ny, nz, m, fm)
    am = np.zeros_like(m)
    L().orc_laplac_adjoint(nx, ny, nz, am, d)
    assert abs(_dot(fm, d) - _dot(m, am)) <= 1e-5 * abs(_dot(fm, d))


def test_deriv_duplicated_first_sample_terms():
    """deriv_3d.cc:40-55 adds the first-sample difference once per plane (to
    trace 0 only), once per trace, and once in the loop."""
    nt, n2, n3, dt = 6, 3, 2, 0.5
    for i2, i3, mult in [(0, 0, 3), (0, 1, 3), (2, 1, 2)]:
        m = np.zeros(nt * n2 * n3, np.float32)
        j = i3 * nt * n2 + i2 * nt
        m[j] = 1.0
        out = np.zeros_like(m)
        L().orc_deriv_forward(nt, n2, n3, dt, m, out)
        want = np.zeros_like(m)
        want[j] = -mult / dt
        want[j + 1] = mult / dt
        assert np.array_equal(out, want)
    # interior samples: plain forward difference
    m = np.zeros(nt * n2 * n3, np.float32)
    m[3] = 1.0
    out = np.zeros_like(m)
    L().orc_deriv_forward(nt, n2, n3, dt, m, out)
    assert out[3] == -2.0 and out[4] == 2.0 and np.count_nonzero(out) == 2


def test_laplac_is_a_vertical_first_difference():
    """laplac_3d.cc:51-75: d[i] -= m[i], d[i + nx ny] += m[i] for i3 < nz-1."""
    nx, ny, nz = 4, 3, 5
    m = np.zeros(nx * ny * nz, np.float32)
    m[(2 * ny + 1) * nx + 2] = 1.0  # (z=2, y=1, x=2)
    d = np.zeros_like(m)
    L().orc_laplac_forward(nx, ny, nz, m, d)
    D = d.reshape(nz, ny, nx)
    assert D[2, 1, 2] == -1 and D[3, 1, 2] == 1 and np.count_nonzero(D) == 2
    m = np.zeros_like(m)
    m[(4 * ny) * nx] = 1.0  # last plane: ignored
    d = np.zeros_like(m)
    L().orc_laplac_forward(nx, ny, nz, m, d)
    assert np.count_nonzero(d) == 0


def test_image_window_round_trip():
    """image_rtm_3d.cc:24-95: extract_sub then add_image is the identity on
    the overlap and zero elsewhere."""
    rng = np.random.default_rng(3)
    n1o, n2o, n3o = 10, 9, 8
    image = rng.standard_normal(n1o * n2o * n3o).astype(np.float32)
    n1i, n2i, n3i = 16, 15, 14
    img = np.zeros(n1i * n2i * n3i, np.float32)
    L().orc_extract_sub(n1o, n2o, n3o, 0, 0, 0, 10, 10, 10, image, n1i, n2i, n3i,
                        -30, -30, -30, img)
    I = img.reshape(n3i, n2i, n1i)
    assert np.array_equal(I[3:11, 3:12, 3:13], image.reshape(n3o, n2o, n1o))
    assert I.sum() == pytest.approx(image.sum(), rel=1e-5)
    out = np.zeros_like(image)
    L().orc_add_image(n1o, n2o, n3o, 0, 0, 0, 10, 10, 10, out, n1i, n2i, n3i, -30, -30, -30, img)
    assert np.array_equal(out, image)


# ---------------------------------------------------------- preparation
def test_padded_size_and_stability_closed_forms():
    """source_func_3d.cc:4-35 and fd_prop_3d.cc:6-33 for the benchmark grids."""
    pad = L().orc_padded_size
    assert pad(100, 20, 4) == 160 and pad(256, 20, 4) == 304
    assert pad(512, 20, 4) == 560 and pad(400, 20, 4) == 448
    assert pad(640, 20, 4) == 688 and pad(500, 20, 4) == 560
    assert pad(100, 50, 4) == 208
    import ctypes as C
    dt, nt = C.c_float(), C.c_int()
    ds = np.array([0.001, 0.001], np.float32)
    L().orc_calc_stability(np.float32(1500 + 15 * np.sqrt(990.0)), 10.0, ds, 2, 800,
                           C.byref(dt), C.byref(nt))
    assert dt.value == pytest.approx(0.001) and nt.value == 800
    # a faster model needs a finer step: jt = 3 sub-steps per sample
    L().orc_calc_stability(6000.0, 10.0, ds, 2, 800, C.byref(dt), C.byref(nt))
    assert dt.value == pytest.approx(0.001 / 2) and nt.value == 799 * 2 + 1


def test_source_stamp_closed_form():
    """source_func_3d.cc:76-109: 3x3 stamp in the plane iz-1 with Gaussian
    weights of the +.5-shifted fractional distances, samples at offset +4."""
    nx, ny = 20, 18
    wdev = np.arange(1, 6, dtype=np.float32)
    locs = np.zeros(9, np.int32)
    nts = 16
    vals = np.zeros(9 * nts, np.float32)
    L().orc_source_func(nx, ny, 0.0, 0.0, 0.0, 10.0, 10.0, 10.0, 52.0, 61.0, 80.0,
                        wdev, 5, nts, locs, vals)
    fx, fy, fz = 52 / 10 + .5, 61 / 10 + .5, 80 / 10 + .5
    ix, iy, iz = int(fx), int(fy), int(fz)
    k = 0
    for y in (iy - 1, iy, iy + 1):
        for x in (ix - 1, ix, ix + 1):
            assert locs[k] == x + y * nx + (iz - 1) * nx * ny
            w = np.exp(-((fz - (iz - 1)) ** 2 + (fx - x) ** 2 + (fy - y) ** 2))
            row = vals.reshape(9, nts)[k]
            assert np.allclose(row[4:9], w * wdev, rtol=1e-6)
            assert np.all(row[:4] == 0) and np.all(row[9:] == 0)
            k += 1


def test_rand_subspace_without_band_is_nearest_lower_cell():
    """vel_fd_3d.cc:29-134 with nrand = 0: (v of the nearest-lower, clamped
    model cell)^2 dt^2, no random number drawn."""
    vn = (5, 4, 3)
    v = (1000 + np.arange(np.prod(vn))).astype(np.float32)
    n1, n2, n3 = 9, 8, 7
    out = np.zeros(n1 * n2 * n3, np.float32)
    dt = np.float32(0.001)
    vmax = L().orc_rand_subspace(7, 0, 0, 0, dt, *vn, 0.0, 0.0, 0.0, 10.0, 10.0, 10.0, v,
                                 n1, n2, n3, -20.0, -20.0, -20.0, 10.0, 10.0, 10.0, out)
    assert vmax == v.max()
    O = out.reshape(n3, n2, n1)
    V = v.reshape(vn[2], vn[1], vn[0])
    def vv(val):
        return np.float32(np.float32(np.float32(val * val) * dt) * dt)
    assert O[0, 0, 0] == vv(V[0, 0, 0])            # before the model: clamped
    assert O[2, 3, 4] == vv(V[0, 1, 2])            # (x-2, y-2, z-2) inside
    assert O[6, 7, 8] == vv(V[2, 3, 4])            # beyond the model: clamped


def _band_dist(n, nr):
    """vel_fd_3d.cc:57-103: squared normalised distance into the band along
    one axis, in float as the reference evaluates it"""
    d2 = np.zeros(n, np.float32)
    for i in range(n):
        if i < nr:
            d2[i] = np.float32((nr - i) * (nr - i)) / np.float32(nr * nr)
        elif i >= n - nr:
            k = i - (n - nr - 1)
            d2[i] = np.float32(k * k) / np.float32(nr * nr)
    return d2


def test_rand_subspace_band_per_cell_envelope():
    """vel_fd_3d.cc:29-134 with a random band: every cell of the band obeys
    the reference's acceptance test (:115-122) -- |v + dev| < vmax * 1.03 *
    (1 - min(1, dist) * .4) and v + dev > 1e-4 -- and the range of its draw,
    dev = (u - .5 - .49 dist) * 1.3 * dist * v with u in [0, 1]; cells with
    dist < 1e-4 keep v exactly.  The values of the draws themselves are parity
    unpinned: the reference calls rand() inside a parallel loop (:50-122)."""
    vn = (14, 12, 10)
    z = np.arange(vn[2], dtype=np.float32)
    v = np.ascontiguousarray(np.broadcast_to((1500 + 120 * z)[:, None, None],
                                             (vn[2], vn[1], vn[0])), np.float32).ravel()
    n1, n2, n3 = 20, 22, 18
    nr = (5, 6, 4)
    o = (-30.0, -40.0, -20.0)  # domain origin: nearest-lower model cells, clamped
    dt = np.float32(0.001)
    out = np.zeros(n1 * n2 * n3, np.float32)
    vmax = L().orc_rand_subspace(42, *nr, dt, *vn, 0.0, 0.0, 0.0, 10.0, 10.0, 10.0, v,
                                 n1, n2, n3, *o, 10.0, 10.0, 10.0, out)
    assert vmax == v.max()
    out2 = np.zeros_like(out)
    L().orc_rand_subspace(42, *nr, dt, *vn, 0.0, 0.0, 0.0, 10.0, 10.0, 10.0, v,
                          n1, n2, n3, *o, 10.0, 10.0, 10.0, out2)
    assert np.array_equal(out, out2)  # the serial rand() order is reproducible
    V = v.reshape(vn[2], vn[1], vn[0])
    ix = np.clip(((o[0] + 10.0 * np.arange(n1)) / 10.0).astype(int), 0, vn[0] - 1)
    iy = np.clip(((o[1] + 10.0 * np.arange(n2)) / 10.0).astype(int), 0, vn[1] - 1)
    iz = np.clip(((o[2] + 10.0 * np.arange(n3)) / 10.0).astype(int), 0, vn[2] - 1)
    val = V[iz][:, iy][:, :, ix].astype(np.float64)
    dx2, dy2, dz2 = _band_dist(n1, nr[0]), _band_dist(n2, nr[1]), _band_dist(n3, nr[2])
    dist = np.sqrt((dx2[None, None, :] + dy2[None, :, None] + dz2[:, None, None])
                   .astype(np.float32)).astype(np.float64)
    O = out.reshape(n3, n2, n1)
    got = np.sqrt(O.astype(np.float64)) / float(dt)  # v + dev
    inner = dist < 1e-4
    f32 = np.float32
    keep = (val.astype(f32) * val.astype(f32) * dt * dt).astype(f32)
    assert inner.sum() > 0 and np.array_equal(O[inner], keep[inner])
    band = ~inner
    lim = float(vmax) * 1.03 * (1.0 - np.minimum(1.0, dist) * 0.4)
    slack = 1e-5 * got  # float rounding of (v + dev)^2 dt^2 and the square root
    assert np.all(np.abs(got[band]) < lim[band] + slack[band])
    assert np.all(got[band] > 1e-4)
    rel = (got - val) / val
    lo = 1.3 * dist * (-0.5 - 0.49 * dist)
    hi = 1.3 * dist * (0.5 - 0.49 * dist)
    assert np.all(rel[band] >= lo[band] - 1e-5) and np.all(rel[band] <= hi[band] + 1e-5)
    # the band is drawn, not copied: most cells move
    assert np.mean(np.abs(rel[band]) > 1e-6) > 0.95


# ------------------------------------------------------------ whole loops
def test_forward_is_linear_in_the_model():
    c = make_case(40, 36, 32, nt=50, nbound=8)
    rng = np.random.default_rng(4)
    m2 = rng.standard_normal(c.model.shape).astype(np.float32)
    outs = []
    for m in (c.model, m2, 2 * c.model - 3 * m2):
        P = oracle_prop(c)
        rec = np.zeros((c.nrec, c.ntb), np.float32)
        P.transfer_receiver_func(c.rec_locs, rec)
        P.rtm_forward(1, m.astype(np.float32), c.nt)
        outs.append(rec)
    assert rel_l2(outs[2], 2 * outs[0] - 3 * outs[1]) < 1e-5
    P = oracle_prop(c)
    rec = np.zeros((c.nrec, c.ntb), np.float32)
    P.transfer_receiver_func(c.rec_locs, rec)
    P.rtm_forward(1, np.zeros_like(c.model), c.nt)
    assert np.count_nonzero(rec) == 0


def test_undamped_leapfrog_is_time_reversible():
    """cpu_prop.cc:106-122 without a source: p_{n+1} = A p_n - p_{n-1} run
    again from the swapped final pair returns to the start (a sign error in
    the leapfrog term would not)."""
    c = make_case(32, 30, 28, nt=40)
    c.src_vals[:] = 0
    P = oracle_prop(c)
    rng = np.random.default_rng(5)
    n = c.nx * c.ny * c.nz
    a0 = np.zeros(n, np.float32)
    a1 = np.zeros(n, np.float32)
    inner = np.zeros((c.nz, c.ny, c.nx), bool)
    inner[4:-4, 4:-4, 4:-4] = True
    a0[inner.ravel()] = rng.standard_normal(inner.sum()) * 1e-3
    a1[inner.ravel()] = rng.standard_normal(inner.sum()) * 1e-3
    s0, s1 = a0.copy(), a1.copy()
    nt = 30  # nt+1 steps; even nt: p0 holds x_{nt+1}, p1 holds x_nt
    P.source_prop(a0, a1, 1, nt)
    # run back: the same recursion with the newest level as the old one
    b0, b1 = a0.copy(), a1.copy()
    P.source_prop(b0, b1, 1, nt)
    assert rel_l2(b0, s0) < 1e-4 and rel_l2(b1, s1) < 1e-4


@pytest.mark.parametrize("scatter", [0, 1])
def test_fp64_exact_adjoint_passes_the_dot_test(scatter):
    c = make_case(40, 36, 32, nt=60, nbound=8)
    rng = np.random.default_rng(6)
    P = oracle_prop(c)
    P.transfer_receiver_func(c.rec_locs, np.zeros((c.nrec, c.ntb), np.float32))
    m = rng.standard_normal(c.model.shape).astype(np.float32)
    w = rng.standard_normal((c.nrec, c.ntb))
    fm = P.born_forward_d(1, m, c.nt, scatter=scatter)
    g = P.exact_adjoint_d(1, w, c.nt, scatter=scatter)
    a, b = _dot(fm, w), _dot(m, g)
    assert abs(a - b) <= 1e-12 * abs(a)
    gf = P.exact_adjoint_f(1, w.astype(np.float32), c.nt, scatter=scatter)
    assert rel_l2(gf, g) < 1e-5


def test_fp32_forward_matches_fp64_born():
    c = make_case(40, 36, 32, nt=60, nbound=8)
    P = oracle_prop(c)
    rec = np.zeros((c.nrec, c.ntb), np.float32)
    P.transfer_receiver_func(c.rec_locs, rec)
    P.rtm_forward(1, c.model, c.nt)
    d64 = P.born_forward_d(1, c.model, c.nt, scatter=0)
    assert rel_l2(rec, d64) < 1e-5


def test_born_matches_finite_difference_of_nonlinear_modelling_fp64():
    c = make_case(40, 36, 32, nt=60, nbound=8)
    rng = np.random.default_rng(7)
    dv = rng.standard_normal(c.model.shape) * 20.0
    v = c.v.astype(np.float64)
    dt = float(c.dt)
    P = oracle_prop(c)
    P.transfer_receiver_func(c.rec_locs, np.zeros((c.nrec, c.ntb), np.float32))
    h = 1e-3
    fd = (P.nonlinear_d(1, (v + h * dv) ** 2 * dt * dt, c.nt) -
          P.nonlinear_d(1, (v - h * dv) ** 2 * dt * dt, c.nt)) / (2 * h)
    dV = (2 * v * dv * dt * dt).astype(np.float32)
    born = P.born_forward_d(1, dV, c.nt, scatter=1)
    assert rel_l2(born, fd) < 1e-5
