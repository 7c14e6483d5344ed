"""Known-answer tests of the oracle's three loops (base/cpu_prop.cc:22-54
rtmForward, :106-129 sourceProp, :56-96 rtmAdjoint), with the expected values
derived by hand from the reference's step lists.

One source point and receivers at a cell c (and at its x neighbour n), exact
one-tap time tables (a delta at tap 3, so step t injects sample t+3 and
extraction writes sample t+3), constant velocities V1 != V2, the reference
sponge (40 cells, bcB = .0005) so that every cell has its own factor b < 1,
and nt <= 3, so each value is a short closed form in the stencil taps, b, V,
the model and the samples.  The forms follow the step lists:

  rtmForward, per step: s_new = V L s + 2 s - s_old; the same for r; damp
  both levels of s, then of r (s_new *= b, s *= b); s_new += src(it);
  r_new += 1e-6 s_new m; traces(it) += r_new at the receivers; swap.
  sourceProp, nt+1 steps: p_new = V1 L p + 2 p - p_old; p_new += src(it);
  swap; odd nt copies the last-but-one level over the last.
  rtmAdjoint, it = nt-1 .. 0: r_new = V2 L r + 2 r - r_old unless it = 0;
  unless it = nt-1: s_new = V1 L s + 2 s - s_old, then s -= src(it+1);
  damp both levels of r; r_new -= trace(it); image += s_new r_new; swap.

Each misordering a plausible restatement could make -- scatter before
injection, damping after injection, extraction before the scatter, the cur
level not damped in place, vel1/vel2 swapped, damping in sourceProp, the
adjoint re-injecting src(it) instead of src(it+1), imaging with the other
background level, r not damped at it = 0 -- changes at least one value
below by far more than the tolerance.
"""
import numpy as np
import pytest

import oracle

NX, NY, NZ = 26, 24, 22
CX, CY, CZ = 11, 12, 9          # source / first receiver cell
V1, V2 = 0.375, 0.25            # v^2 dt^2 of the background and Born fields
D = 10.0
TOL = 2e-6                      # float32 loops against float64 closed forms

C00 = 8.0 / (5.0 * D * D)
C0 = -1025.0 / 576.0 * 3 * C00  # isotropic: the three axes share the taps
TAPS = [C00, -C00 / 8, C00 / 63, -C00 / 896]
AXES = [(1, 0, 0), (0, 1, 0), (0, 0, 1)]


def idx(x, y, z):
    return x + NX * (y + NY * z)


def bfac(x, y, z):
    """reference sponge factor (cpu_prop.cc:141-165, 329-332): bound[edge],
    edge = distance to the interior face, the far side offset by one"""
    e = min(x - 4, NX - 4 - x, y - 4, NY - 4 - y, z - 4, NZ - 4 - z)
    return float(np.float32(np.exp(-0.0005 * (40 - e)))) if e < 40 else 1.0


def star(fn):
    """sum over the 24 off-centre taps: t_k * (f(c + k e) + f(c - k e))"""
    s = 0.0
    for ax in AXES:
        for k in range(1, 5):
            p = (CX + k * ax[0], CY + k * ax[1], CZ + k * ax[2])
            q = (CX - k * ax[0], CY - k * ax[1], CZ - k * ax[2])
            s += TAPS[k - 1] * (fn(*p) + fn(*q))
    return s


def tap_at(x, y, z):
    """the stencil tap linking c to (x, y, z): c0 at c, t_k on the star, 0 off it"""
    d = (x - CX, y - CY, z - CZ)
    nz = [v for v in d if v != 0]
    if not nz:
        return C0
    if len(nz) == 1 and abs(nz[0]) <= 4:
        return TAPS[abs(nz[0]) - 1]
    return 0.0


def prop(nsrc_pts=1, nrec=2, ntb=16, seed=0):
    rng = np.random.default_rng(seed)
    P = oracle.Prop(nthreads=2)
    P.create_space(D, D, D, NX, NY, NZ)
    e3 = np.zeros((1, 8), np.float32)
    e3[0, 3] = 1.0
    P.transfer_sinc_table_s(e3)
    P.transfer_sinc_table_d(e3)
    src = (rng.uniform(0.5, 1.5, ntb) * rng.choice([-1, 1], ntb)).astype(np.float32)
    P.transfer_source_func(np.array([idx(CX, CY, CZ)], np.int32), src[None, :])
    P.transfer_vel_func1(np.full(NX * NY * NZ, V1, np.float32))
    P.transfer_vel_func2(np.full(NX * NY * NZ, V2, np.float32))
    locs = np.array([idx(CX, CY, CZ), idx(CX + 1, CY, CZ)][:nrec], np.int32)
    return P, src.astype(np.float64), locs


def close(got, want, what):
    assert abs(got - want) <= TOL * max(abs(want), 1e-12), "%s: %r vs %r" % (what, got, want)


def test_rtm_forward_impulse_response():
    nt = 3
    P, src, locs = prop(ntb=nt + 8)
    rng = np.random.default_rng(5)
    m = rng.uniform(0.5, 2.0, NX * NY * NZ).astype(np.float32)
    rec = np.zeros((2, nt + 8), np.float32)
    P.transfer_receiver_func(locs, rec)
    P.rtm_forward(1, m, nt)
    a0, a1, a2 = src[3], src[4], src[5]
    M = lambda x, y, z: float(m[idx(x, y, z)])
    mc, bc = M(CX, CY, CZ), bfac(CX, CY, CZ)
    # step 0: s1 = a0 at c; r1 = 1e-6 s1 m
    r1c = 1e-6 * a0 * mc
    # step 1
    s2 = lambda x, y, z: (bfac(x, y, z) * (V2 * C0 * a0 + 2 * a0) + a1
                          if (x, y, z) == (CX, CY, CZ)
                          else bfac(x, y, z) * V2 * tap_at(x, y, z) * a0)
    r2 = lambda x, y, z: ((bfac(x, y, z) * (V2 * C0 * r1c + 2 * r1c)
                           if (x, y, z) == (CX, CY, CZ)
                           else bfac(x, y, z) * V2 * tap_at(x, y, z) * r1c)
                          + 1e-6 * s2(x, y, z) * M(x, y, z))
    # step 2 at c: old levels are the damped levels 1 (b * a0, b * r1c)
    s3c = bc * (V2 * (C0 * s2(CX, CY, CZ) + star(s2)) + 2 * s2(CX, CY, CZ) - bc * a0) + a2
    r3c = (bc * (V2 * (C0 * r2(CX, CY, CZ) + star(r2)) + 2 * r2(CX, CY, CZ) - bc * r1c)
           + 1e-6 * s3c * mc)
    close(rec[0, 3], r1c, "receiver c, step 0")
    close(rec[0, 4], r2(CX, CY, CZ), "receiver c, step 1")
    close(rec[0, 5], r3c, "receiver c, step 2")
    assert rec[1, 3] == 0.0
    close(rec[1, 4], r2(CX + 1, CY, CZ), "receiver c+x, step 1")
    assert not rec[0, :3].any() and not rec[0, 6:].any()


@pytest.mark.parametrize("nt", [1, 2, 3])
def test_source_prop_impulse_response(nt):
    P, src, _ = prop(ntb=nt + 9)
    p0 = np.zeros(NX * NY * NZ, np.float32)
    p1 = np.zeros_like(p0)
    P.source_prop(p0, p1, 1, nt)
    a0, a1, a2 = src[3], src[4], src[5]
    c = idx(CX, CY, CZ)
    lev1 = lambda x, y, z: a0 if (x, y, z) == (CX, CY, CZ) else 0.0
    lev2 = lambda x, y, z: (V1 * C0 * a0 + 2 * a0 + a1 if (x, y, z) == (CX, CY, CZ)
                            else V1 * tap_at(x, y, z) * a0)
    lev3c = V1 * (C0 * lev2(CX, CY, CZ) + star(lev2)) + 2 * lev2(CX, CY, CZ) - a0 + a2
    if nt == 1:
        # two steps; odd nt: both outputs are level 1 (cpu_prop.cc:123-128)
        for p in (p0, p1):
            close(p[c], a0, "level 1")
            assert np.count_nonzero(p) == 1
    elif nt == 2:
        close(p0[c], lev3c, "p0 = level 3")
        close(p1[c], lev2(CX, CY, CZ), "p1 = level 2")
        close(p1[idx(CX, CY, CZ + 2)], V1 * TAPS[1] * a0, "p1 at c + 2z")
    else:
        # four steps; odd nt: level 4 is lost, both outputs are level 3
        close(p0[c], lev3c, "p0 = level 3")
        assert np.array_equal(p0, p1)


@pytest.mark.parametrize("nt", [1, 2, 3])
def test_rtm_adjoint_impulse_response(nt):
    P, src, locs = prop(ntb=nt + 8)
    # _jt is set by sourceProp (cpu_prop.cc:108), as in the reference call order
    P.source_prop(np.zeros(NX * NY * NZ, np.float32), np.zeros(NX * NY * NZ, np.float32), 1, 1)
    rng = np.random.default_rng(9)
    rec = rng.uniform(-1, 1, (2, nt + 8)).astype(np.float32)
    P.transfer_receiver_func(locs, rec)
    A, B = 1.25, -0.5
    c, n = idx(CX, CY, CZ), idx(CX + 1, CY, CZ)
    s0 = np.zeros(NX * NY * NZ, np.float32)
    s1 = np.zeros_like(s0)
    s0[c], s1[c] = A, B
    img = np.zeros_like(s0)
    P.rtm_adjoint(s0, s1, img, nt)
    w = rec[0].astype(np.float64)[3:]  # trace of receiver c at step t
    u = rec[1].astype(np.float64)[3:]  # receiver n = c + x
    sg = src[3:]
    bc, bn = bfac(CX, CY, CZ), bfac(CX + 1, CY, CZ)
    if nt == 1:
        close(img[c], -A * w[0], "image c")
        assert img[n] == 0.0
        return
    # S1: background after one reverse step from (s0, s1) = (A, B) at c
    S1 = lambda x, y, z: (V1 * C0 * A + 2 * A - B if (x, y, z) == (CX, CY, CZ)
                          else V1 * tap_at(x, y, z) * A)
    if nt == 2:
        close(img[c], -A * w[1] - w[0] * S1(CX, CY, CZ), "image c")
        close(img[n], -u[0] * S1(CX + 1, CY, CZ), "image n")
        return
    # nt = 3
    R1c = bc * (V2 * (C0 * -w[2] + TAPS[0] * -u[2]) + 2 * -w[2]) - w[1]
    R1n = bn * (V2 * (C0 * -u[2] + TAPS[0] * -w[2]) + 2 * -u[2]) - u[1]
    # level before S1 after the re-injection of src(it+1) = src(2) at it = 1
    S0c = V1 * (C0 * S1(CX, CY, CZ) + star(S1)) + 2 * S1(CX, CY, CZ) - (A - sg[2])
    # n = c + x: its x neighbours are c + (1 +- k) x on the star, its y and z
    # neighbours are off it
    Ls1n = C0 * S1(CX + 1, CY, CZ) + sum(
        TAPS[k - 1] * (S1(CX + 1 + k, CY, CZ) + S1(CX + 1 - k, CY, CZ)) for k in range(1, 5))
    S0n = V1 * Ls1n + 2 * S1(CX + 1, CY, CZ)
    R0c = -bc * bc * w[2] - w[0]  # r of step 2 damped at it = 1 and again at it = 0
    R0n = -bn * bn * u[2] - u[0]
    close(img[c], -A * w[2] + S1(CX, CY, CZ) * R1c + S0c * R0c, "image c")
    close(img[n], S1(CX + 1, CY, CZ) * R1n + S0n * R0n, "image n")
