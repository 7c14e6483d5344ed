"""Small seeded Born problems for the parity tests.

Inputs are built with the oracle's restatement of the reference's
preparation code (source stamp, sinc table, receiver mapping), so the oracle
and the HIP library receive identical baseProp inputs.
"""
import numpy as np

import oracle


def ricker(n, dt, f, t0):
    t = np.arange(n) * dt - t0
    a = (np.pi * f * t) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


class Case:
    pass


def make_case(nx=64, ny=48, nz=40, nt=100, d=10.0, dt=1e-3, seed=0, rec_step=2,
              rec_z=6, src_z=8, fpeak=25.0, nbound=None, rand_band=0, jt=1,
              vel_law="vz", aniso=None):
    """A domain of nx*ny*nz cells (x fastest) with v(z) = 1500 + 15 sqrt(z)
    clipped to [1500, 3000], a 9-point Ricker source and a receiver grid."""
    rng = np.random.default_rng(seed)
    c = Case()
    c.nx, c.ny, c.nz, c.nt, c.jt = nx, ny, nz, nt, jt
    d1, d2, d3 = (d, d, d) if aniso is None else aniso
    c.d = (d1, d2, d3)
    c.dt = np.float32(dt)
    zm = np.arange(nz, dtype=np.float32) * np.float32(d3)
    if vel_law == "vz":
        vz = np.clip(1500 + 15 * np.sqrt(zm), 1500, 3000).astype(np.float32)
    else:
        vz = np.full(nz, 2000.0, np.float32)
    v = np.broadcast_to(vz[:, None, None], (nz, ny, nx)).astype(np.float32)
    c.v = v
    # (v+dev)^2 dt^2 rounded in float as vel_fd_3d.cc:125
    c.V2 = (v * v * c.dt * c.dt).astype(np.float32)
    if rand_band:
        out = np.zeros(nz * ny * nx, np.float32)
        oracle.lib().orc_rand_subspace(
            1234, rand_band, rand_band, rand_band, c.dt, nx, ny, nz, 0.0, 0.0, 0.0,
            d1, d2, d3, np.ascontiguousarray(v.ravel()), nx, ny, nz, 0.0, 0.0, 0.0,
            d1, d2, d3, out)
        c.V1 = out.reshape(nz, ny, nx)
    else:
        c.V1 = c.V2.copy()
    # source: derivative of a Ricker, 3x3 stamp (source_func_3d.cc:76-127)
    c.ntb = (nt - 1) // jt + 1 + 8
    nw = min(120, c.ntb - 8)
    w = ricker(nw, dt * jt, fpeak, 1.2 / fpeak)
    wdev = np.zeros(nw, np.float32)
    oracle.lib().orc_wavelet_deriv(nw, np.float32(dt * jt), w, wdev)
    c.src_locs = np.zeros(9, np.int32)
    c.src_vals = np.zeros(9 * c.ntb, np.float32)
    sx, sy, sz = (nx // 2 + 0.3) * d1, (ny // 2 - 0.2) * d2, (src_z + 0.4) * d3
    oracle.lib().orc_source_func(nx, ny, 0.0, 0.0, 0.0, d1, d2, d3, sx, sy, sz, wdev, nw,
                                 c.ntb, c.src_locs, c.src_vals)
    c.src_vals = c.src_vals.reshape(9, c.ntb)
    # receivers on one plane
    xs = np.arange(6, nx - 6, rec_step)
    ys = np.arange(6, ny - 6, rec_step)
    c.rec_nx, c.rec_ny = len(xs), len(ys)
    c.rec_locs = np.array([x + y * nx + rec_z * nx * ny for y in ys for x in xs], np.int32)
    c.nrec = len(c.rec_locs)
    c.tab = oracle.sinc_table(jt)
    # model: smoothed noise, zero near the top
    m = rng.standard_normal((nz, ny, nx)).astype(np.float32)
    k = np.ones(3, np.float32) / 3
    for ax in range(3):
        m = np.apply_along_axis(lambda a: np.convolve(a, k, mode="same"), ax, m)
    m[: min(nz, 10)] = 0
    c.model = np.ascontiguousarray(m, np.float32)
    c.nbound = nbound
    return c


def oracle_prop(c, nthreads=0):
    P = oracle.Prop(nthreads)
    P.create_space(*c.d, c.nx, c.ny, c.nz)
    if c.nbound is not None:
        P.set_sponge(c.nbound)
    P.transfer_sinc_table_s(c.tab)
    P.transfer_sinc_table_d(c.tab)
    P.transfer_source_func(c.src_locs, c.src_vals)
    P.transfer_vel_func1(c.V1)
    P.transfer_vel_func2(c.V2)
    return P


def hip_prop(pyborn, c):
    P = pyborn.Propagator(0)
    P.create_space(*c.d, c.nx, c.ny, c.nz)
    if c.nbound is not None:
        P.set_sponge(c.nbound)
    P.transfer_sinc_table_s(c.tab)
    P.transfer_sinc_table_d(c.tab)
    P.transfer_source_func(c.src_locs, c.src_vals)
    P.transfer_vel_func1(c.V1)
    P.transfer_vel_func2(c.V2)
    return P


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
