"""Seeded synthetic Born problems for the benchmark configurations.

Each problem is built with the host preparation library (libborn_host.so):
domain sizing, v^2 dt^2 over the padded domain, the stamped Ricker derivative
per shot, receiver cells and the sinc table -- the inputs the reference's
rtm_zero_op hands to its propagator (base/rtm_zero_op_3d.cc:137-185).

Velocity law v(z) = clip(1500 + 15 sqrt(z_m), 1500, 3000); model
perturbations are seeded N(0, 50^2) noise, 3-cell box smoothed, zeroed in
the top 10 cells.  Random numbers come from numpy's PCG64 with the
configuration's seed.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HOST_LIB = os.path.join(_HERE, "libborn_host.so")
_H = None

f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def host():
    global _H
    if _H is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("%s is not built" % HOST_LIB)
        H = C.CDLL(HOST_LIB)
        H.born_host_sinc_table.argtypes = [C.c_int, C.c_int, f32p]
        H.born_host_padded_size.restype = C.c_int
        H.born_host_padded_size.argtypes = [C.c_int] * 3
        H.born_host_wavelet_deriv.argtypes = [C.c_int, C.c_float, f32p, f32p]
        H.born_host_source_func.argtypes = [C.c_int, C.c_int] + [C.c_float] * 9 + [
            f32p, C.c_int, C.c_int, i32p, f32p]
        H.born_host_rand_subspace.restype = C.c_float
        H.born_host_rand_subspace.argtypes = [C.c_int] * 4 + [C.c_float] + [C.c_int] * 3 + [
            C.c_float] * 6 + [f32p] + [C.c_int] * 3 + [C.c_float] * 6 + [f32p]
        H.born_host_receiver_locs.argtypes = [C.c_int, C.c_int] + [C.c_float] * 5 + [
            C.c_int, C.c_int] + [C.c_float] * 6 + [i32p]
        H.born_host_calc_stability.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int,
                                               C.POINTER(C.c_float), C.POINTER(C.c_int)]
        for name in ("born_host_laplac_forward", "born_host_laplac_adjoint"):
            getattr(H, name).argtypes = [C.c_int] * 3 + [f32p, f32p]
        for name in ("born_host_extract_sub", "born_host_add_image"):
            getattr(H, name).argtypes = [C.c_int] * 3 + [C.c_float] * 6 + [f32p] + [
                C.c_int] * 3 + [C.c_float] * 3 + [f32p]
        _H = H
    return _H


def ricker(n, dt, fpeak, delay):
    t = np.arange(n, dtype=np.float64) * dt - delay
    a = (np.pi * fpeak * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def velocity_law(nz, dz):
    z = np.arange(nz, dtype=np.float32) * np.float32(dz)
    return np.clip(1500 + 15 * np.sqrt(z), 1500, 3000).astype(np.float32)


def smoothed_noise(rng, shape, sigma=50.0, ztop=10):
    from scipy.ndimage import uniform_filter

    dv = rng.standard_normal(shape, dtype=np.float32) * np.float32(sigma)
    dv = uniform_filter(dv, size=3, mode="constant")
    dv[:ztop] = 0
    return dv.astype(np.float32)


class Problem:
    """A single-GPU Born workload: padded domain, vnone = v^2 dt^2, shots,
    receivers, sinc table and a domain-size model image."""


def make_problem(n=(256, 256, 256), d=10.0, nbound=20, nt=1500, dt=0.001, nshots=3,
                 rec=(128, 128), rec_d=20.0, rec_o=10.0, rec_depth=10.0, src_depth=20.0,
                 fpeak=12.0, delay=0.1, nw=200, seed=1002, model_seed=None, src_xy=None):
    H = host()
    rng = np.random.default_rng(seed)
    p = Problem()
    mx, my, mz = n
    p.model_shape = (mz, my, mx)
    p.d = d
    p.nt = nt
    p.dt = np.float32(dt)
    pad = nbound + 4
    p.nx, p.ny, p.nz = (H.born_host_padded_size(k, nbound, 4) for k in (mx, my, mz))
    p.o = -pad * d
    vz = velocity_law(mz, d)
    vel = np.ascontiguousarray(np.broadcast_to(vz[:, None, None], (mz, my, mx)), np.float32)
    p.vmax = float(vel.max())
    # vnone: the propagation velocity of the Born fields (rtm_zero_op_3d.cc:157)
    p.V = np.zeros(p.nx * p.ny * p.nz, np.float32)
    H.born_host_rand_subspace(0, 0, 0, 0, p.dt, mx, my, mz, 0.0, 0.0, 0.0, d, d, d,
                              vel.ravel(), p.nx, p.ny, p.nz, p.o, p.o, p.o, d, d, d, p.V)
    # wavelet derivative and shots at seeded surface positions
    w = ricker(nw, dt, fpeak, delay)
    p.wdev = np.zeros(nw, np.float32)
    H.born_host_wavelet_deriv(nw, p.dt, w, p.wdev)
    p.nt_big = nt + 8  # rtm_zero_op_3d.cc:114 with jt = 1
    p.shots = []
    for k in range(nshots):
        if src_xy is not None:
            sx, sy = src_xy[k]
        else:
            sx = float(rng.uniform(0.25, 0.75) * (mx - 1) * d)
            sy = float(rng.uniform(0.25, 0.75) * (my - 1) * d)
        locs = np.zeros(9, np.int32)
        vals = np.zeros(9 * p.nt_big, np.float32)
        H.born_host_source_func(p.nx, p.ny, p.o, p.o, p.o, d, d, d, sx, sy, src_depth,
                                p.wdev, nw, p.nt_big, locs, vals)
        p.shots.append((locs, vals.reshape(9, p.nt_big)))
    p.rec_nx, p.rec_ny = rec
    p.nrec = rec[0] * rec[1]
    p.rec_locs = np.zeros(p.nrec, np.int32)
    H.born_host_receiver_locs(rec[0], rec[1], rec_o, rec_o, rec_d, rec_d, rec_depth,
                              p.nx, p.ny, p.o, p.o, p.o, d, d, d, p.rec_locs)
    p.tab = np.zeros((1, 8), np.float32)
    H.born_host_sinc_table(1, 8, p.tab)
    # model: smoothed dv, windowed into the domain, vertical difference
    # (rtm_zero_op_3d.cc:161-180)
    mrng = np.random.default_rng(seed if model_seed is None else model_seed)
    dv = smoothed_noise(mrng, (mz, my, mx))
    img = np.zeros(p.nx * p.ny * p.nz, np.float32)
    H.born_host_extract_sub(mx, my, mz, 0.0, 0.0, 0.0, d, d, d, dv.ravel(), p.nx, p.ny,
                            p.nz, p.o, p.o, p.o, img)
    p.model = np.zeros_like(img)
    H.born_host_laplac_forward(p.nx, p.ny, p.nz, img, p.model)
    p.interior_cells = (p.nx - 8) * (p.ny - 8) * (p.nz - 8)
    return p


# Benchmark configurations (sizes of the model before padding)
CONFIGS = {
    "c1": dict(n=(100, 100, 100), nt=800, nshots=1, rec=(50, 50), rec_d=20.0,
               rec_o=10.0, seed=1001, src_xy=[(495.0, 495.0)]),
    "c2": dict(n=(256, 256, 256), nt=1500, nshots=3, rec=(128, 128), rec_d=20.0,
               rec_o=10.0, seed=1002),
}
