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


def gaussian_anomalies(shape, d, count, amp, radius, seed, window=None):
    """count Gaussian velocity anomalies exp(-(r/radius)^2) of +-amp m/s
    (radius in cells), centres and signs seeded; window=(x0, x1, y0, y1,
    z0, z1) restricts the centres to a sub-box of the grid."""
    rng = np.random.default_rng(seed)
    mz, my, mx = shape
    x0, x1, y0, y1, z0, z1 = window or (0, mx, 0, my, 0, mz)
    z = np.arange(mz, dtype=np.float32)[:, None, None]
    y = np.arange(my, dtype=np.float32)[None, :, None]
    x = np.arange(mx, dtype=np.float32)[None, None, :]
    dv = np.zeros(shape, np.float32)
    for _ in range(count):
        cx, cy, cz = rng.uniform(x0, x1), rng.uniform(y0, y1), rng.uniform(z0, z1)
        sign = 1.0 if rng.random() < 0.5 else -1.0
        r2 = ((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / np.float32(radius * radius)
        dv += np.float32(sign * amp) * np.exp(-r2)
    return dv


def make_problem(n=(256, 256, 256), d=10.0, nbound=20, nt=1500, dt=0.001, nshots=3,
                 rec=(128, 128), rec_d=20.0, rec_o=10.0, rec_depth=10.0, src_depth=20.0,
                 fpeak=12.0, delay=0.1, nw=200, seed=1002, model_seed=None, src_xy=None,
                 anomalies=None):
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
    if anomalies is not None:
        vel = (vel + gaussian_anomalies((mz, my, mx), d, **anomalies)).astype(np.float32)
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


class Shot:
    """One shot of a batch: its own padded window, V = v^2 dt^2, model,
    source stamp and receiver patch."""


def make_batch(nshots=32, nxy=(200, 320), nz=300, d=10.0, nbound=20, nt=2000, dt=0.001,
               rec=(96, 96), rec_d=10.0, rec_depth=10.0, src_depth=20.0, fpeak=12.0,
               delay=0.1, nw=200, seed=1004):
    """Shots on windowed subgrids: shot k's window is nx x ny x nz with nx, ny
    drawn uniformly from nxy (seed + k), its own velocity law v(z) (seeded
    surface velocity and gradient), a perturbation cut from one seeded
    smoothed-noise volume at a seeded offset and scaled per shot, the source
    at a seeded position in the window's central half and a receiver patch
    centred on the source."""
    H = host()
    big_rng = np.random.default_rng(seed)
    big = smoothed_noise(big_rng, (nz, nxy[1] + 64, nxy[1] + 64))
    w = ricker(nw, dt, fpeak, delay)
    wdev = np.zeros(nw, np.float32)
    H.born_host_wavelet_deriv(nw, np.float32(dt), w, wdev)
    tab = np.zeros((1, 8), np.float32)
    H.born_host_sinc_table(1, 8, tab)
    shots = []
    for k in range(nshots):
        rng = np.random.default_rng(seed + k)
        mx, my = (int(v) for v in rng.integers(nxy[0], nxy[1] + 1, size=2))
        mz = nz
        s = Shot()
        s.model_shape = (mz, my, mx)
        pad = nbound + 4
        s.nx, s.ny, s.nz = (H.born_host_padded_size(n, nbound, 4) for n in (mx, my, mz))
        s.o = -pad * d
        zm = np.arange(mz, dtype=np.float32) * np.float32(d)
        v0 = np.float32(1500 + rng.uniform(-50, 50))
        gz = np.float32(15 * rng.uniform(0.9, 1.1))
        vz = np.clip(v0 + gz * np.sqrt(zm), 1400, 3000).astype(np.float32)
        vel = np.ascontiguousarray(np.broadcast_to(vz[:, None, None], (mz, my, mx)), np.float32)
        s.V = np.zeros(s.nx * s.ny * s.nz, np.float32)
        H.born_host_rand_subspace(0, 0, 0, 0, np.float32(dt), mx, my, mz, 0.0, 0.0, 0.0, d, d,
                                  d, vel.ravel(), s.nx, s.ny, s.nz, s.o, s.o, s.o, d, d, d,
                                  s.V)
        oy, ox = (int(v) for v in rng.integers(0, 64, size=2))
        dv = np.ascontiguousarray(big[:, oy:oy + my, ox:ox + mx] * np.float32(rng.uniform(0.5, 1.5)))
        img = np.zeros(s.nx * s.ny * s.nz, np.float32)
        H.born_host_extract_sub(mx, my, mz, 0.0, 0.0, 0.0, d, d, d, dv.ravel(), s.nx, s.ny,
                                s.nz, s.o, s.o, s.o, img)
        s.model = np.zeros_like(img)
        H.born_host_laplac_forward(s.nx, s.ny, s.nz, img, s.model)
        sx = float(rng.uniform(0.25, 0.75) * (mx - 1) * d)
        sy = float(rng.uniform(0.25, 0.75) * (my - 1) * d)
        s.nt_big = nt + 8
        s.src_locs = np.zeros(9, np.int32)
        vals = np.zeros(9 * s.nt_big, np.float32)
        H.born_host_source_func(s.nx, s.ny, s.o, s.o, s.o, d, d, d, sx, sy, src_depth, wdev,
                                nw, s.nt_big, s.src_locs, vals)
        s.src_vals = vals.reshape(9, s.nt_big)
        s.rec_nx, s.rec_ny = rec
        s.nrec = rec[0] * rec[1]
        s.rec_locs = np.zeros(s.nrec, np.int32)
        H.born_host_receiver_locs(rec[0], rec[1], sx - rec_d * (rec[0] - 1) / 2,
                                  sy - rec_d * (rec[1] - 1) / 2, rec_d, rec_d, rec_depth,
                                  s.nx, s.ny, s.o, s.o, s.o, d, d, d, s.rec_locs)
        s.interior_cells = (s.nx - 8) * (s.ny - 8) * (s.nz - 8)
        shots.append(s)
    return shots, tab


# Benchmark configurations (sizes of the model before padding)
CONFIGS = {
    "c1": dict(n=(100, 100, 100), nt=800, nshots=1, rec=(50, 50), rec_d=20.0,
               rec_o=10.0, seed=1001, src_xy=[(495.0, 495.0)]),
    "c2": dict(n=(256, 256, 256), nt=1500, nshots=3, rec=(128, 128), rec_d=20.0,
               rec_o=10.0, seed=1002),
}

# Operator-level form of the 256^3 pair (born_op_* / pyborn.BornOp): the
# velocity model and the rtm_zero_op parameters of the C2 problem above (the
# shot positions are make_problem's draws for the same seed)
def op_config(name="c2", mode=1, scatter=0, storage="f32"):
    cfg = CONFIGS[name]
    mx, my, mz = cfg["n"]
    d = 10.0
    vz = velocity_law(mz, d)
    vel = np.ascontiguousarray(np.broadcast_to(vz[:, None, None], (mz, my, mx)), np.float32)
    rng = np.random.default_rng(cfg["seed"])
    sx, sy = [], []
    for k in range(cfg["nshots"]):
        if "src_xy" in cfg:
            x, y = cfg["src_xy"][k]
        else:
            x = float(rng.uniform(0.25, 0.75) * (mx - 1) * d)
            y = float(rng.uniform(0.25, 0.75) * (my - 1) * d)
        sx.append(x)
        sy.append(y)
    kw = dict(vo=(0.0, 0.0, 0.0), vd=(d, d, d), wavelet=ricker(200, 0.001, 12.0, 0.1),
              wdt=0.001, src_x=np.array(sx, np.float32), src_y=np.array(sy, np.float32),
              src_depth=20.0, nt=cfg["nt"], t0=0.0, dt=0.001, rec_n=cfg["rec"],
              rec_o=(cfg["rec_o"], cfg["rec_o"]), rec_d=(cfg["rec_d"], cfg["rec_d"]),
              rec_depth=10.0, nbound=20, mode=mode, scatter=scatter, storage=storage)
    return vel, kw


# FP16 error study: the 512x512x400 grid's velocity (v(z) plus 4 seeded
# Gaussian anomalies of +-300 m/s, radius 40 cells), 2500 steps at 0.8 ms,
# receivers every 20 m, restricted to a 128^3 crop around the shot (the FP64
# oracle's share); the anomaly centres are drawn inside the crop.
C3_CROP = dict(n=(128, 128, 128), nt=2500, dt=0.0008, nshots=1, rec=(64, 64), rec_d=20.0,
               rec_o=10.0, seed=1003, src_xy=[(635.0, 635.0)],
               anomalies=dict(count=4, amp=300.0, radius=40.0, seed=1003))

# Batched shots: 32 windows of [200, 320]^2 x 300 cells at 10 m, 2000 steps
BATCHES = {
    "c4": dict(nshots=32, nxy=(200, 320), nz=300, nt=2000, rec=(96, 96), seed=1004),
}
