#!/usr/bin/env python3
"""Least-squares RTM with the Born operator on one MI355X.

The reference runs lin_solver + cgstep (base/lin_solver.cc:104-131,
base/cgstep.cc:11-54) over rtm_zero_op (base/rtm_zero_op_3d.cc).  This driver
does the same through the operator handle of libborn_host.so (born_op_lsrtm):
the exact adjoint (mode 1) so that CG sees a true transpose, the CG scalars
in FP64, the background rebuilt from HBM checkpoints when its history does
not fit.

Synthetic problem (default: the 640x640x500 LSRTM configuration at half the
size per axis): background v(z) = clip(1500 + 15 sqrt(z), 1500, 3000); true
model = a layered perturbation with 20 seeded interfaces of +-100 m/s; shots
at seeded surface positions, a surface receiver grid.  Observed data are the
Born data of the true model, so the residual can go to zero.

Prints one JSON line: residual norm per iteration over the data norm,
model error against the true model, seconds per iteration.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    spec = importlib.util.spec_from_file_location("sc17born_" + name,
                                                  os.path.join(_HERE, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


def layered_model(shape, d, ninterfaces, amp, seed):
    """A perturbation constant between seeded depths, each interface adding
    a seeded jump of +-amp m/s (zero above the first interface)."""
    rng = np.random.default_rng(seed)
    mz = shape[0]
    depths = np.sort(rng.integers(10, mz - 5, size=ninterfaces))
    jumps = rng.choice([-amp, amp], size=ninterfaces) * rng.uniform(0.5, 1.0, ninterfaces)
    prof = np.zeros(mz, np.float32)
    for z, j in zip(depths, jumps):
        prof[z:] += np.float32(j)
    return np.ascontiguousarray(np.broadcast_to(prof[:, None, None], shape), np.float32)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, nargs=3, default=[320, 320, 250], help="model x y z")
    ap.add_argument("--d", type=float, default=10.0)
    ap.add_argument("--shots", type=int, default=8)
    ap.add_argument("--rec", type=int, nargs=2, default=[160, 160])
    ap.add_argument("--nt", type=int, default=1500)
    ap.add_argument("--dt", type=float, default=0.001)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16"])
    ap.add_argument("--seed", type=int, default=1005)
    args = ap.parse_args()

    # host buffers only: no torch in this process, so the library brings
    # the only HIP runtime (and a fresh machine skips torch's start-up)
    pyborn = _load("pyborn")
    configs = _load("configs")
    mx, my, mz = args.n
    d = args.d
    vz = configs.velocity_law(mz, d)
    vel = np.ascontiguousarray(np.broadcast_to(vz[:, None, None], (mz, my, mx)), np.float32)
    m_true = layered_model((mz, my, mx), d, 20, 100.0, args.seed)
    rng = np.random.default_rng(args.seed)
    sx = rng.uniform(0.25, 0.75, args.shots) * (mx - 1) * d
    sy = rng.uniform(0.25, 0.75, args.shots) * (my - 1) * d
    w = configs.ricker(200, args.dt, 12.0, 0.1)
    rd = (mx - 1) * d / args.rec[0]
    op = pyborn.BornOp(vel, (0.0, 0.0, 0.0), (d, d, d), w, args.dt, sx, sy, 20.0, args.nt, 0.0,
                       args.dt, tuple(args.rec), (rd / 2, rd / 2), (rd, rd), 10.0, nbound=20,
                       mode=1, scatter=0, storage=args.storage)
    print("[lsrtm] domain %s, %d shots, %d steps" % (op.domain, args.shots, op.nt),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    data = op.forward(m_true)
    t_fwd = time.perf_counter() - t0
    print("[lsrtm] observed data in %.1f s" % t_fwd, file=sys.stderr, flush=True)
    os.environ.setdefault("BORN_VERBOSE", "1")
    t0 = time.perf_counter()
    m, res = op.lsrtm(data, args.iters)
    t_cg = time.perf_counter() - t0
    op.close()
    dn = float(np.linalg.norm(data.astype(np.float64)))
    merr = float(np.linalg.norm((m - m_true).astype(np.float64)) /
                 np.linalg.norm(m_true.astype(np.float64)))
    print(json.dumps({
        "workload": "LSRTM, %dx%dx%d model at %g m, %d shots, %dx%d receivers, %d steps, "
                    "%d CG iterations, %s storage" % (mx, my, mz, d, args.shots, args.rec[0],
                                                      args.rec[1], args.nt, args.iters,
                                                      args.storage),
        # cgstep reports rr.rr after each step: |r| / |d|
        "residual_over_data": [round(float(np.sqrt(r)) / dn, 6) if r >= 0 else None
                               for r in res],
        "model_error_rel_l2": round(merr, 6),
        "seconds_per_iteration": round(t_cg / max(args.iters, 1), 2),
        "seconds_forward": round(t_fwd, 2),
    }), flush=True)


if __name__ == "__main__":
    main()
