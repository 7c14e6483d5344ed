"""16-bit-based storage error on the 512x512x400 configuration's physics, on
a 128^3 crop (the FP64 oracle's share): data, image and gradient of the Born
pair against the FP64 oracle, 2500 steps at 0.8 ms, for FP32 levels, FP16
levels (an approximate mode) and 24-bit levels (FP32 rounded to 24 bits,
held to the 1e-3 bar).  The oracle takes a few minutes of CPU time.

  data     F m                    (m: smoothed noise, z-differenced)
  image    F^T d64                (d64: the FP64 data)
  gradient F^T (F m0 - d64) at m0 = m / 2; in FP64 it equals -image64 / 2
"""
import json
import os

import numpy as np
import pytest

import oracle
from cases import rel_l2
from conftest import load_pkg_module

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_fp16_errors_on_a_128_cube_crop(pyborn):
    os.environ["ORC_PROGRESS"] = "1"  # the oracle loops report their steps
    configs = load_pkg_module("configs")
    p = configs.make_problem(**configs.C3_CROP)
    nt = p.nt
    locs, vals = p.shots[0]
    env = os.environ.get("OMP_NUM_THREADS")
    P = oracle.Prop(int(env) if env else len(os.sched_getaffinity(0)))
    P.create_space(p.d, p.d, p.d, p.nx, p.ny, p.nz)
    P.transfer_sinc_table_s(p.tab)
    P.transfer_sinc_table_d(p.tab)
    P.transfer_source_func(locs, vals)
    P.transfer_vel_func1(p.V)
    P.transfer_vel_func2(p.V)
    P.transfer_receiver_func(p.rec_locs, np.zeros((p.nrec, p.nt_big), np.float32))
    d64 = P.born_forward_d(1, p.model, nt, scatter=0)
    img64 = P.exact_adjoint_d(1, d64, nt, scatter=0).ravel()
    g64 = -0.5 * img64
    d32in = d64.astype(np.float32)
    m0 = (0.5 * p.model).astype(np.float32)
    out = {"domain": [p.nx, p.ny, p.nz], "steps": nt}
    for storage in ("f32", "f16", "f24"):
        H = pyborn.Propagator(0)
        H.create_space(p.d, p.d, p.d, p.nx, p.ny, p.nz)
        H.transfer_sinc_table_s(p.tab)
        H.transfer_sinc_table_d(p.tab)
        H.transfer_source_func(locs, vals)
        H.transfer_vel_func1(p.V)
        H.transfer_vel_func2(p.V)
        H.set_storage(storage)
        H.transfer_receiver_func(p.rec_nx, p.rec_ny, p.rec_locs,
                                 np.zeros((p.nrec, p.nt_big), np.float32))
        d = np.zeros((p.nrec, p.nt_big), np.float32)
        H.forward_scatter(1, p.model, d, nt, 0)
        img = np.zeros(p.nx * p.ny * p.nz, np.float32)
        H.adjoint_exact(1, d32in, img, nt, scatter=0)
        dm0 = np.zeros_like(d)
        H.forward_scatter(1, m0, dm0, nt, 0)
        g = np.zeros_like(img)
        H.adjoint_exact(1, (dm0.astype(np.float64) - d64).astype(np.float32), g, nt, scatter=0)
        H.close()
        out[storage] = {"data": rel_l2(d, d64), "image": rel_l2(img, img64),
                        "gradient": rel_l2(g, g64)}
        print(storage, out[storage], flush=True)
    print("crop study", json.dumps(out))
    assert max(out["f32"].values()) <= 1e-4
    # FP16 levels are an approximate mode: rounding noise injected at every
    # step grows with the step count (measured 9.5e-3 / 4.4e-3 / 6.9e-3)
    assert max(out["f16"].values()) <= 1.5e-2
    assert max(out["f24"].values()) <= 1e-3
