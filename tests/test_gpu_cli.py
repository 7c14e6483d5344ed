"""Model3D and RTM3D (the reference's model.cc and rtm.cc) on hypercube
files: the programs build the operator from the files' axes and
parameters exactly as the library's own callers do, so their data and
images equal those of born_op_forward / born_op_adjoint on the same inputs
bit for bit (the operator itself is pinned to the oracle by
test_gpu_operator.py)."""
import os
import subprocess

import numpy as np
import pytest

import hcube
from conftest import load_pkg_module

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(__file__), "..", "sc17-born_amd", "bin")


def run(prog, *args, cwd):
    r = subprocess.run([os.path.join(BIN, prog), *args], cwd=cwd, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    nx, ny, nz, h = 44, 40, 36, 10.0
    z = np.arange(nz, dtype=np.float32) * h
    vel = np.broadcast_to((1500 + 15 * np.sqrt(z))[:, None, None], (nz, ny, nx))
    vel = np.ascontiguousarray(vel, np.float32)
    rng = np.random.default_rng(4)
    img = np.zeros((nz, ny, nx), np.float32)
    img[10:, :, :] = rng.standard_normal((nz - 10, ny, nx)).astype(np.float32) * 20
    hcube.write_json(str(d / "velocity.json"), vel, [(nx, 0.0, h), (ny, 0.0, h), (nz, 0.0, h)])
    hcube.write_json(str(d / "image.json"), img, [(nx, 0.0, h), (ny, 0.0, h), (nz, 0.0, h)])
    run("Wavelet", "out=wavelet.json", "nt=60", "dt=.002", "fund=12", cwd=d)
    # 90 samples at 2 ms, 10 x 8 receivers every 30 m, 2 x 1 shots
    nt, dt = 90, 0.002
    axes = [(nt, 0.0, dt), (10, 60.0, 30.0), (8, 60.0, 30.0), (2, 150.0, 120.0), (1, 200.0, 10.0)]
    hcube.write_json(str(d / "data.json"), np.zeros(nt * 10 * 8 * 2, np.float32), axes)
    return d, vel, img, axes


def _op(pyborn, d, vel, axes, mode):
    w, waxes = hcube.read(str(d / "wavelet.json"))
    (nt, t0, dt), (rnx, rox, rdx), (rny, roy, rdy), (snx, sox, sdx), (sny, soy, sdy) = axes
    sx = [sox + sdx * i for j in range(sny) for i in range(snx)]
    sy = [soy + sdy * j for j in range(sny) for i in range(snx)]
    return pyborn.BornOp(vel, (0.0, 0.0, 0.0), (10.0, 10.0, 10.0), w, waxes[0][2], sx, sy, 20.0,
                         nt, t0, dt, (rnx, rny), (rox, roy), (rdx, rdy), 10.0, nbound=10,
                         mode=mode)


def test_model3d_equals_the_operator_forward(files):
    d, vel, img, axes = files
    pyborn = load_pkg_module("pyborn")
    run("Model3D", "velocity=velocity.json", "wavelet=wavelet.json", "image=image.json",
        "data=data.json", "src_depth=20", "rec_depth=10", "nbound=10", cwd=d)
    got, gaxes = hcube.read(str(d / "data.json"))
    assert [a[0] for a in gaxes] == [a[0] for a in axes]
    op = _op(pyborn, d, vel, axes, 0)
    want = op.forward(img)
    op.close()
    assert np.abs(want).max() > 0
    assert np.array_equal(got, want.ravel())


@pytest.mark.parametrize("mode", [0, 1])
def test_rtm3d_equals_the_operator_adjoint(files, mode):
    d, vel, img, axes = files
    pyborn = load_pkg_module("pyborn")
    data = np.random.default_rng(6).standard_normal(90 * 10 * 8 * 2).astype(np.float32)
    hcube.write_json(str(d / "resid.json"), data, axes)
    run("RTM3D", "velocity=velocity.json", "wavelet=wavelet.json", "data=resid.json",
        "image=mig.json", "src_depth=20", "rec_depth=10", "nbound=10", "mode=%d" % mode, cwd=d)
    got, gaxes = hcube.read(str(d / "mig.json"))
    assert [a[0] for a in gaxes] == [44, 40, 36]
    op = _op(pyborn, d, vel, axes, mode)
    want = op.adjoint(data)
    op.close()
    assert np.abs(want).max() > 0
    assert np.array_equal(got, want.ravel())
