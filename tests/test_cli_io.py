"""The command-line programs around the Born path (sc17-born_amd/bin, built
from sc17-born_amd/cli): hypercube files in the genericIO JSON and SEP
formats, parameters from arguments and parameter files, and the reference's
util generators Wavelet, Vel, Data and Add (util/*.cc), checked against
their closed forms."""
import json
import os
import subprocess

import numpy as np
import pytest

import hcube

BIN = os.path.join(os.path.dirname(__file__), "..", "sc17-born_amd", "bin")


def run(prog, *args, cwd):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        pytest.fail("%s is not built (__graft_entry__.build())" % exe)
    return subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True)


def test_add_reads_json_and_big_endian_sep_and_writes_json(tmp_path):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(7 * 5 * 3).astype(np.float32)
    b = rng.standard_normal(7 * 5 * 3).astype(np.float32)
    axes = [(7, 0.0, 0.5), (5, 1.0, 2.0), (3, -1.0, 4.0)]
    hcube.write_json(str(tmp_path / "a.json"), a, axes)
    hcube.write_sep(str(tmp_path / "b.H"), b, [(7, 9.0, 9.0), (5, 0.0, 1.0), (3, 0.0, 1.0)])
    r = run("Add", "in1=a.json", "in2=b.H", "out=c.json", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    c, caxes = hcube.read(str(tmp_path / "c.json"))
    assert np.array_equal(c, a + b)
    assert caxes == [(n, pytest.approx(o), pytest.approx(d)) for n, o, d in axes]
    # SEP output (native floats), and out = in2 as the job script does
    r = run("Add", "in1=c.json", "in2=b.H", "out=b.H", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    bb, _ = hcube.read(str(tmp_path / "b.H"))
    assert np.array_equal(bb, (a + b) + b)


def test_add_rejects_inputs_of_different_sizes(tmp_path):
    hcube.write_json(str(tmp_path / "a.json"), np.zeros(6), [(6, 0, 1)])
    hcube.write_json(str(tmp_path / "b.json"), np.zeros(6), [(3, 0, 1), (2, 0, 1)])
    r = run("Add", "in1=a.json", "in2=b.json", "out=c.json", cwd=tmp_path)
    assert r.returncode == 1 and "differ in size" in r.stderr


def test_wavelet_is_the_ricker_of_the_reference(tmp_path):
    r = run("Wavelet", "out=w.json", "nt=100", "dt=.002", "fund=15", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    w, axes = hcube.read(str(tmp_path / "w.json"))
    assert axes == [(100, 0.0, pytest.approx(0.002))]
    t = -0.064 + np.arange(100) * np.float32(0.002)
    a = (np.pi * 15.0 * t) ** 2
    np.testing.assert_allclose(w, (1 - 2 * a) * np.exp(-a), atol=2e-6)
    assert w.argmax() == 32  # 0.064 s


def test_vel_and_parameter_files(tmp_path):
    # json= then par= then arguments: later settings win
    (tmp_path / "p.json").write_text(json.dumps({"nx": 6, "ny": 5, "nz": 4, "vconst": 1.5,
                                                 "out": "v.json", "nested": {"nx": 99}}))
    (tmp_path / "p.txt").write_text("# SEP parameters\nnz=3 dz=.02\n")
    r = run("Vel", "json=p.json", "par=p.txt", "vconst=2.5", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    v, axes = hcube.read(str(tmp_path / "v.json"))
    assert [a[0] for a in axes] == [6, 5, 3]
    assert axes[2][2] == pytest.approx(0.02)
    assert v.size == 90 and np.all(v == np.float32(2.5))


def test_data_spikes(tmp_path):
    r = run("Data", "out=d.H", "nt=120", "nx=45", "ny=41", "dt=.004", cwd=tmp_path)
    assert r.returncode == 0, r.stderr
    d, axes = hcube.read(str(tmp_path / "d.H"))
    assert [a[0] for a in axes] == [120, 45, 41, 1, 1]
    assert axes[3][1] == pytest.approx(0.5 + 0.01 * 45 / 2)
    d = d.reshape(41, 45, 120)
    want = np.zeros_like(d)
    want[20::20, 20::20, 30::50] = 1
    want[41 // 2, 45 // 2, 60] = 1
    assert np.array_equal(d, want)


def test_missing_parameters_are_reported(tmp_path):
    r = run("RTM3D", "velocity=v.json", cwd=tmp_path)
    assert r.returncode == 1 and "RTM3D:" in r.stderr
    r = run("Wavelet", cwd=tmp_path)
    assert r.returncode == 1 and "missing parameter out=" in r.stderr
