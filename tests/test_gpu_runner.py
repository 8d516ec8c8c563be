"""The headless runner akua_run on the GPU, checked against the FP32 CPU oracle.

  * A trajectory written by the runner holds, at every written step, the oracle's state bit for
    bit (positions, velocities, masses, densities, lambdas and the logical order's ids).
  * The reference application's scene under its own loop (dt 0.016, frames of 0.04 s, at most 3
    steps per frame, Application.cpp:37-70) runs the number of steps the loop gives and lands on
    the oracle's state after that many steps.
  * Resuming from a written frame (--init) continues the trajectory bit for bit and keeps the
    original particle ids.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from akuaengine_amd import scenes
from akuaengine_amd import trajectory as T
from akuaengine_amd.config import LambdaCorrParams, PBFConfig, dam_break_params

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(*args):
    _say(f"[runner] akua_run {' '.join(args)}")
    r = subprocess.run([T.RUNNER, *args], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise AssertionError(f"akua_run exited {r.returncode}\n{r.stdout}\n{r.stderr}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def _assert_frame_equals_oracle(fr, o):
    r = o.state()
    for k in ("pos", "vel", "mass", "density", "lam", "ids"):
        a = fr[k]
        b = r[k].astype(a.dtype)
        if not np.array_equal(a, b):
            d = np.abs(a.astype(np.float64) - r[k])
            raise AssertionError(f"{k} differs: {np.count_nonzero(d)} entries, max |diff| {d.max():.3e}")


def test_runner_trajectory_matches_oracle(oracle_mod, tmp_path):
    path = str(tmp_path / "run.akt")
    out = _run("--scene", "dam_break_27k", "--steps", "6", "--every", "3", "--out", path, "--fields", "all")
    assert out["steps"] == 6 and out["frames_written"] == 3 and out["particles"] == 27000
    assert out["params"] == "dam_break" and out["particle_substeps_per_s"] > 0
    f = T.read_trajectory(path)
    assert len(f) == 3 and [f[k]["step"] for k in range(3)] == [0, 3, 6]
    sc = scenes.dam_break_27k()
    cfg, corr = dam_break_params(4)
    o = oracle_mod.Oracle(sc.n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel, sc.mass)
    _assert_frame_equals_oracle(f[0], o)
    for k in (1, 2):
        o.step(0.0083, sc.box_min, sc.box_max, 3)
        _assert_frame_equals_oracle(f[k], o)
        assert f[k]["time"] == pytest.approx(3 * k * float(np.float32(0.0083)), rel=1e-12)
    emax, _ = o.density_error(cfg.restDensity)
    assert out["max_density_error"] == pytest.approx(emax, rel=1e-5, abs=1e-9)


def test_runner_application_loop_matches_oracle(oracle_mod, tmp_path):
    path = str(tmp_path / "app.akt")
    out = _run("--scene", "app_dam_break", "--frames", "4", "--frame-time", "0.04", "--out", path,
               "--every", "1000", "--fields", "all")
    # frames of 0.04 s at dt 0.016 in float32: 2, 2, 3, 2 steps
    assert out["params"] == "reference" and out["steps"] == 9
    sc = scenes.app_dam_break()
    cfg, corr = PBFConfig(), LambdaCorrParams()
    o = oracle_mod.Oracle(sc.n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel, sc.mass)
    o.step(0.016, sc.box_min, sc.box_max, 9)
    f = T.read_trajectory(path)
    assert len(f) == 1  # only the initial frame (every 1000 steps)
    # the final state: a second run writing every step, last frame
    path2 = str(tmp_path / "app2.akt")
    _run("--scene", "app_dam_break", "--frames", "4", "--frame-time", "0.04", "--out", path2, "--fields", "all")
    f2 = T.read_trajectory(path2)
    assert len(f2) == 10 and f2[9]["step"] == 9
    _assert_frame_equals_oracle(f2[9], o)


def test_runner_resume_continues_bit_for_bit(tmp_path):
    full = str(tmp_path / "full.akt")
    _run("--scene", "cube:20", "--steps", "6", "--every", "3", "--out", full, "--fields", "all")
    part = str(tmp_path / "part.akt")
    out = _run("--init", full + "@1", "--steps", "3", "--every", "3", "--out", part, "--fields", "all")
    assert out["source"].endswith("@3") and out["steps"] == 3
    a, b = T.read_trajectory(full), T.read_trajectory(part)
    assert len(b) == 2 and b[1]["step"] == 6
    for k in ("pos", "vel", "mass", "density", "lam", "ids"):
        np.testing.assert_array_equal(b[1][k], a[2][k], err_msg=k)
    np.testing.assert_array_equal(b.by_id(1, "pos"), a.by_id(2, "pos"))


def test_runner_boundary_matches_oracle(oracle_mod, tmp_path):
    path = str(tmp_path / "bnd.akt")
    out = _run("--scene", "app_dam_break", "--steps", "8", "--every", "8", "--boundary", "2",
               "--out", path, "--fields", "all")
    assert out["boundary_layers"] == 2 and out["steps"] == 8
    sc = scenes.app_dam_break()
    o = oracle_mod.Oracle(sc.n, PBFConfig(), LambdaCorrParams(), 32)
    o.set_state(sc.pos, sc.vel, sc.mass)
    o.set_boundary(2)
    o.step(0.016, sc.box_min, sc.box_max, 8)
    f = T.read_trajectory(path)
    assert len(f) == 2 and f[1]["step"] == 8
    _assert_frame_equals_oracle(f[1], o)


def _read_ppm(path):
    data = open(path, "rb").read()
    parts = data.split(b"\n", 3)
    assert parts[0] == b"P6" and parts[2] == b"255"
    w, h = (int(v) for v in parts[1].split())
    return np.frombuffer(parts[3], np.uint8).reshape(h, w, 3)


def test_runner_images_equal_the_library_render(tmp_path):
    # frame 0 of the runner is the library's render of the scene as uploaded, with the
    # reference's camera and the application's sprite size
    from akuaengine_amd.solver import PBFSolver, default_camera
    out = _run("--scene", "app_dam_break", "--steps", "4", "--render-every", "2", "--render-dir",
               str(tmp_path), "--width", "320", "--height", "180")
    assert out["images"] == 3
    frames = sorted(p.name for p in tmp_path.iterdir())
    assert frames == ["frame_00000000.ppm", "frame_00000002.ppm", "frame_00000004.ppm"]
    sc = scenes.app_dam_break()
    s = PBFSolver(sc.n, PBFConfig(), LambdaCorrParams())
    s.set_state(sc.pos, sc.vel, sc.mass)
    img = s.render(default_camera(), 320, 180, point_size=50.0)
    np.testing.assert_array_equal(_read_ppm(str(tmp_path / frames[0])), img)
    s.step_n(0.016, sc.box_min, sc.box_max, 4)
    np.testing.assert_array_equal(_read_ppm(str(tmp_path / frames[2])), s.render(default_camera(), 320, 180))
