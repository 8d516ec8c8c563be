"""Scenes, trajectory files and the runner's host logic (no GPU).

  * libakua.so's scene generators (akua_scene_generate) equal the Python generators bit for bit,
    and the reference application's own scene matches known answers from Application.cpp.
  * Trajectory files written by the library read back identically through the library's reader
    and through the numpy reader, whose byte layout is checked against the specification in
    include/akua_io.h; malformed and truncated files are reported.
  * akua_run: the step plan of --frames/--frame-time equals the Application loop
    (Application.cpp:37-70) restated here in float32, and --dry-run writes the initial frame.
"""
import json
import os
import struct
import subprocess

import numpy as np
import pytest

from akuaengine_amd import _lib, scenes
from akuaengine_amd import trajectory as T


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


@pytest.fixture(scope="module")
def lib():
    return _lib.load()


# ---- scenes -------------------------------------------------------------------------------------

@pytest.mark.parametrize("name,make", [
    ("app_dam_break", lambda: scenes.app_dam_break()),
    ("dam_break_27k", lambda: scenes.dam_break_27k()),
    ("dam_break_1m", lambda: scenes.dam_break_1m()),
    ("cube:20", lambda: scenes.dam_break_scaled(20)),
    ("cube:7", lambda: scenes.dam_break_scaled(7, seed=5)),
], ids=["app", "27k", "1m", "cube20", "cube7_seed5"])
def test_native_scene_equals_python(lib, name, make):
    py = make()
    seed = 5 if name == "cube:7" else -1
    pos, info = T.native_scene(name, seed=seed)
    assert info.numParticles == py.n
    np.testing.assert_array_equal(_bits(pos), _bits(py.pos))
    np.testing.assert_array_equal(np.array(info.boxMin, np.float32), np.float32(py.box_min))
    np.testing.assert_array_equal(np.array(info.boxMax, np.float32), np.float32(py.box_max))


@pytest.mark.slow
def test_native_8m_scene_equals_python(lib):
    py = scenes.dam_break_8m()
    pos, info = T.native_scene("dam_break_8m")
    assert info.numParticles == 8_000_000
    np.testing.assert_array_equal(_bits(pos), _bits(py.pos))


def test_app_dam_break_known_answers(lib):
    # Application.cpp:159-182 with the default particle_spacing 0.05: numX = int(1.5f / 0.05f)
    # = 30 per axis (27,000 = NUM_PARTICLES, Application.cpp:12); position = minPos + i * s.
    pos, info = T.native_scene("app_dam_break")
    assert info.numParticles == 27000 and info.referenceParams == 1
    assert np.float32(info.deltaTime) == np.float32(0.016)
    assert tuple(info.boxMin) == (1.5, 0.0, 1.5) and tuple(info.boxMax) == (4.5, 4.0, 4.5)
    s = np.float32(0.05)
    np.testing.assert_array_equal(pos[0], np.float32([2.0, 1.0, 2.0]))
    last = np.float32([2.0, 1.0, 2.0]) + np.float32(29) * s
    np.testing.assert_array_equal(pos[-1], last)
    # z runs fastest, then y, then x (the loop nest x { y { z } })
    np.testing.assert_array_equal(pos[1], np.float32([2.0, 1.0, 2.0 + np.float32(1) * s]))
    np.testing.assert_array_equal(pos[30], np.float32([2.0, 1.0 + np.float32(1) * s, 2.0]))


def test_app_dam_break_spacing_by_plain_loops(lib):
    # the reference's loop nest written out literally in float32, at a spacing giving 6^3
    s = np.float32(0.25)
    n = int(np.float32(np.float32(1.5) / s))
    want = []
    for x in range(n):
        for y in range(n):
            for z in range(n):
                want.append([np.float32(2.0) + np.float32(x) * s, np.float32(1.0) + np.float32(y) * s,
                             np.float32(2.0) + np.float32(z) * s])
    pos, info = T.native_scene("app_dam_break", spacing=0.25)
    assert info.numParticles == n ** 3 == 216
    np.testing.assert_array_equal(_bits(pos), _bits(np.array(want, np.float32)))
    np.testing.assert_array_equal(_bits(pos), _bits(scenes.app_dam_break(0.25).pos))
    # 1.5f / 0.04f = 37.5 in float32: 37 per axis
    _, info = T.native_scene("app_dam_break", spacing=0.04)
    assert info.numParticles == 37 ** 3


def test_unknown_scene_and_wrong_count(lib):
    with pytest.raises(_lib.AkuaError, match="unknown scene"):
        T.native_scene("no_such_scene")
    with pytest.raises(_lib.AkuaError, match="cube"):
        T.native_scene("cube:0")
    pos = np.empty((10, 3), np.float32)
    st = lib.akua_scene_generate(b"dam_break_27k", 0.0, -1, pos.ctypes.data, 10)
    assert st == 1 and b"27000" in lib.akua_last_error()


# ---- trajectory files ---------------------------------------------------------------------------

def _random_frames(n, k, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for f in range(k):
        out.append(dict(step=10 * f, time=0.5 * f, pos=rng.standard_normal((n, 3)).astype(np.float32),
                        vel=rng.standard_normal((n, 3)).astype(np.float32),
                        mass=rng.random(n).astype(np.float32), density=rng.random(n).astype(np.float32),
                        lam=rng.standard_normal(n).astype(np.float32),
                        ids=rng.permutation(n).astype(np.uint32)))
    return out


def _write(path, frames, fields=("all",), n=None):
    n = n or len(frames[0]["pos"])
    with T.TrajectoryWriter(path, n, fields, dt=0.0083, box_min=(0, 0, 0), box_max=(2, 1.5, 1),
                            gravity=(0, -9.8, 0)) as w:
        for fr in frames:
            w.append(fr["step"], fr["time"], fr["pos"], fr["vel"], fr["mass"], fr["density"], fr["lam"], fr["ids"])


def test_trajectory_round_trip_all_fields(lib, tmp_path):
    path = str(tmp_path / "t.akt")
    frames = _random_frames(257, 4)
    _write(path, frames)
    f = T.read_trajectory(path)
    assert len(f) == 4 and not f.info.partial_frame and f.info.n == 257 and f.info.fields == T.ALL
    assert f.info.dt == float(np.float32(0.0083)) and f.info.box_max == (2.0, 1.5, 1.0)
    with T.TrajectoryReader(path) as r:
        assert r.frames == 4 and r.n == 257
        for k, fr in enumerate(frames):
            a, b = f[k], r.read(k)
            for name in ("pos", "vel", "mass", "density", "lam", "ids"):
                np.testing.assert_array_equal(a[name], fr[name])
                np.testing.assert_array_equal(b[name], fr[name])
            assert a["step"] == b["step"] == fr["step"] and a["time"] == b["time"] == fr["time"]
        np.testing.assert_array_equal(r.read(-1)["pos"], frames[-1]["pos"])
        with pytest.raises(_lib.AkuaError):
            r.read(4)
    # by original index
    fr = frames[2]
    want = np.empty_like(fr["pos"])
    want[fr["ids"]] = fr["pos"]
    np.testing.assert_array_equal(f.by_id(2, "pos"), want)


def test_trajectory_byte_layout(lib, tmp_path):
    # the offsets documented in include/akua_io.h
    path = str(tmp_path / "t.akt")
    frames = _random_frames(5, 1, seed=3)
    _write(path, frames, fields=("pos", "density", "id"))
    raw = open(path, "rb").read()
    assert raw[0:8] == b"AKUATRJ\0"
    version, hb, n, fields = struct.unpack_from("<IIiI", raw, 8)
    assert (version, hb, n, fields) == (1, 128, 5, T.POSITION | T.DENSITY | T.ID)
    assert struct.unpack_from("<f", raw, 24)[0] == np.float32(0.0083)
    assert struct.unpack_from("<3f", raw, 40) == (2.0, 1.5, 1.0)
    assert struct.unpack_from("<3f", raw, 52)[1] == np.float32(-9.8)
    fb = struct.unpack_from("<Q", raw, 64)[0]
    assert fb == 16 + 5 * (12 + 4 + 4) and len(raw) == 128 + fb
    assert raw[72:128] == bytes(56)
    step, time = struct.unpack_from("<qd", raw, 128)
    assert (step, time) == (0, 0.0)
    pos = np.frombuffer(raw, "<f4", 15, 144).reshape(5, 3)
    dens = np.frombuffer(raw, "<f4", 5, 144 + 60)
    ids = np.frombuffer(raw, "<u4", 5, 144 + 80)
    np.testing.assert_array_equal(pos, frames[0]["pos"])
    np.testing.assert_array_equal(dens, frames[0]["density"])
    np.testing.assert_array_equal(ids, frames[0]["ids"])


def test_trajectory_field_subsets_and_missing_arrays(lib, tmp_path):
    path = str(tmp_path / "t.akt")
    fr = _random_frames(9, 1)[0]
    with T.TrajectoryWriter(path, 9, ("pos",)) as w:
        w.append(1, 0.1, pos=fr["pos"])
        with pytest.raises(_lib.AkuaError):
            w.append(2, 0.2)                       # the file holds positions: NULL is an error
    f = T.read_trajectory(path)
    assert set(f[0]) == {"step", "time", "pos"}
    with T.TrajectoryReader(path) as r:
        assert set(r.read(0)) == {"step", "time", "pos"}
        vel = np.empty((9, 3), np.float32)
        st = lib.akua_traj_read(r._h, 0, None, None, None, vel.ctypes.data, None, None, None, None)
        assert st == 1 and b"no velocity" in lib.akua_last_error()
    with pytest.raises(_lib.AkuaError):
        T.TrajectoryWriter(str(tmp_path / "z.akt"), 9, ())         # no fields
    with pytest.raises(_lib.AkuaError):
        T.TrajectoryWriter(str(tmp_path / "z.akt"), 0, ("pos",))   # no particles


def test_trajectory_partial_frame_and_bad_files(lib, tmp_path):
    path = str(tmp_path / "t.akt")
    _write(path, _random_frames(33, 3))
    raw = open(path, "rb").read()
    cut = str(tmp_path / "cut.akt")
    open(cut, "wb").write(raw[:-100])              # a writer that stopped inside frame 2
    f = T.read_trajectory(cut)
    assert len(f) == 2 and f.info.partial_frame
    with T.TrajectoryReader(cut) as r:
        assert r.frames == 2 and r.partial_frame
    bad = str(tmp_path / "bad.akt")
    open(bad, "wb").write(b"NOTATRAJ" + raw[8:])
    with pytest.raises(T.TrajectoryFormatError, match="magic"):
        T.read_trajectory(bad)
    with pytest.raises(_lib.AkuaError, match="magic"):
        T.TrajectoryReader(bad)
    v2 = str(tmp_path / "v2.akt")
    open(v2, "wb").write(raw[:8] + struct.pack("<I", 2) + raw[12:])
    with pytest.raises(T.TrajectoryFormatError, match="version"):
        T.read_trajectory(v2)
    with pytest.raises(_lib.AkuaError) as e:
        T.TrajectoryReader(v2)
    assert e.value.status == 6                    # AKUA_ERR_UNSUPPORTED
    inc = str(tmp_path / "inc.akt")
    open(inc, "wb").write(raw[:64] + struct.pack("<Q", 12345) + raw[72:])
    with pytest.raises(T.TrajectoryFormatError, match="inconsistent"):
        T.read_trajectory(inc)
    with pytest.raises(_lib.AkuaError, match="inconsistent"):
        T.TrajectoryReader(inc)
    short = str(tmp_path / "short.akt")
    open(short, "wb").write(raw[:50])
    with pytest.raises(T.TrajectoryFormatError):
        T.read_trajectory(short)
    with pytest.raises(_lib.AkuaError):
        T.TrajectoryReader(short)


# ---- the runner (host logic) --------------------------------------------------------------------

def _run(*args, check=True):
    assert os.path.exists(T.RUNNER), "akua_run is built by akuaengine_amd/build.py"
    r = subprocess.run([T.RUNNER, *args], capture_output=True, text=True, timeout=120)
    if check and r.returncode != 0:
        raise AssertionError(f"akua_run {args} -> {r.returncode}\n{r.stderr}")
    return r


def _application_plan(frames, frame_time, dt, max_steps):
    """Application::run (Application.cpp:37-70) with a fixed frame time, in float32."""
    acc = np.float32(0.0)
    ft, d = np.float32(frame_time), np.float32(dt)
    plan = []
    for _ in range(frames):
        acc = np.float32(acc + ft)
        k = 0
        while acc >= d and k < max_steps:
            acc = np.float32(acc - d)
            k += 1
        plan.append(k)
    return plan


def test_runner_usage_errors():
    assert _run("--help").returncode == 0
    r = _run("--bogus", check=False)
    assert r.returncode == 2 and "unknown option" in r.stderr
    r = _run("--steps", "3", "--frames", "2", check=False)
    assert r.returncode == 2
    r = _run("--scene", "nope", "--dry-run", check=False)
    assert r.returncode == 1 and "unknown scene" in r.stderr


@pytest.mark.parametrize("frames,frame_time,max_steps", [(40, 0.04, 3), (30, 1 / 60, 3), (25, 0.1, 3),
                                                         (20, 0.05, 1)])
def test_runner_step_plan_is_the_application_loop(frames, frame_time, max_steps):
    r = _run("--scene", "app_dam_break", "--frames", str(frames), "--frame-time", repr(frame_time),
             "--max-steps-per-frame", str(max_steps), "--dry-run")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    want = _application_plan(frames, frame_time, 0.016, max_steps)
    assert out["steps_per_frame"] == want and out["steps"] == sum(want)
    if frame_time == 0.1:
        assert max(want) == 3          # the cap binds: the accumulator falls behind


def test_runner_dry_run_initial_frame(tmp_path):
    path = str(tmp_path / "init.akt")
    r = _run("--scene", "dam_break_27k", "--steps", "7", "--dry-run", "--out", path, "--fields", "all")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["steps"] == 7 and out["steps_per_frame"] == [1] * 7 and out["frames_written"] == 1
    f = T.read_trajectory(path)
    sc = scenes.dam_break_27k()
    assert len(f) == 1 and f.info.n == sc.n
    fr = f[0]
    np.testing.assert_array_equal(_bits(fr["pos"]), _bits(sc.pos))
    np.testing.assert_array_equal(fr["vel"], 0.0)
    np.testing.assert_array_equal(fr["mass"], 1.0)
    np.testing.assert_array_equal(fr["ids"], np.arange(sc.n, dtype=np.uint32))
    assert fr["step"] == 0 and f.info.dt == float(np.float32(0.0083))


def test_runner_dry_run_resume_keeps_ids(tmp_path):
    src = str(tmp_path / "src.akt")
    frames = _random_frames(64, 3, seed=7)
    _write(src, frames)
    dst = str(tmp_path / "dst.akt")
    r = _run("--init", src + "@1", "--steps", "2", "--dry-run", "--out", dst, "--fields", "all")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["particles"] == 64 and out["steps"] == 2
    f = T.read_trajectory(dst)
    fr = f[0]
    for name in ("pos", "vel", "mass", "density", "lam", "ids"):
        np.testing.assert_array_equal(fr[name], frames[1][name])
    assert fr["step"] == frames[1]["step"] and fr["time"] == frames[1]["time"]
    # a file without a whole frame is refused before anything is allocated
    empty = str(tmp_path / "empty.akt")
    open(empty, "wb").write(open(src, "rb").read()[:128])
    r = _run("--init", empty, "--steps", "1", "--dry-run", check=False)
    assert r.returncode == 2 and "no whole frame" in r.stderr
    # ids that are not a permutation are refused
    bad = [dict(fr, ids=np.zeros(64, np.uint32)) for fr in frames[:1]]
    _write(src, bad)
    r = _run("--init", src, "--steps", "1", "--dry-run", check=False)
    assert r.returncode == 2 and "permutation" in r.stderr


# ---- the renderer's camera (host) -----------------------------------------------------------------

def test_camera_matrix_is_glm_lookat_times_perspective(lib):
    # glm::lookAt (right-handed) and glm::perspective (depth in [-1, 1]) written out in float64,
    # for the reference's camera (CameraConfig.h:8-16) and two others
    from akuaengine_amd import solver as S
    for pos, yaw, pitch, fov, W, H in [((3, 2, 6), -90.0, 0.0, 45.0, 1280, 720),
                                       ((5, 4, 9), -110.0, -25.0, 60.0, 320, 240),
                                       ((-1, 0.5, 2), 30.0, 10.0, 35.0, 100, 300)]:
        c = S.default_camera()
        c.position[:] = pos
        c.yaw, c.pitch, c.fovY = yaw, pitch, fov
        m = S.camera_matrix(c, W, H).astype(np.float64)
        y, p = np.radians(yaw), np.radians(pitch)
        f = np.array([np.cos(y) * np.cos(p), np.sin(p), np.sin(y) * np.cos(p)])
        f /= np.linalg.norm(f)
        r = np.cross(f, [0, 1, 0]); r /= np.linalg.norm(r)
        u = np.cross(r, f); u /= np.linalg.norm(u)
        s = np.cross(f, u); s /= np.linalg.norm(s)
        uu = np.cross(s, f)
        e = np.array(pos, float)
        V = np.array([[*s, -s @ e], [*uu, -uu @ e], [*(-f), f @ e], [0, 0, 0, 1]])
        t = np.tan(np.radians(fov) / 2)
        n_, f_ = 0.1, 160.0
        P = np.array([[1 / (W / H * t), 0, 0, 0], [0, 1 / t, 0, 0],
                      [0, 0, -(f_ + n_) / (f_ - n_), -2 * f_ * n_ / (f_ - n_)], [0, 0, -1, 0]])
        np.testing.assert_allclose(m, P @ V, rtol=1e-6, atol=1e-6)
        # the gaze point lands in the middle of the image
        g = m @ np.array([*(e + 2 * f), 1.0])
        assert abs(g[0] / g[3]) < 1e-6 and abs(g[1] / g[3]) < 1e-6
