"""The reference's 108-byte particle records (Particle.h:8-31) through the C ABI, field by field
against the FP32 CPU oracle, bit for bit.

After a step the reference's record holds: position = new_position = x*, velocity after the wall
damping, vorticity confinement and XSPH, new_velocity = (x* - x) / dt from the update
(IntegrationCUDA.cu:45), position_delta of the last Jacobi iteration (ConstraintSolverCUDA.cu:129),
vorticity, density, lambda and hash of the last step.  A step without iterations keeps the
record's position_delta, which moves with its particle.  step_device is ordered on the caller's
stream, including the null stream, with no host synchronisation in between.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
DT = 0.0083


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def gpu():
    import torch
    if torch.cuda.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    from akuaengine_amd.solver import PBFSolver
    return PBFSolver


def _records(sc, rng=None):
    from akuaengine_amd import _lib
    recs = (_lib.AkuaParticle * sc.n)()
    for i in range(sc.n):
        recs[i].position[:] = sc.pos[i].tolist()
        recs[i].velocity[:] = sc.vel[i].tolist()
        recs[i].mass = 1.0
        recs[i].color[:] = [0.25, 0.5, 1.0, float(i)]
        recs[i].size = 40.0
        if rng is not None:
            recs[i].new_velocity[:] = rng.normal(size=3).tolist()
            recs[i].position_delta[:] = rng.normal(size=3).tolist()
    return recs


def _fields(recs):
    a = np.frombuffer(bytes(recs), dtype=np.float32).reshape(len(recs), 27)
    h = np.frombuffer(bytes(recs), dtype=np.uint32).reshape(len(recs), 27)
    return dict(pos=a[:, 0:3], vel=a[:, 3:6], new_pos=a[:, 6:9], new_vel=a[:, 9:12],
                dp=a[:, 12:15], vort=a[:, 15:18], mass=a[:, 18], density=a[:, 19], lam=a[:, 20],
                hash=h[:, 21], color=a[:, 22:26], size=a[:, 26])


def _assert_records_equal_oracle(f, o):
    st, sa, t = o.state(), o.stage_arrays(), o.taps()
    want = dict(pos=st["pos"], vel=st["vel"], new_pos=st["pos"], new_vel=sa["vs"], dp=sa["dp"],
                vort=sa["vort"], mass=st["mass"], density=st["density"], lam=st["lam"])
    for k, w in want.items():
        w32 = w.astype(np.float32)
        if not np.array_equal(f[k], w32):
            d = np.abs(f[k].astype(np.float64) - w)
            raise AssertionError(f"record field {k} differs: {np.count_nonzero(d)} entries, "
                                 f"max |diff| {d.max():.3e}")
    np.testing.assert_array_equal(f["hash"], t["sorted_hash"])
    np.testing.assert_array_equal(f["color"][:, 3], st["ids"].astype(np.float32))


@pytest.mark.parametrize("iters", [4, 1])
def test_record_fields_after_steps(gpu, oracle_mod, dam_params, iters):
    from akuaengine_amd import scenes
    from akuaengine_amd.config import with_iterations
    cfg, corr = dam_params
    cfg = with_iterations(cfg, iters)
    sc = scenes.dam_break_scaled(12, seed=3)
    _say(f"records: {sc.n} particles, {iters} iterations")
    s = gpu(sc.n, cfg, corr)
    s.set_particles(_records(sc))
    o = oracle_mod.Oracle(sc.n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel)
    for k in range(3):
        s.step_n(DT, sc.box_min, sc.box_max, 4)
        o.step(DT, sc.box_min, sc.box_max, 4)
        f = _fields(s.get_particles())
        _assert_records_equal_oracle(f, o)
    assert np.abs(f["dp"]).max() > 0 and np.abs(f["new_vel"] - f["vel"]).max() > 0


def test_step_without_iterations_keeps_position_delta(gpu, oracle_mod, dam_params):
    """With solverIterations 0 the uploaded position_delta moves with its particle and
    new_velocity is the step's (x* - x) / dt."""
    from akuaengine_amd import scenes
    from akuaengine_amd.config import with_iterations
    cfg, corr = dam_params
    cfg = with_iterations(cfg, 0)
    sc = scenes.dam_break_scaled(10, seed=5)
    rng = np.random.default_rng(2)
    recs = _records(sc, rng)
    dp0 = _fields(recs)["dp"].copy()
    rho = np.full(sc.n, cfg.restDensity, np.float32)
    for i in range(sc.n):
        recs[i].density = float(rho[i])
    s = gpu(sc.n, cfg, corr)
    s.set_particles(recs)
    f0 = _fields(s.get_particles())
    np.testing.assert_array_equal(f0["dp"], dp0)       # returned unchanged before a step
    o = oracle_mod.Oracle(sc.n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel, None, rho)
    s.step_n(DT, sc.box_min, sc.box_max, 3)
    o.step(DT, sc.box_min, sc.box_max, 3)
    f = _fields(s.get_particles())
    st, sa = o.state(), o.stage_arrays()
    np.testing.assert_array_equal(f["dp"], dp0[st["ids"]])
    np.testing.assert_array_equal(f["new_vel"], sa["vs"].astype(np.float32))
    np.testing.assert_array_equal(f["pos"], st["pos"].astype(np.float32))


@pytest.mark.parametrize("which", ["null", "user"])
def test_step_device_is_ordered_on_the_callers_stream(gpu, oracle_mod, dam_params, which):
    """Records written asynchronously on the caller's stream (behind queued work), stepped with
    step_device on that stream and read back on it, with no host synchronisation in between,
    equal the oracle's state, record fields included."""
    import torch
    from akuaengine_amd import _lib, scenes
    cfg, corr = dam_params
    sc = scenes.dam_break_scaled(12, seed=9)
    n = sc.n
    _say(f"step_device on the {which} stream: {n} particles")
    host = torch.from_numpy(np.frombuffer(bytes(_records(sc)), dtype=np.uint8).copy()).pin_memory()
    dev = torch.zeros_like(host, device="cuda")
    s = gpu(n, cfg, corr)
    o = oracle_mod.Oracle(n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel)
    torch.cuda.synchronize()
    stream = torch.cuda.Stream() if which == "user" else torch.cuda.default_stream()
    handle = stream.cuda_stream if which == "user" else 0
    big = torch.ones(1 << 24, device="cuda")
    with torch.cuda.stream(stream):
        for _ in range(20):            # queued work ahead of the copy
            big.mul_(1.0000001).add_(1e-7)
        dev.copy_(host, non_blocking=True)
        for step in range(3):
            s.step_device(dev.data_ptr(), DT, sc.box_min, sc.box_max, stream=handle)
            o.step(DT, sc.box_min, sc.box_max)
        out = dev.to("cpu", non_blocking=True)
        done = torch.cuda.Event()
        done.record(stream)
    done.synchronize()
    recs = (_lib.AkuaParticle * n).from_buffer_copy(out.numpy().tobytes())
    _assert_records_equal_oracle(_fields(recs), o)


@pytest.mark.parametrize("field,bad", [("pos", np.nan), ("vel", np.inf), ("mass", -np.inf),
                                       ("record", np.nan)])
def test_non_finite_uploads_are_rejected(gpu, dam_params, field, bad):
    """NaN or Inf in uploaded particle data is refused with AKUA_ERR_NON_FINITE, and the solver
    keeps its previous state."""
    from akuaengine_amd import _lib, scenes
    cfg, corr = dam_params
    sc = scenes.dam_break_scaled(6, seed=1)
    s = gpu(sc.n, cfg, corr)
    s.set_state(sc.pos, sc.vel)
    s.step(DT, sc.box_min, sc.box_max)
    before = s.get_state()
    pos, vel, mass = sc.pos.copy(), sc.vel.copy(), np.ones(sc.n, np.float32)
    with pytest.raises(_lib.AkuaError) as e:
        if field == "record":
            recs = _records(sc)
            recs[sc.n // 2].position[1] = float(bad)
            s.set_particles(recs)
        else:
            {"pos": pos, "vel": vel, "mass": mass}[field].reshape(-1)[7] = bad
            s.set_state(pos, vel, mass)
    assert e.value.status == 4                      # AKUA_ERR_NON_FINITE
    after = s.get_state()
    for k in ("pos", "vel", "ids"):
        np.testing.assert_array_equal(after[k], before[k])
    s.step(DT, sc.box_min, sc.box_max)                # still usable
    assert np.isfinite(s.get_state()["pos"]).all()
