"""Boundary samples on the walls (opt-in), HIP path against the FP32 CPU oracle, bit for bit.

The samples, their volume weights psi, and every step's positions, velocities, densities and
lambdas must equal the oracle's exactly: the kernels visit the samples of each particle in the
oracle's order (ascending sample index) with the same FP32 operations.  Covered: 1-3 layers, the
reference application's scene and parameters, a dam break with negative coordinates, a box that
changes between steps, switching the samples off again, every launch setting, graph and
kernel-by-kernel launches, and the combinations the library refuses.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _assert_state_equal(s, o, what=("pos", "vel", "density", "lam", "ids")):
    g, r = s.get_state(), o.state()
    for k in what:
        a = g[k]
        b = r[k].astype(a.dtype)
        if not np.array_equal(a, b):
            d = np.abs(a.astype(np.float64) - r[k])
            raise AssertionError(f"{k} differs: {np.count_nonzero(d)} entries, max |diff| {d.max():.3e}")


def _pair(PBFSolver, O, sc, cfg, corr, layers, spacing=0.0):
    s = PBFSolver(sc.n, cfg, corr)
    s.set_state(sc.pos, sc.vel, sc.mass)
    s.set_boundary(layers, spacing)
    o = O.Oracle(sc.n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel, sc.mass)
    o.set_boundary(layers, spacing)
    return s, o


def _scene(name):
    from akuaengine_amd import scenes
    from akuaengine_amd.config import LambdaCorrParams, PBFConfig, dam_break_params
    if name == "app":
        return scenes.app_dam_break(), PBFConfig(), LambdaCorrParams(), 0.016
    if name == "negative":
        sc = scenes.dam_break_scaled(12, seed=4)
        sc.pos = sc.pos - np.float32(0.35)
        sc.box_min, sc.box_max = (-0.35, -0.35, -0.35), (0.85, 0.4, 0.25)
        return (sc, *dam_break_params(4), 0.0083)
    return (scenes.dam_break_27k(), *dam_break_params(4), 0.0083)


def test_samples_match_oracle(gpu, oracle_mod):
    sc, cfg, corr, dt = _scene("27k")
    s, o = _pair(gpu, oracle_mod, sc, cfg, corr, 2)
    s.step(dt, sc.box_min, sc.box_max)
    pos, psi = s.debug_boundary()
    opos, opsi = o.boundary_samples(sc.box_min, sc.box_max)
    assert len(psi) == len(opsi) > 0
    np.testing.assert_array_equal(pos, opos.astype(np.float32))
    np.testing.assert_array_equal(psi, opsi.astype(np.float32))


@pytest.mark.parametrize("name,layers", [("27k", 1), ("27k", 2), ("app", 2), ("app", 3), ("negative", 2)])
def test_steps_bit_exact(gpu, oracle_mod, name, layers):
    sc, cfg, corr, dt = _scene(name)
    s, o = _pair(gpu, oracle_mod, sc, cfg, corr, layers)
    for chunk in range(3):
        _say(f"[boundary] {name} layers {layers} steps {chunk * 10}..{chunk * 10 + 10}")
        s.step_n(dt, sc.box_min, sc.box_max, 10)
        o.step(dt, sc.box_min, sc.box_max, 10)
        _assert_state_equal(s, o)


def test_box_change_rebuilds_samples(gpu, oracle_mod):
    sc, cfg, corr, dt = _scene("27k")
    s, o = _pair(gpu, oracle_mod, sc, cfg, corr, 2)
    boxes = [(sc.box_min, sc.box_max), ((0.0, 0.0, 0.0), (1.6, 1.5, 1.0)), ((-0.05, 0.0, 0.0), (1.6, 1.4, 1.05))]
    for lo, hi in boxes:
        _say(f"[boundary] box {lo} {hi}")
        s.step_n(dt, lo, hi, 5)
        o.step(dt, lo, hi, 5)
        _assert_state_equal(s, o)
        pos, psi = s.debug_boundary()
        opos, opsi = o.boundary_samples(lo, hi)
        np.testing.assert_array_equal(psi, opsi.astype(np.float32))


def test_switching_off_returns_to_the_reference_walls(gpu, oracle_mod):
    sc, cfg, corr, dt = _scene("27k")
    s, o = _pair(gpu, oracle_mod, sc, cfg, corr, 2)
    s.step_n(dt, sc.box_min, sc.box_max, 5)
    o.step(dt, sc.box_min, sc.box_max, 5)
    s.set_boundary(0)
    o.set_boundary(0)
    s.step_n(dt, sc.box_min, sc.box_max, 5)
    o.step(dt, sc.box_min, sc.box_max, 5)
    _assert_state_equal(s, o)


def test_launch_settings_and_graphs_do_not_change_results(gpu):
    sc, cfg, corr, dt = _scene("app")
    ref = None
    for tuning, graph in [(dict(block=256, layoutLevel=1), True), (dict(block=128, layoutLevel=0), True),
                          (dict(block=512, layoutLevel=2, layoutInterval=3), True), (dict(block=256), False)]:
        s = gpu(sc.n, cfg, corr)
        s.set_state(sc.pos, sc.vel, sc.mass)
        s.set_boundary(2)
        s.set_tuning(**tuning)
        s.set_graph(graph)
        s.step_n(dt, sc.box_min, sc.box_max, 12)
        st = s.get_state()
        if ref is None:
            ref = st
        else:
            for k in ("pos", "vel", "density", "lam", "ids"):
                np.testing.assert_array_equal(st[k], ref[k], err_msg=f"{tuning} graph={graph} {k}")


def test_refused_combinations(gpu):
    from akuaengine_amd import _lib
    from akuaengine_amd.solver import BatchSolver
    sc, cfg, corr, dt = _scene("27k")
    s = gpu(sc.n, cfg, corr)
    s.set_state(sc.pos, sc.vel, sc.mass)
    with pytest.raises(_lib.AkuaError):
        s.set_boundary(5)
    with pytest.raises(_lib.AkuaError):
        s.set_boundary(2, -1.0)
    s.set_boundary(2)
    with pytest.raises(_lib.AkuaError):
        s.debug_boundary()           # no box passed yet: no admissible samples
    s.set_boundary(2, 1e-4)          # far too fine for the box: refused at the next step
    with pytest.raises(_lib.AkuaError, match="boundary"):
        s.step(dt, sc.box_min, sc.box_max)
    s.set_boundary(2)
    import os
    from akuaengine_amd.learned import load_network
    net = load_network(os.path.join(os.path.dirname(_lib.LIB_PATH), "networks", "dam_break_1m_reference_J1.npy"))
    s.set_network(net)
    with pytest.raises(_lib.AkuaError) as e:
        s.set_learned(True, 1.0, 0.01, "max", 1)
    assert e.value.status == 6                                   # AKUA_ERR_UNSUPPORTED
    b = BatchSolver([1000, 2000], cfg, corr)
    with pytest.raises(_lib.AkuaError) as e:
        b.set_boundary(2)
    assert e.value.status == 6


@pytest.mark.parametrize("iters", [0, 1, 7])
def test_iteration_counts_with_samples(gpu, oracle_mod, iters):
    from akuaengine_amd.config import with_iterations
    sc, cfg, corr, dt = _scene("27k")
    cfg = with_iterations(cfg, iters)
    # with no iteration the step reuses the stored density (vorticity and XSPH divide by it)
    rho = np.full(sc.n, cfg.restDensity, np.float32)
    s = gpu(sc.n, cfg, corr)
    s.set_state(sc.pos, sc.vel, sc.mass, rho)
    s.set_boundary(2)
    o = oracle_mod.Oracle(sc.n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel, sc.mass, rho)
    o.set_boundary(2)
    for _ in range(2):
        s.step_n(dt, sc.box_min, sc.box_max, 5)
        o.step(dt, sc.box_min, sc.box_max, 5)
        _assert_state_equal(s, o)


def test_non_unit_masses_with_samples(gpu, oracle_mod):
    # masses other than 1 select the kernels that multiply by each neighbour's mass; the
    # boundary terms must join those sums in the same order
    sc, cfg, corr, dt = _scene("app")
    rng = np.random.default_rng(8)
    sc.mass = rng.uniform(0.5, 2.0, sc.n).astype(np.float32)
    s, o = _pair(gpu, oracle_mod, sc, cfg, corr, 2)
    for _ in range(2):
        s.step_n(dt, sc.box_min, sc.box_max, 6)
        o.step(dt, sc.box_min, sc.box_max, 6)
        _assert_state_equal(s, o)


def test_record_stepping_with_samples(gpu, oracle_mod):
    # the caller-owned 108-byte record path (the masses come from device memory, so the general
    # kernels run) with samples on, against the oracle
    import torch
    from akuaengine_amd import _lib
    sc, cfg, corr, dt = _scene("27k")
    n = sc.n
    recs = (_lib.AkuaParticle * n)()
    for i in range(n):
        recs[i].position[:] = sc.pos[i].tolist()
        recs[i].mass = 1.0
    dev = torch.from_numpy(np.frombuffer(bytes(recs), dtype=np.uint8).copy()).cuda()
    s = gpu(n, cfg, corr)
    s.set_boundary(2)
    o = oracle_mod.Oracle(n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel, sc.mass)
    o.set_boundary(2)
    torch.cuda.synchronize()
    for _ in range(3):
        s.step_device(dev.data_ptr(), dt, sc.box_min, sc.box_max)
        s.synchronize()
        o.step(dt, sc.box_min, sc.box_max)
    back = (_lib.AkuaParticle * n).from_buffer_copy(dev.cpu().numpy().tobytes())
    got = np.array([list(back[i].position) for i in range(n)], np.float32)
    np.testing.assert_array_equal(got, o.state()["pos"].astype(np.float32))
