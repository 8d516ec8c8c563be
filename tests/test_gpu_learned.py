"""The learned warm start on the device: the BF16-MFMA forward, gradient and Adam kernels against
the numpy restatement, the features against the C++ oracle, the learned step's continuation
against the oracle, and the fallback against the plain solver."""
import numpy as np
import pytest

from oracle import mlp_oracle as M

pytestmark = pytest.mark.gpu
DT = 0.0083


@pytest.fixture(scope="module")
def dev():
    import torch
    if torch.cuda.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    from akuaengine_amd import learned
    from akuaengine_amd.solver import PBFSolver
    return learned, PBFSolver


def _data(rng, n):
    F = rng.normal(size=(n, 16)) * rng.uniform(0.1, 10, 16) + rng.normal(size=16)
    T = rng.normal(size=(n, 4)) * np.array([0.3, 1.0, 0.1, 2.0])
    return F.astype(np.float32), T.astype(np.float32)


@pytest.mark.parametrize("n", [1, 16, 17, 1000, 65536 + 5])
def test_forward_matches_oracle(dev, n):
    learned, _ = dev
    rng = np.random.default_rng(n)
    F, T = _data(rng, n)
    p = M.random_params(rng, F, T).astype(np.float32)
    y = learned.mlp_forward(p, F)
    ye = M.forward(p, F, exact=False)
    yx = M.forward(p, F, exact=True)
    scale = np.abs(ye).max(0) + 1e-6
    # accumulation order only (and rarely one BF16 ulp of an activation)
    assert (np.abs(y - ye) / scale).max() < 2e-3
    assert (np.abs(y - yx) / scale).max() < 5e-2


@pytest.mark.parametrize("n", [16, 333, 20000])
def test_gradient_matches_oracle(dev, n):
    learned, _ = dev
    rng = np.random.default_rng(100 + n)
    F, T = _data(rng, n)
    p = M.random_params(rng, F, T).astype(np.float32)
    w = (1.0, 0.5, 2.0, 0.25)
    loss, g = learned.mlp_gradient(p, F, T, w)
    le, ge = M.loss_and_grad(p, F, T, w, exact=False)
    lx, gx = M.loss_and_grad(p, F, T, w, exact=True)
    assert loss == pytest.approx(le, rel=2e-3)
    assert np.linalg.norm(g - ge) < 2e-3 * np.linalg.norm(ge)
    for a, b in ((M.OFF_W1, M.OFF_B1), (M.OFF_W2, M.OFF_B2), (M.OFF_W3, M.OFF_B3), (M.OFF_B1, M.OFF_W2)):
        assert np.linalg.norm(g[a:b] - ge[a:b]) < 5e-3 * np.linalg.norm(ge[a:b]) + 1e-9, (a, b)
    assert g @ gx / np.linalg.norm(g) / np.linalg.norm(gx) > 0.999
    assert loss == pytest.approx(lx, rel=2e-2)


def test_gradient_with_clipped_targets_matches_oracle(dev):
    """Outlier targets beyond 8 normalised units enter the loss clipped."""
    learned, _ = dev
    rng = np.random.default_rng(12)
    F, T = _data(rng, 5000)
    p = M.random_params(rng, F, T).astype(np.float32)
    T[::20] *= 100.0
    assert (np.abs(T / p[M.OFF_YMUL:]) > M.TARGET_CLIP).any()
    loss, g = learned.mlp_gradient(p, F, T)
    le, ge = M.loss_and_grad(p, F, T, exact=False)
    assert loss == pytest.approx(le, rel=2e-3)
    assert np.linalg.norm(g - ge) < 2e-3 * np.linalg.norm(ge)


def test_adam_step_matches_oracle(dev):
    learned, _ = dev
    rng = np.random.default_rng(7)
    F, T = _data(rng, 4096)
    p = M.random_params(rng, F, T).astype(np.float32)
    m = (rng.normal(size=M.TRAINABLE) * 1e-3).astype(np.float32)
    v = (rng.uniform(size=M.TRAINABLE) * 1e-6).astype(np.float32)
    _, g = learned.mlp_gradient(p, F, T)
    p1, m1, v1, _ = learned.mlp_adam_step(p, m, v, 5, 1e-3, F, T)
    pe, me, ve = M.adam(p, m, v, g.astype(np.float64), 5, 1e-3)
    # FP32 arithmetic on the device: tolerances relative to each vector's scale
    np.testing.assert_allclose(m1, me, rtol=1e-5, atol=1e-6 * np.abs(me).max())
    # (1 - beta2) formed in FP32 is 1.3e-5 away from 0.001
    np.testing.assert_allclose(v1, ve, rtol=3e-5, atol=1e-6 * np.abs(ve).max())
    # the update, to FP32 rounding of the stored parameter (one ulp) plus 1e-4 of the step
    T_ = M.TRAINABLE
    step_gpu = p1[:T_].astype(np.float64) - p[:T_]
    step_ref = pe[:T_] - p[:T_]
    ulp = np.spacing(np.abs(p[:T_]).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(step_gpu - step_ref) <= 1e-4 * np.abs(step_ref) + ulp)
    np.testing.assert_array_equal(p1[T_:], p[T_:])


def test_training_kernels_reduce_the_loss(dev):
    learned, _ = dev
    rng = np.random.default_rng(8)
    F = rng.normal(size=(8192, 16)).astype(np.float32)
    T = np.stack([np.tanh(F[:, 0] + F[:, 1]), F[:, 2] * F[:, 3], np.abs(F[:, 4]), F[:, 5]], 1).astype(np.float32)
    p = M.random_params(rng, F, T).astype(np.float32)
    m = np.zeros(M.TRAINABLE, np.float32)
    v = np.zeros(M.TRAINABLE, np.float32)
    losses = []
    for t in range(1, 201):
        p, m, v, loss = learned.mlp_adam_step(p, m, v, t, 3e-3, F, T)
        losses.append(loss)
    assert losses[-1] < 0.2 * losses[0]


def _scene_pair(PBFSolver, oracle_mod, dam_params, n_side=14):
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    sc = scenes.dam_break_scaled(n_side)
    s = PBFSolver(sc.n, cfg, corr)
    s.set_state(sc.pos, sc.vel)
    o = oracle_mod.Oracle(sc.n, cfg, corr)
    o.set_state(sc.pos, sc.vel)
    return sc, s, o


def _random_net(rng):
    F = rng.normal(size=(64, 16))
    p = M.random_params(rng)
    p[M.OFF_INV:M.OFF_YMUL] = 0.5
    p[M.OFF_YMUL:] = [0.05, 0.05, 0.05, 1e-4]
    return p.astype(np.float32)


@pytest.mark.parametrize("fuse", [0, 1])
@pytest.mark.parametrize("J", [1, 2, 4])
def test_features_and_learned_step_match_oracle(dev, oracle_mod, dam_params, J, fuse):
    """After a plain step, one learned step: the features equal the oracle's (bit for bit), the
    warm start equals x*_pred + h * network(features), and the step continued by the oracle from
    that warm start (J iterations, then the rest of the step) equals the device's.  fuse 1: the
    list build, the first plain sweep and the feature sweep share their launches."""
    learned, PBFSolver = dev
    sc, s, o = _scene_pair(PBFSolver, oracle_mod, dam_params)
    s.set_tuning(fuseSweep=fuse)
    s.step(DT, sc.box_min, sc.box_max)
    o.step(DT, sc.box_min, sc.box_max)
    net = _random_net(np.random.default_rng(1))
    s.set_network(net)
    s.set_learned(True, ratio=1.0, tolerance=1e30, iterations=J)   # never falls back
    s.step(DT, sc.box_min, sc.box_max)
    st = s.learned_stats()
    assert st["steps"] == 1 and st["fallbacks"] == 0
    F = s.debug_features()
    xp, xw, lam_hat = s.debug_warm_start()
    for name in ("predict", "hashes", "sort", "neighbours", "features"):
        o.stage(name, DT, sc.box_min, sc.box_max)
    Fo = o.features().astype(np.float32)
    np.testing.assert_array_equal(F, Fo)
    np.testing.assert_array_equal(xp, o.stage_arrays()["xs"].astype(np.float32))
    y = learned.mlp_forward(net, F)
    h = np.float32(dam_params[0].smoothRadius)
    ymul = net[M.OFF_YMUL:M.OFF_YMUL + 3]
    yc = np.clip(y[:, :3] / ymul, -M.TARGET_CLIP, M.TARGET_CLIP).astype(np.float32) * ymul
    np.testing.assert_array_equal(xw, xp + h * yc)
    np.testing.assert_array_equal(lam_hat, y[:, 3])
    o.set_xs(xw)
    names = ("densities", "lambdas", "deltas", "correct") * J + (
        "update", "damping", "vorticity", "confinement", "xsph")
    for name in names:
        o.stage(name, DT, sc.box_min, sc.box_max)
    got, want = s.get_state(), o.state()
    np.testing.assert_array_equal(got["ids"], want["ids"])
    np.testing.assert_array_equal(got["pos"], want["pos"].astype(np.float32))
    np.testing.assert_array_equal(got["vel"], want["vel"].astype(np.float32))
    assert np.abs(xw - xp).max() > 0  # the warm start moved the particles


@pytest.mark.parametrize("J", [1, 3])
def test_fallback_reproduces_the_plain_solver(dev, dam_params, J):
    """With a gate that always fails, every learned step falls back: the trajectory is the plain
    solver's, bit for bit."""
    _, PBFSolver = dev
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    sc = scenes.dam_break_27k()
    a = PBFSolver(sc.n, cfg, corr)
    b = PBFSolver(sc.n, cfg, corr)
    for s in (a, b):
        s.set_state(sc.pos, sc.vel)
    b.set_network(_random_net(np.random.default_rng(2)))
    b.set_learned(True, ratio=0.0, tolerance=-1.0, iterations=J)
    a.step_n(DT, sc.box_min, sc.box_max, 10)
    b.step_n(DT, sc.box_min, sc.box_max, 10)
    st = b.learned_stats()
    assert st["steps"] == 10 and st["fallbacks"] == 10
    ga, gb = a.get_state(), b.get_state()
    for k in ("pos", "vel", "density", "lam", "ids"):
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)


def test_gate_decides_per_scene(dev, dam_params):
    """In a batch, the gate is per scene: scenes that pass take the learned result, the others
    the plain one, each equal to the same scene stepped alone in that mode."""
    _, PBFSolver = dev
    from akuaengine_amd import scenes
    from akuaengine_amd.solver import BatchSolver
    cfg, corr = dam_params
    counts = [3000, 5000]
    scs = [scenes.batched_scene(n, i, seed=11) for i, n in enumerate(counts)]
    net = _random_net(np.random.default_rng(3))
    b = BatchSolver(counts, cfg, corr)
    b.set_state(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]))
    for i, s in enumerate(scs):
        b.set_box(i, s.box_min, s.box_max)
    b.set_network(net)
    b.set_learned(True, ratio=1.0, tolerance=1e30)
    b.step_batch(DT, 2)
    got = b.get_state()
    off = 0
    for i, s in enumerate(scs):
        one = PBFSolver(s.n, cfg, corr)
        one.set_state(s.pos, s.vel)
        one.set_network(net)
        one.set_learned(True, ratio=1.0, tolerance=1e30)
        one.step_n(DT, s.box_min, s.box_max, 2)
        want = one.get_state()
        sl = slice(off, off + s.n)
        for k in ("pos", "vel"):
            np.testing.assert_array_equal(got[k][sl], want[k], err_msg=f"scene {i} {k}")
        off += s.n


def test_trained_network_gains_over_the_step_without_it(dev):
    """The committed J = 1 network (trained on the GPU by train_warm_start.py) against the same
    learned step with a zero network (no warm start: one Jacobi iteration), on the 1M evaluation
    scene with the reference application's parameters at a 2 ms substep, 150 steps, no fallback.
    The trained warm start must cut the time-averaged mean density error; the max is reported."""
    learned, PBFSolver = dev
    from akuaengine_amd import scenes
    from akuaengine_amd.config import LambdaCorrParams, PBFConfig
    import os
    path = os.path.join(learned.NETWORK_DIR, "dam_break_1m_reference_J1.npy")
    net = learned.load_network(path)
    zero = net.copy()
    zero[:M.TRAINABLE] = 0.0
    sc = scenes.dam_break_1m()
    res = {}
    for name, p in (("trained", net), ("zero", zero)):
        s = PBFSolver(sc.n, PBFConfig(), LambdaCorrParams())
        s.set_state(sc.pos, sc.vel)
        s.set_network(p)
        s.set_learned(True, ratio=1.0, tolerance=1e30, iterations=1)
        errs = []
        for _ in range(6):
            s.step_n(0.002, sc.box_min, sc.box_max, 25)
            errs.append(s.density_error())
        res[name] = np.array(errs).mean(0)
    print("time-averaged density error (max, mean):", res)
    assert res["trained"][1] < 0.7 * res["zero"][1]


@pytest.mark.parametrize("fuse", [0, 1])
def test_training_records_plain_steps_and_its_targets(dev, oracle_mod, dam_params, fuse):
    """akua_solver_train: the recorded steps advance the state exactly as plain steps do; the
    features and targets of the last recorded step equal the oracle's staged values (features
    at x*_pred; targets (x*_{I-J} - x*_pred) / h and the lambda computed at x*_{I-J}); two runs
    from the same state and seed give the same network, bit for bit; the learned-step iteration
    count is left alone."""
    learned, PBFSolver = dev
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    sc = scenes.dam_break_scaled(16, seed=6)
    net0 = _random_net(np.random.default_rng(4))
    steps, J = 3, 2
    nets, states = [], []
    for run in range(2):
        s = PBFSolver(sc.n, cfg, corr)
        s.set_tuning(fuseSweep=fuse)
        s.set_state(sc.pos, sc.vel)
        s.set_network(net0)
        loss, resid = s.train(DT, sc.box_min, sc.box_max, steps, samplesPerStep=4096,
                              minibatch=1024, epochs=2, seed=7, targetIterations=J)
        assert np.isfinite(loss).all() and np.isfinite(resid).all()
        nets.append(s.get_network())
        states.append(s.get_state())
        F, T = s.debug_features(targets=True)
    np.testing.assert_array_equal(nets[0], nets[1])
    assert np.abs(nets[0] - net0).max() > 0            # training moved the weights
    plain = PBFSolver(sc.n, cfg, corr)
    plain.set_state(sc.pos, sc.vel)
    plain.step_n(DT, sc.box_min, sc.box_max, steps)
    want = plain.get_state()
    for k in ("pos", "vel", "density", "lam", "ids"):
        np.testing.assert_array_equal(states[0][k], want[k], err_msg=k)
    # the oracle: steps - 1 plain steps, then the last step's stages
    o = oracle_mod.Oracle(sc.n, cfg, corr)
    o.set_state(sc.pos, sc.vel)
    o.step(DT, sc.box_min, sc.box_max, steps - 1)
    for name in ("predict", "hashes", "sort", "neighbours", "features"):
        o.stage(name, DT, sc.box_min, sc.box_max)
    np.testing.assert_array_equal(F, o.features().astype(np.float32))
    xpred = o.stage_arrays()["xs"].astype(np.float32)
    I = cfg.solverIterations
    for name in ("densities", "lambdas", "deltas", "correct") * (I - J):
        o.stage(name, DT, sc.box_min, sc.box_max)
    o.stage("densities", DT, sc.box_min, sc.box_max)
    o.stage("lambdas", DT, sc.box_min, sc.box_max)
    xlast = o.stage_arrays()["xs"].astype(np.float32)
    inv_h = np.float32(1.0) / np.float32(cfg.smoothRadius)
    np.testing.assert_array_equal(T[:, :3], (xlast - xpred) * inv_h)
    np.testing.assert_array_equal(T[:, 3], o.state()["lam"].astype(np.float32))
    # set_learned's iteration count survives training: after training with J = 2, a learned
    # step still runs one iteration, as a fresh solver in the same state does
    s.set_learned(True, ratio=1.0, tolerance=1e30, iterations=1)
    s.train(DT, sc.box_min, sc.box_max, 1, samplesPerStep=4096, minibatch=1024, seed=8,
            targetIterations=J)
    st = s.get_state()
    ref = PBFSolver(sc.n, cfg, corr)
    ref.set_state(st["pos"], st["vel"], st["mass"], st["density"], st["lam"])
    ref.set_network(s.get_network())
    ref.set_learned(True, ratio=1.0, tolerance=1e30, iterations=1)
    s.step(DT, sc.box_min, sc.box_max)
    ref.step(DT, sc.box_min, sc.box_max)
    a, b = s.get_state(), ref.get_state()
    np.testing.assert_array_equal(a["pos"], b["pos"])
    np.testing.assert_array_equal(st["ids"][b["ids"]], a["ids"])


def test_learned_step_at_8m_matches_oracle(dev, oracle_mod, dam_params):
    """The learned step on the 8M two-column scene with the committed J = 1 network: features
    equal the oracle's bit for bit, x*_w = x*_pred + h * network(features), and the step the
    oracle continues from x*_w equals the device's."""
    learned, PBFSolver = dev
    import os
    from akuaengine_amd import scenes
    print("[learned 8m] scene", flush=True)
    cfg, corr = dam_params
    sc = scenes.dam_break_8m()
    net = learned.load_network(os.path.join(learned.NETWORK_DIR, "dam_break_1m_reference_J1.npy"))
    s = PBFSolver(sc.n, cfg, corr)
    s.set_state(sc.pos, sc.vel)
    s.step(DT, sc.box_min, sc.box_max)
    s.set_network(net)
    s.set_learned(True, ratio=1.0, tolerance=1e30, iterations=1)   # never falls back
    s.step(DT, sc.box_min, sc.box_max)
    assert s.learned_stats()["fallbacks"] == 0
    F = s.debug_features()
    xp, xw, lam_hat = s.debug_warm_start()
    got = s.get_state()
    del s
    print("[learned 8m] oracle plain step", flush=True)
    o = oracle_mod.Oracle(sc.n, cfg, corr)
    o.set_state(sc.pos, sc.vel)
    o.step(DT, sc.box_min, sc.box_max)
    for name in ("predict", "hashes", "sort", "neighbours", "features"):
        print(f"[learned 8m] oracle {name}", flush=True)
        o.stage(name, DT, sc.box_min, sc.box_max)
    np.testing.assert_array_equal(F, o.features().astype(np.float32))
    np.testing.assert_array_equal(xp, o.stage_arrays()["xs"].astype(np.float32))
    y = learned.mlp_forward(net, F)
    h = np.float32(cfg.smoothRadius)
    ymul = net[M.OFF_YMUL:M.OFF_YMUL + 3]
    yc = np.clip(y[:, :3] / ymul, -M.TARGET_CLIP, M.TARGET_CLIP).astype(np.float32) * ymul
    np.testing.assert_array_equal(xw, xp + h * yc)
    np.testing.assert_array_equal(lam_hat, y[:, 3])
    o.set_xs(xw)
    for name in ("densities", "lambdas", "deltas", "correct", "update", "damping", "vorticity",
                 "confinement", "xsph"):
        print(f"[learned 8m] oracle {name}", flush=True)
        o.stage(name, DT, sc.box_min, sc.box_max)
    want = o.state()
    np.testing.assert_array_equal(got["ids"], want["ids"])
    np.testing.assert_array_equal(got["pos"], want["pos"].astype(np.float32))
    np.testing.assert_array_equal(got["vel"], want["vel"].astype(np.float32))
