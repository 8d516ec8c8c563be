"""The numpy restatement of the warm-start network (oracle/mlp_oracle.py) and the C++ feature
restatement are pinned here: gradients by central finite differences, forward passes and Adam by
closed forms, features by closed forms on hand-placed particles."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def _data(rng, n=41):
    F = rng.normal(size=(n, 16)) * rng.uniform(0.1, 10, 16) + rng.normal(size=16)
    T = rng.normal(size=(n, 4)) * np.array([0.3, 1.0, 0.1, 2.0])
    return F, T


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_exact_gradient_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    F, T = _data(rng)
    p = M.random_params(rng, F, T)
    w = (1.0, 0.5, 2.0, 0.25)
    _, g = M.loss_and_grad(p, F, T, w)
    idx = np.concatenate([rng.choice(M.TRAINABLE, 60, replace=False),
                          np.arange(M.OFF_B3, M.TRAINABLE), np.arange(M.OFF_B1, M.OFF_B1 + 4)])
    for k in idx:
        e = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += e
        pm[k] -= e
        fd = (M.loss_only(pp, F, T, w) - M.loss_only(pm, F, T, w)) / (2 * e)
        assert abs(fd - g[k]) <= 1e-5 * (abs(fd) + 1e-3), (k, fd, g[k])


def test_clipped_targets_keep_the_gradient_exact():
    rng = np.random.default_rng(9)
    F, T = _data(rng)
    p = M.random_params(rng, F, T)
    T[::5] *= 50.0
    assert (np.abs(T / p[M.OFF_YMUL:]) > M.TARGET_CLIP).any()
    _, g = M.loss_and_grad(p, F, T)
    for k in rng.choice(M.TRAINABLE, 40, replace=False):
        e = 1e-6 * max(1.0, abs(p[k]))
        pp, pm = p.copy(), p.copy()
        pp[k] += e
        pm[k] -= e
        fd = (M.loss_only(pp, F, T) - M.loss_only(pm, F, T)) / (2 * e)
        assert abs(fd - g[k]) <= 1e-5 * (abs(fd) + 1e-3)


def test_a_plant_in_the_backward_pass_is_caught():
    """The finite-difference check fails for a gradient with a dropped ReLU mask."""
    rng = np.random.default_rng(5)
    F, T = _data(rng)
    p = M.random_params(rng, F, T)
    _, c = M.forward(p, F, cache=True)
    q = c["q"]
    diff = c["y"] - T / q["ymul"]
    G3 = 2.0 * diff / len(F)
    G2 = G3 @ q["W3"].T  # mask dropped
    bad = (c["h1"].T @ G2).ravel()
    _, g = M.loss_and_grad(p, F, T)
    assert np.abs(bad - g[M.OFF_W2:M.OFF_B2]).max() > 1e-3 * np.abs(g[M.OFF_W2:M.OFF_B2]).max()


def test_forward_closed_form_single_path():
    """A network with one active path computes y0 = ymul0 * (w3 * relu(w2 * relu(w1 * xn + b1) + b2) + b3)."""
    p = np.zeros(M.NPARAMS)
    p[M.OFF_INV:M.OFF_YMUL] = 1.0
    p[M.OFF_MEAN + 2] = 0.5
    p[M.OFF_INV + 2] = 4.0
    p[M.OFF_W1 + 2 * 64 + 7] = 1.5   # feature 2 -> unit 7
    p[M.OFF_B1 + 7] = 0.25
    p[M.OFF_W2 + 7 * 64 + 3] = -2.0  # unit 7 -> unit 3
    p[M.OFF_B2 + 3] = 10.0
    p[M.OFF_W3 + 3 * 4 + 1] = 0.125  # unit 3 -> output 1
    p[M.OFF_B3 + 1] = -1.0
    p[M.OFF_YMUL:] = [1.0, 3.0, 1.0, 1.0]
    F = np.zeros((3, 16))
    F[:, 2] = [0.5, 1.0, -1.0]
    y = M.forward(p, F)
    for r, f in enumerate(F[:, 2]):
        h1 = max(0.0, 1.5 * (f - 0.5) * 4.0 + 0.25)
        h2 = max(0.0, -2.0 * h1 + 10.0)
        assert y[r, 1] == pytest.approx(3.0 * (0.125 * h2 - 1.0))
        assert y[r, 0] == 0.0 and y[r, 2] == 0.0 and y[r, 3] == 0.0


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-39], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                       # tie -> even
    assert r[2] == 1.0 + 2 ** -6             # tie -> even (up)
    assert r[3] == 1.0                       # below half an ulp
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-39) < 2 ** -133   # subnormal kept to bf16 precision


def test_emulated_rounding_stays_close_to_exact():
    rng = np.random.default_rng(3)
    F, T = _data(rng, 200)
    p = M.random_params(rng, F, T)
    y0 = M.forward(p, F, exact=True)
    y1 = M.forward(p, F, exact=False)
    assert np.abs(y0 - y1).max() < 0.05 * np.abs(y0).max()
    _, g0 = M.loss_and_grad(p, F, T, exact=True)
    _, g1 = M.loss_and_grad(p, F, T, exact=False)
    cos = g0 @ g1 / np.linalg.norm(g0) / np.linalg.norm(g1)
    assert cos > 0.999


def test_adam_first_step_closed_form():
    """At t = 1 the bias-corrected step is lr * g / (|g| + eps): about lr * sign(g)."""
    rng = np.random.default_rng(4)
    p = rng.normal(size=M.NPARAMS)
    g = rng.normal(size=M.TRAINABLE)
    q, m, v = M.adam(p, np.zeros(M.TRAINABLE), np.zeros(M.TRAINABLE), g, 1, 1e-3)
    np.testing.assert_allclose(q[:M.TRAINABLE], p[:M.TRAINABLE] - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    np.testing.assert_array_equal(q[M.TRAINABLE:], p[M.TRAINABLE:])
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_oracle_features_closed_forms(oracle_mod):
    """Two particles a distance r apart along x, far from the walls: every feature in closed form
    (FP64 oracle), and features 0-1 equal the density and lambda stages (FP32 oracle)."""
    from akuaengine_amd.config import LambdaCorrParams, PBFConfig
    cfg = PBFConfig(restDensity=1000.0, gravity=(0.0, -9.8, 0.0))
    corr = LambdaCorrParams(k=0.1, n=4.0, delta_q=0.02)
    f = lambda v: float(np.float32(v))  # the oracle takes every parameter as a float
    h, r, dt, pi = f(0.1), 0.06, f(0.01), f(3.14)
    pos = np.array([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]], np.float32)
    vel = np.array([[0.5, 0.0, 0.0], [0.0, 0.0, 0.0]], np.float32)
    bmin, bmax = (0.0, 0.0, 0.0), (1.05, 3.0, 3.0)
    o = oracle_mod.Oracle(2, cfg, corr, prec=64)
    o.set_state(pos, vel, lam=np.array([0.25, -0.5], np.float32))
    for st in ("predict", "hashes", "sort", "neighbours", "features"):
        o.stage(st, dt, bmin, bmax)
    F = o.features()
    s = o.state()
    order = s["ids"]
    xs = pos.astype(np.float64) + (vel + np.array([0.0, f(-9.8) * dt, 0.0])) * dt
    xs = xs[order]
    d = np.linalg.norm(xs[0] - xs[1])
    W = lambda q2: 315.0 / (64 * pi * h ** 9) * (h * h - q2) ** 3
    dW = 45.0 / (pi * h ** 6) * (h - d) ** 2            # |grad W|
    rho = W(0.0) + W(d * d)
    C = rho / 1000.0 - 1.0
    gmag = dW / 1000.0
    lam = -C / (gmag ** 2 + gmag ** 2 + 600.0)
    scorr = -f(0.1) * (W(d * d) / W(f(0.02) ** 2)) ** 4
    for k in range(2):
        u = (xs[k] - xs[1 - k]) / d                      # grad W_ij points from j to i with -
        gvec = -dW * u
        assert F[k, 0] == pytest.approx(C, rel=1e-9)
        assert F[k, 1] == pytest.approx(lam, rel=1e-9)
        assert F[k, 2] == 1.0
        assert F[k, 3] == [0.25, -0.5][order[k]]
        np.testing.assert_allclose(F[k, 4:7], (xs[k] - pos[order[k]]) / h, rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(F[k, 7:10], 2 * lam * gvec / 1000.0 / h, rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(F[k, 10:13], scorr * gvec / 1000.0 / h, rtol=1e-9, atol=1e-15)
        for a in range(3):
            lo = min(max((xs[k, a] - f(bmin[a])) / h, 0.0), 2.0)
            hi = min(max((f(bmax[a]) - xs[k, a]) / h, 0.0), 2.0)
            assert F[k, 13 + a] == pytest.approx(lo - hi, abs=1e-12)
    # FP32: features 0 and 1 are exactly the first iteration's density constraint and lambda
    o32 = oracle_mod.Oracle(2, cfg, corr, prec=32)
    o32.set_state(pos, vel)
    for st in ("predict", "hashes", "sort", "neighbours", "features", "densities", "lambdas"):
        o32.stage(st, dt, bmin, bmax)
    F32, s32 = o32.features(), o32.state()
    np.testing.assert_array_equal(F32[:, 1], s32["lam"])
    rho32 = s32["density"].astype(np.float32)
    np.testing.assert_array_equal(F32[:, 0].astype(np.float32),
                                  rho32 * np.float32(1.0 / 1000.0) - np.float32(1.0))
