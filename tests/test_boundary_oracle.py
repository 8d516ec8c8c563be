"""The oracle's boundary samples (opt-in wall samples after Akinci et al. 2012), pinned on the CPU.

  * Sums over the index window equal brute-force sums over every sample (same samples, ascending
    index, float32 terms from the oracle's own kernels), at points inside, near walls and
    corners, and outside the box.
  * psi equals its brute-force definition rho0 / sum_k W(x_b - x_k) bit for bit, and satisfies
    the invariant sum_k psi_k W(x_b - x_k) = rho0 on the flat part of a wall.
  * The sample layout follows its closed form: counts per wall, layer planes, unique positions,
    every edge and corner covered.
  * One particle near the floor: density, lambda and position delta equal their closed forms
    from the boundary sums; a particle far from every wall sees no boundary term.
  * A resting block on the floor: with the samples its bottom layer's density comes close to the
    interior's; without them it is far below.
"""
import numpy as np
import pytest

from akuaengine_amd.config import LambdaCorrParams, PBFConfig

H = np.float32(0.1)


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _oracle(O, pts, layers=2, spacing=0.0, cfg=None):
    cfg = cfg or PBFConfig()
    o = O.Oracle(len(pts), cfg, LambdaCorrParams(), 32)
    o.set_state(np.asarray(pts, np.float32), np.zeros((len(pts), 3), np.float32))
    o.set_boundary(layers, spacing)
    return o


def _brute_sums(O, x, pos, psi):
    """Sum over all samples in index order with |x - x_b|^2 < h^2, float32, oracle kernels."""
    x = np.asarray(x, np.float32)
    sep = (x[None, :] - pos.astype(np.float32)).astype(np.float32)
    d2 = ((sep[:, 0] * sep[:, 0]) + (sep[:, 1] * sep[:, 1])) + (sep[:, 2] * sep[:, 2])
    rb = np.float32(0.0)
    gb = np.zeros(3, np.float32)
    for k in np.nonzero(d2 < H * H)[0]:
        w = np.float32(O.poly6(float(d2[k]), float(H)))
        g = O.grad_spiky(sep[k].astype(np.float64), float(H)).astype(np.float32)
        p = np.float32(psi[k])
        rb = np.float32(rb + p * w)
        gb = (gb + p * g).astype(np.float32)
    return rb, gb


BOX = ((0.0, 0.0, 0.0), (0.6, 0.45, 0.3))


def test_window_sums_equal_brute_force(O):
    rng = np.random.default_rng(5)
    lo, hi = np.array(BOX[0]), np.array(BOX[1])
    pts = [lo + rng.random(3) * (hi - lo) for _ in range(40)]                  # anywhere inside
    pts += [np.array([lo[0] + 0.01 + 0.07 * rng.random(), rng.random() * hi[1], rng.random() * hi[2]])
            for _ in range(20)]                                                 # near x-lo
    pts += [lo + 0.03 * rng.random(3) for _ in range(10)]                       # corner lo
    pts += [hi - 0.03 * rng.random(3) for _ in range(10)]                       # corner hi
    pts += [lo - 0.05 * rng.random(3) for _ in range(10)]                       # outside
    pts += [np.array([0.3, 0.2, 0.15])]                                         # far from walls
    pts = np.array(pts, np.float32)
    for layers in (1, 2, 3):
        o = _oracle(O, pts, layers)
        pos, psi = o.boundary_samples(*BOX)
        rb, gb = o.boundary_sums(*BOX)
        st = o.state()
        for i in range(len(pts)):
            want_r, want_g = _brute_sums(O, st["pos"][i], pos, psi)
            assert np.float32(rb[i]) == want_r, (layers, i)
            np.testing.assert_array_equal(gb[i].astype(np.float32), want_g)
        assert rb[-1] == 0.0 and not gb[-1].any()


def test_psi_equals_brute_force_definition(O):
    o = _oracle(O, [[0.3, 0.2, 0.15]], 2)
    pos, psi = o.boundary_samples(*BOX)
    rest = np.float32(PBFConfig().restDensity)
    p32 = pos.astype(np.float32)
    for b in range(0, len(pos), 7):
        sep = (p32[b][None, :] - p32).astype(np.float32)
        d2 = ((sep[:, 0] * sep[:, 0]) + (sep[:, 1] * sep[:, 1])) + (sep[:, 2] * sep[:, 2])
        delta = np.float32(0.0)
        for k in np.nonzero(d2 < H * H)[0]:
            delta = np.float32(delta + np.float32(O.poly6(float(d2[k]), float(H))))
        assert np.float32(rest / delta) == np.float32(psi[b]), b


def test_psi_invariant_on_a_flat_wall(O):
    # in the middle of a large wall every neighbouring sample has the same psi, so
    # sum_k psi_k W_bk = psi_b delta_b = rho0
    box = ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    o = _oracle(O, [[0.5, 0.5, 0.5]], 2)
    pos, psi = o.boundary_samples(*box)
    rest = PBFConfig().restDensity
    mid = np.nonzero((pos[:, 1] < -0.0) & (np.abs(pos[:, 0] - 0.5) < 0.2) & (np.abs(pos[:, 2] - 0.5) < 0.2))[0]
    mid = mid if len(mid) else np.nonzero((np.abs(pos[:, 1]) < 1e-6) & (np.abs(pos[:, 0] - 0.5) < 0.2)
                                          & (np.abs(pos[:, 2] - 0.5) < 0.2))[0]
    assert len(mid) > 10
    for b in mid[:20]:
        d2 = np.sum((pos - pos[b]) ** 2, axis=1)
        near = np.nonzero(d2 < 0.01)[0]
        total = sum(psi[k] * O.poly6(float(d2[k]), 0.1, prec=64) for k in near)
        assert total == pytest.approx(rest, rel=2e-5)
    # on the flat part psi is uniform (to the rounding of float32 coordinates); near edges and
    # corners fewer neighbours raise it
    assert np.ptp(psi[mid]) < 1e-5 * psi[mid].max()
    assert psi.max() > psi[mid].max()


@pytest.mark.parametrize("layers", [1, 2, 3])
def test_sample_layout_closed_form(O, layers):
    lo, hi = np.float32(BOX[0]), np.float32(BOX[1])
    s = np.float32(0.05)
    n = [max(2, int(np.ceil(np.float32(np.float32(hi[a] - lo[a]) / s))) + 1) for a in range(3)]
    count = 0
    for w in range(6):
        a = w // 2
        tang = [e for e in range(3) if e != a]
        dims = [n[e] + 2 * (layers - 1) if e > a else n[e] - 2 for e in tang]
        count += layers * dims[0] * dims[1]
    o = _oracle(O, [[0.3, 0.2, 0.15]], layers)
    pos, psi = o.boundary_samples(*BOX)
    assert len(pos) == count
    assert len(np.unique(np.round(pos, 6), axis=0)) == count        # no sample twice
    inside = np.all((pos > lo + 1e-6) & (pos < hi - 1e-6), axis=1)
    assert not inside.any()                                         # all on or outside the walls
    planes = {np.float32(lo[0] - np.float32(l) * s) for l in range(layers)}
    xlo = pos[:, 0] <= lo[0] + 1e-7
    assert {np.float32(v) for v in pos[xlo, 0]} <= planes | {np.float32(lo[0])}
    # every point of every edge has a sample within the largest spacing
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = lo + rng.random(3) * (hi - lo)
        a, b = rng.choice(3, 2, replace=False)
        p[a] = lo[a] if rng.random() < 0.5 else hi[a]
        p[b] = lo[b] if rng.random() < 0.5 else hi[b]
        assert np.min(np.linalg.norm(pos - p, axis=1)) <= 0.05 + 1e-6
    assert np.all(psi > 0)


def test_single_particle_closed_form(O):
    cfg = PBFConfig()
    x = np.array([[0.3, 0.02, 0.15]], np.float32)         # near the floor, far from other walls
    o = _oracle(O, x, 2, cfg=cfg)
    rb, gb = o.boundary_sums(*BOX)
    for name in ("hashes", "sort", "neighbours", "densities", "lambdas", "deltas"):
        o.stage(name, 0.016, *BOX)
    st, sa = o.state(), o.stage_arrays()
    w0 = np.float32(O.poly6(0.0, float(H)))
    rho = np.float32(np.float32(1.0) * w0) + np.float32(rb[0])
    assert np.float32(st["density"][0]) == rho
    inv = np.float32(np.float32(1.0) / np.float32(cfg.restDensity))
    gi = (gb[0].astype(np.float32) * inv).astype(np.float32)
    C = np.float32(rho * inv - np.float32(1.0))
    dd = np.float32(np.float32(np.float32(0.0) + np.float32(np.dot(gi.astype(np.float64), gi))) + np.float32(cfg.relaxation))
    lam = np.float32(-C / dd)
    assert np.float32(st["lam"][0]) == pytest.approx(lam, rel=1e-6)
    dp = (np.float32(st["lam"][0]) * gb[0].astype(np.float32)).astype(np.float32) * inv
    np.testing.assert_allclose(sa["dp"][0], dp, rtol=1e-6, atol=1e-12)
    # the spiky gradient of x_i - x_b points towards the sample (its coefficient is negative), so
    # above the floor gB points down with no tangential part (the samples are symmetric about x);
    # dp = lambda gB / rho0 pushes a compressed particle (lambda < 0) up and pulls this lone,
    # under-dense one (C < 0, lambda > 0) down
    assert gb[0][1] < 0 and abs(gb[0][0]) < 1e-5 * abs(gb[0][1]) and abs(gb[0][2]) < 1e-5 * abs(gb[0][1])
    assert st["lam"][0] > 0 and sa["dp"][0][1] < 0
    # far from every wall: no boundary term at all
    o2 = _oracle(O, [[0.3, 0.2, 0.15]], 2, cfg=cfg)
    rb2, gb2 = o2.boundary_sums(*BOX)
    assert rb2[0] == 0.0 and not gb2[0].any()


def test_resting_block_bottom_density(O):
    # a lattice block on the floor at the particle spacing: the bottom layer lacks the
    # neighbours below it and the samples add density there; particles farther than h from
    # every wall keep exactly the fluid-only density.  A single sheet supplies about rho0 at its
    # own plane (psi's definition), so more layers spread the same support and add less at the
    # bottom layer's distance.
    s = 0.05
    ix, iy, iz = np.meshgrid(np.arange(8), np.arange(6), np.arange(8), indexing="ij")
    pts = np.stack([0.2 + ix * s, 0.025 + iy * s, 0.2 + iz * s], axis=-1).reshape(-1, 3).astype(np.float32)
    box = ((0.0, 0.0, 0.0), (0.8, 0.6, 0.8))
    rho = {}
    for layers in (0, 1, 2, 3):
        o = _oracle(O, pts, layers)
        for name in ("hashes", "sort", "neighbours", "densities"):
            o.stage(name, 0.016, *box)
        st = o.state()
        order = np.argsort(st["ids"])
        rho[layers] = st["density"][order]
    y = pts[:, 1]
    far = y > 0.1 + 1e-6                       # farther than h from the floor (and the other walls)
    for layers in (1, 2, 3):
        np.testing.assert_array_equal(rho[layers][far], rho[0][far])
    bottom = y < 0.03
    core = (np.abs(pts[:, 0] - 0.375) < 0.06) & (np.abs(pts[:, 2] - 0.375) < 0.06)
    interior = core & (np.abs(y - 0.125) < 0.03)
    assert (rho[0][bottom & core] < 0.85 * rho[0][interior].mean()).all()  # under-dense without samples
    for layers in (1, 2, 3):
        assert (rho[layers][bottom] > rho[0][bottom]).all()
    assert rho[1][bottom].mean() > rho[2][bottom].mean() >= rho[3][bottom].mean()


def test_thin_box_walls_without_interior_rows(O):
    # an axis shorter than the spacing has n = 2 points, so walls whose tangential range on that
    # axis is interior-only hold no samples; sums still equal brute force
    box = ((0.0, 0.0, 0.0), (0.6, 0.02, 0.3))
    rng = np.random.default_rng(9)
    pts = np.stack([rng.random(30) * 0.6, rng.random(30) * 0.02, rng.random(30) * 0.3], axis=1)
    pts = np.concatenate([pts, [[0.01, 0.01, 0.01], [0.59, 0.019, 0.29]]]).astype(np.float32)
    o = _oracle(O, pts, 2)
    pos, psi = o.boundary_samples(*box)
    # z-walls own no edge with y (interior-only on y, n_y = 2): they have no samples at all
    y = pos[:, 1].astype(np.float32)
    assert not np.any((np.abs(pos[:, 2]) < 1e-7) & (y > np.float32(0.0)) & (y < np.float32(0.02)))
    rb, gb = o.boundary_sums(*box)
    st = o.state()
    for i in range(len(pts)):
        want_r, want_g = _brute_sums(O, st["pos"][i], pos, psi)
        assert np.float32(rb[i]) == want_r, i
        np.testing.assert_array_equal(gb[i].astype(np.float32), want_g)


@pytest.mark.parametrize("box,layers,spacing", [
    (((0.0, 0.0, 0.0), (0.6, 0.45, 0.3)), 2, 0.0),
    (((0.0, 0.0, 0.0), (2.0, 1.5, 1.0)), 1, 0.0),
    (((-0.3, -0.3, -0.3), (1.0, 0.6, 0.2)), 3, 0.0),
    (((0.0, 0.0, 0.0), (0.6, 0.02, 0.3)), 2, 0.0),
    (((1.5, 0.0, 1.5), (4.5, 4.0, 4.5)), 2, 0.07),
])
def test_library_samples_equal_the_oracle(O, box, layers, spacing):
    # libakua.so's host construction (the one the solver uploads) against the oracle's, bit for bit
    from akuaengine_amd.solver import boundary_samples
    cfg = PBFConfig()
    pos, psi = boundary_samples(cfg, *box, layers, spacing)
    o = _oracle(O, [[0.0, 0.0, 0.0]], layers, spacing, cfg=cfg)
    opos, opsi = o.boundary_samples(*box)
    np.testing.assert_array_equal(pos, opos.astype(np.float32))
    np.testing.assert_array_equal(psi, opsi.astype(np.float32))
