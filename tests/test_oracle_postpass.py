"""Closed-form pins of the oracle's passes after the constraint solve: vorticity, vorticity
confinement, XSPH viscosity and the six-wall velocity damping (IntegrationCUDA.cu:51-195).

Each check derives its expected values from the mathematics of the reference's formula on a
configuration where it has a closed form (rigid rotation on a lattice, a vorticity magnitude that
grows linearly, two or three hand-placed particles, one particle per wall).  Each check also runs
against the oracle built with a planted mistake (oracle/liboracle_plant.so, mistakes listed in
pbf_oracle.hpp) and must fail there: a sign flip, swapped cross-product operands, m_i in place of
m_j, a dropped 1/rho, Gauss-Seidel in place of Jacobi, a wrong wall normal or point, or walls
applied out of sequence.
"""
import dataclasses

import numpy as np
import pytest

PI32 = float(np.float32(3.14))          # the reference's pi as the FP32 literal 3.14f
H = float(np.float32(0.1))              # smoothRadius as the oracle receives it
KEEP = 1.0 - float(np.float32(0.95))    # tangential factor of friction 0.95 (PBFSolver.cpp:64)
S = 0.052                               # lattice spacing: 2S > h keeps the second shell out


def _cfg(dam_params, **kw):
    cfg, corr = dam_params
    return dataclasses.replace(cfg, **kw), corr


def _maker(O, plants):
    def make(pos, vel, cfg, corr, mass=None, density=None, prec=64):
        pos = np.asarray(pos, np.float32)
        o = O.Oracle(len(pos), cfg, corr, prec, 0, plants=plants)
        o.set_state(pos, np.asarray(vel, np.float32), mass, density)
        o.set_gravity((0.0, 0.0, 0.0))
        return o
    return make


def _stages(o, names, dt, box_min=(-100, -100, -100), box_max=(100, 100, 100)):
    for n in names:
        o.stage(n, dt, box_min, box_max)


def _lattice(n=12, origin=1.0):
    """An n^3 lattice at spacing S (FP32 points) and the integer index of every point."""
    g = np.arange(n)
    idx = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = (np.float32(origin) + idx.astype(np.float32) * np.float32(S)).astype(np.float32)
    return pos, idx


def _interior_clean(st, t, n=12, origin=1.0):
    """Sorted slots whose 26 lattice neighbours are all present, each listed once."""
    idx = np.rint((st["pos"] - origin) / S).astype(int)
    inner = np.all((idx >= 1) & (idx <= n - 2), axis=1)
    clean = np.array([t["cnt"][i] == 26 and len(set(t["nbr"][i, :26])) == 26
                      for i in range(len(idx))])
    return inner & clean, idx


def _grad_norm_times_r(r):
    """|grad W_spiky(r)| * r = 45 / (3.14 h^6) (h - r)^2 r, for 1e-5 <= r <= h."""
    return 45.0 / (PI32 * H ** 6) * (H - r) ** 2 * r


def _grad_spiky(r):
    """grad W_spiky(r) = -45 / (3.14 h^6) (h - |r|)^2 r / |r| (SmoothingKernelsCUDA.h:23-28)."""
    m = np.linalg.norm(r)
    if m > H or m < 1e-5:
        return np.zeros(3)
    return -45.0 / (PI32 * H ** 6) * (H - m) ** 2 * r / m


def _poly6(d2):
    """W_poly6 = 315 / (64 3.14 h^9) (h^2 - d^2)^3 (SmoothingKernelsCUDA.h:16-21)."""
    return 0.0 if d2 > H * H else 315.0 / (64.0 * PI32 * H ** 9) * (H * H - d2) ** 3


SHELLS = [(6, S), (12, S * np.sqrt(2.0)), (8, S * np.sqrt(3.0))]  # face, edge, corner neighbours


# ---- vorticity ----------------------------------------------------------------------------------

def check_vorticity_rigid_rotation(make, dam_params, prec=64):
    """v = Omega x (x - c): v_j - v_i = -Omega x r with r = x_i - x_j, and grad W = -F(r) r, so
    term j is m_j F(r) (Omega r^2 - r (Omega . r)); over a cubic shell sum r r^T = (r^2 / 3) n I,
    hence w_i = (2/3) Omega sum_shells m_shell n |grad W(r)| r.  With checkerboard masses the face
    and corner shells carry the other parity's mass and the edge shell this parity's."""
    cfg, corr = _cfg(dam_params)
    pos, idx = _lattice()
    parity = idx.sum(1) % 2
    mass = np.where(parity == 0, 1.5, 2.5).astype(np.float32)
    om = np.array([0.3, -0.7, 1.1])
    vel = np.cross(om, pos.astype(np.float64) - pos.mean(0))
    o = make(pos, vel, cfg, corr, mass=mass, prec=prec)
    _stages(o, ["predict", "hashes", "sort", "neighbours", "vorticity"], 0.0)
    st, t = o.state(), o.taps()
    sel, sidx = _interior_clean(st, t)
    assert sel.sum() > 500
    par = sidx.sum(1) % 2
    m_same = np.where(par == 0, 1.5, 2.5)
    m_other = np.where(par == 0, 2.5, 1.5)
    k = (m_other * SHELLS[0][0] * _grad_norm_times_r(SHELLS[0][1])
         + m_same * SHELLS[1][0] * _grad_norm_times_r(SHELLS[1][1])
         + m_other * SHELLS[2][0] * _grad_norm_times_r(SHELLS[2][1]))
    want = (2.0 / 3.0) * k[:, None] * om[None, :]
    got = o.stage_arrays()["vort"]
    rtol = 1e-5 if prec == 64 else 1e-4
    np.testing.assert_allclose(got[sel], want[sel], rtol=rtol, atol=rtol * np.abs(want).max())


# ---- vorticity confinement ----------------------------------------------------------------------

def _confinement_lattice(make, dam_params, a, b, rho, eps, dt, prec):
    """Lattice at rest with w_i = (0, 0, a + b x_i) and density rho; returns the velocities after
    the confinement pass, the interior slots and |w_i|."""
    cfg, corr = _cfg(dam_params, vorticityEpsilon=eps)
    pos, _ = _lattice()
    dens = np.full(len(pos), rho, np.float32)
    o = make(pos, np.zeros_like(pos), cfg, corr, density=dens, prec=prec)
    _stages(o, ["predict", "hashes", "sort", "neighbours"], dt)
    st, t = o.state(), o.taps()
    wz = (a + b * st["pos"][:, 0]).astype(np.float32)
    w = np.zeros((len(pos), 3), np.float32)
    w[:, 2] = wz
    o.set_vort(w)
    o.stage("confinement", dt, (-100, -100, -100), (100, 100, 100))
    sel, _ = _interior_clean(st, t)
    assert sel.sum() > 500
    return o.state()["vel"], sel, np.abs(wz.astype(np.float64))


def check_confinement_gradient(make, dam_params, prec=64):
    """|w_i| - |w_j| = b (x_i - x_j), so eta = -(b / (3 rho_i)) sum_shells m n |grad W| r x-hat:
    N = -x-hat for b > 0, and N x (|w_i| z-hat) = |w_i| y-hat, so v gains dt eps |w_i| y-hat."""
    eps, dt = 0.5, 0.01
    v, sel, wabs = _confinement_lattice(make, dam_params, 2.0, 3.0, 8000.0, eps, dt, prec)
    want = np.zeros_like(v)
    want[:, 1] = dt * eps * wabs
    rtol = 1e-6 if prec == 64 else 1e-5
    np.testing.assert_allclose(v[sel], want[sel], rtol=rtol, atol=rtol * want[:, 1].max())


def check_confinement_threshold(make, dam_params, prec=64):
    """|eta| = b K / (3 rho) with K = sum_shells n |grad W(r)| r (unit masses).  The pass skips a
    particle with |eta| < 1e-5: at |eta| = 5e-6 the velocity stays zero, at 2e-5 it moves."""
    rho, eps, dt = 8000.0, 0.5, 0.01
    K = sum(n * _grad_norm_times_r(r) for n, r in SHELLS)
    for target, moves in ((0.5e-5, False), (2e-5, True)):
        b = target * 3.0 * rho / K
        v, sel, wabs = _confinement_lattice(make, dam_params, 1.0, b, rho, eps, dt, prec)
        if moves:
            np.testing.assert_allclose(v[sel, 1], dt * eps * wabs[sel], rtol=1e-5)
        else:
            assert np.all(v[sel] == 0.0)


def check_confinement_three_particles(make, dam_params, prec=64):
    """Three particles with unequal masses, densities and vorticities (all within h):
    eta_p = (1/rho_p) sum_q m_q (|w_p| - |w_q|) grad W(x_p - x_q), N = eta / |eta|,
    v_p = dt eps N x w_p."""
    eps, dt = 0.5, 0.01
    cfg, corr = _cfg(dam_params, vorticityEpsilon=eps)
    pos = np.array([[1.0, 1.0, 1.0], [1.05, 1.0, 1.0], [1.0, 1.06, 1.02]], np.float32)
    mass = np.array([1.0, 1.7, 0.6], np.float32)
    dens = np.array([5000.0, 7000.0, 9000.0], np.float32)
    w0 = np.array([[0.2, -0.4, 1.0], [0.5, 0.1, -0.3], [-1.2, 0.7, 0.4]], np.float32)
    o = make(pos, np.zeros_like(pos), cfg, corr, mass=mass, density=dens, prec=prec)
    _stages(o, ["predict", "hashes", "sort", "neighbours"], dt)
    ids = o.state()["ids"]
    o.set_vort(w0[ids])
    o.stage("confinement", dt, (-100, -100, -100), (100, 100, 100))
    got = np.zeros((3, 3))
    got[ids] = o.state()["vel"]
    p64, w64 = pos.astype(np.float64), w0.astype(np.float64)
    for p in range(3):
        eta = np.zeros(3)
        for q in range(3):
            if q != p:
                diff = np.linalg.norm(w64[p]) - np.linalg.norm(w64[q])
                eta += float(mass[q]) * diff * _grad_spiky(p64[p] - p64[q])
        eta /= float(dens[p])
        assert np.linalg.norm(eta) > 1e-3
        want = float(np.float32(dt)) * eps * np.cross(eta / np.linalg.norm(eta), w64[p])
        np.testing.assert_allclose(got[p], want, rtol=1e-9 if prec == 64 else 1e-5, atol=1e-12)


# ---- XSPH ---------------------------------------------------------------------------------------

def _xsph_run(make, dam_params, pos, vel, mass, dens, c, prec):
    cfg, corr = _cfg(dam_params, viscosity=c)
    o = make(pos, vel, cfg, corr, mass=mass, density=dens, prec=prec)
    _stages(o, ["predict", "hashes", "sort", "neighbours", "xsph"], 0.0)
    st = o.state()
    got = np.zeros((len(pos), 3))
    got[st["ids"]] = st["vel"]
    return got


def check_xsph_pair(make, dam_params, prec=64):
    """Two particles: v_i += c (m_j / rho_j) (v_j - v_i) W(r) and v_j += c (m_i / rho_i) (v_i - v_j) W(r)."""
    c = 0.25
    pos = np.array([[1.0, 1.0, 1.0], [1.04, 1.03, 0.99]], np.float32)
    vel = np.array([[0.1, -0.2, 0.3], [-0.5, 0.4, 0.25]], np.float32)
    mass = np.array([1.3, 0.7], np.float32)
    dens = np.array([6000.0, 9000.0], np.float32)
    got = _xsph_run(make, dam_params, pos, vel, mass, dens, c, prec)
    r = pos[0].astype(np.float64) - pos[1]
    w = _poly6(float(r @ r))
    v = vel.astype(np.float64)
    want = np.array([v[0] + c * (float(mass[1]) / float(dens[1])) * (v[1] - v[0]) * w,
                     v[1] + c * (float(mass[0]) / float(dens[0])) * (v[0] - v[1]) * w])
    np.testing.assert_allclose(got, want, rtol=1e-9 if prec == 64 else 1e-6)
    assert np.abs(want - v).max() > 1e-3  # the pass moves both velocities


def check_xsph_chain(make, dam_params, prec=64):
    """a - b - c on a line, 0.06 apart (a and c are 0.12 > h apart): every particle reads the
    velocities from before the pass (Jacobi), so b's change uses the old v_a and v_c."""
    c = 0.25
    pos = np.array([[1.0, 1.0, 1.0], [1.06, 1.0, 1.0], [1.12, 1.0, 1.0]], np.float32)
    vel = np.array([[0.4, 0.0, 0.1], [-0.3, 0.2, 0.0], [0.1, -0.6, 0.5]], np.float32)
    mass = np.array([1.0, 1.4, 0.8], np.float32)
    dens = np.array([7000.0, 6500.0, 8000.0], np.float32)
    got = _xsph_run(make, dam_params, pos, vel, mass, dens, c, prec)
    v, p = vel.astype(np.float64), pos.astype(np.float64)
    wab = _poly6(float((p[0] - p[1]) @ (p[0] - p[1])))
    wbc = _poly6(float((p[1] - p[2]) @ (p[1] - p[2])))
    mr = mass.astype(np.float64) / dens
    want = np.array([v[0] + c * mr[1] * (v[1] - v[0]) * wab,
                     v[1] + c * (mr[0] * (v[0] - v[1]) * wab + mr[2] * (v[2] - v[1]) * wbc),
                     v[2] + c * mr[1] * (v[1] - v[2]) * wbc])
    np.testing.assert_allclose(got, want, rtol=1e-9 if prec == 64 else 1e-5, atol=1e-12)


# ---- wall damping -------------------------------------------------------------------------------

BOX_MIN = np.array([-0.5, 0.2, -1.0])
BOX_MAX = np.array([1.0, 2.0, 0.3])


def _damping_cases():
    """(position, velocity, expected velocity) for each wall (approaching, resting, leaving), an
    interior particle and a corner reached by two walls in turn."""
    ctr = (BOX_MIN + BOX_MAX) / 2
    cases = []
    tang = [1.0, 3.0]
    for a in range(3):
        others = [b for b in range(3) if b != a]
        for side in (0, 1):
            x = ctr.copy()
            x[a] = BOX_MIN[a] + 0.01 if side == 0 else BOX_MAX[a] - 0.01
            inward = 1.0 if side == 0 else -1.0  # the wall's normal along axis a
            for vn, kind in ((-2.0 * inward, "approaching"), (0.0, "resting"),
                             (2.0 * inward, "leaving")):
                v = np.zeros(3)
                v[a] = vn
                v[others[0]], v[others[1]] = tang
                if kind == "leaving":
                    want = v.copy()
                else:
                    want = np.zeros(3)
                    want[others[0]], want[others[1]] = KEEP * tang[0], KEEP * tang[1]
                cases.append((x, v, want))
    cases.append((ctr.copy(), np.array([1.5, -2.5, 0.75]), np.array([1.5, -2.5, 0.75])))
    corner = ctr.copy()
    corner[0], corner[1] = BOX_MIN[0] + 0.01, BOX_MIN[1] + 0.01
    # x-min first: (0, -2 k, 3 k); then y-min: (0, 0, 3 k^2)
    cases.append((corner, np.array([-1.0, -2.0, 3.0]), np.array([0.0, 0.0, 3.0 * KEEP * KEEP])))
    return cases


def check_damping_six_walls(make, dam_params, prec=64):
    cfg, corr = _cfg(dam_params)
    cases = _damping_cases()
    pos = np.array([c[0] for c in cases], np.float32)
    vel = np.array([c[1] for c in cases], np.float32)
    want = np.array([c[2] for c in cases])
    # predict, then the update gives back v up to rounding (no neighbour stage: no correction);
    # in 10 ms every particle stays within 0.025 of the walls it started near
    o = make(pos, vel, cfg, corr, prec=prec)
    for n in ("predict", "update", "damping"):
        o.stage(n, 0.01, BOX_MIN, BOX_MAX)
    got = o.state()["vel"]
    tol = 1e-9 if prec == 64 else 1e-4
    np.testing.assert_allclose(got, want, rtol=tol, atol=tol)


CHECKS = {
    "vorticity_rigid_rotation": check_vorticity_rigid_rotation,
    "confinement_gradient": check_confinement_gradient,
    "confinement_threshold": check_confinement_threshold,
    "confinement_three_particles": check_confinement_three_particles,
    "xsph_pair": check_xsph_pair,
    "xsph_chain": check_xsph_chain,
    "damping_six_walls": check_damping_six_walls,
}


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("name", sorted(CHECKS))
def test_postpass_closed_form(oracle_mod, dam_params, name, prec):
    CHECKS[name](_maker(oracle_mod, False), dam_params, prec)


def test_planted_build_passes_without_a_plant(oracle_mod, dam_params):
    oracle_mod.set_plant(0)
    for name in sorted(CHECKS):
        CHECKS[name](_maker(oracle_mod, True), dam_params)


# (planted mistake in pbf_oracle.hpp, the check that must catch it)
PLANTED = [
    (1, "vorticity_rigid_rotation"), (2, "vorticity_rigid_rotation"),
    (3, "vorticity_rigid_rotation"),
    (4, "confinement_gradient"), (5, "confinement_gradient"), (6, "confinement_threshold"),
    (7, "confinement_three_particles"),
    (8, "xsph_pair"), (9, "xsph_pair"), (10, "xsph_chain"),
    (11, "damping_six_walls"), (12, "damping_six_walls"), (13, "damping_six_walls"),
    (14, "damping_six_walls"),
]


@pytest.mark.parametrize("plant,name", PLANTED)
def test_a_planted_mistake_is_caught(oracle_mod, dam_params, plant, name):
    oracle_mod.set_plant(plant)
    try:
        with pytest.raises(AssertionError):
            CHECKS[name](_maker(oracle_mod, True), dam_params)
    finally:
        oracle_mod.set_plant(0)
