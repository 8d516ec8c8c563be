"""Headless rendering (akua_render) against a numpy restatement of the reference's GL drawing.

The reference draws point sprites (fluidParticle.vert: size / distance pixels; .frag: discard
outside the disc, colour * exp(-3 d^2)) depth-tested (GL_LESS, buffer order) over a checkered
floor (floor.frag).  The restatement below follows GL's rules for points (clip-volume test of the
centre, fragment centres inside the sprite square), in float32 with the library's own camera
matrix.  It differs from the device only where the two exp implementations round differently, so
the images must agree exactly in 99% of pixels and within one level everywhere (a floating-point
kernel: the tolerance is that of exp).  Known answers pin the geometry: a sprite's size, the
nearer sprite winning, equal depth resolved by draw order, the floor's checker.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = np.float32


@pytest.fixture(scope="module")
def gpu():
    import torch
    if torch.cuda.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    from akuaengine_amd import solver
    return solver


def _reference_render(S, cam, pos, col, size, W, H, floor=True):
    """GL's point sprites and depth test, then the floor and clear colour, in float32."""
    m = S.camera_matrix(cam, W, H).astype(F)
    eye = np.array(cam.position, F)
    key = np.full(W * H, np.iinfo(np.uint64).max, np.uint64)
    info = {}
    for i in range(len(pos)):
        px, py, pz = (F(v) for v in pos[i])
        c = [F(F(F(F(m[r, 0] * px) + F(m[r, 1] * py)) + F(m[r, 2] * pz)) + m[r, 3]) for r in range(4)]
        cw = c[3]
        if not (cw > 0 and abs(c[0]) <= cw and abs(c[1]) <= cw and abs(c[2]) <= cw):
            continue
        nx, ny, nz = F(c[0] / cw), F(c[1] / cw), F(c[2] / cw)
        xw = F(F(F(nx * F(0.5)) + F(0.5)) * F(W))
        yw = F(F(F(ny * F(0.5)) + F(0.5)) * F(H))
        zw = F(F(nz * F(0.5)) + F(0.5))
        d = np.array([px, py, pz], F) - eye
        dist = np.sqrt(F(F(F(d[0] * d[0]) + F(d[1] * d[1])) + F(d[2] * d[2])))
        s = F(min(max(F(F(size[i]) / dist), F(1.0)), F(256.0)))
        x0, y0 = F(xw - F(F(0.5) * s)), F(yw - F(F(0.5) * s))
        info[i] = (x0, y0, s)
        fx0, fx1 = int(np.ceil(F(x0 - F(0.5)))), int(np.ceil(F(F(x0 + s) - F(0.5)))) - 1
        fy0, fy1 = int(np.ceil(F(y0 - F(0.5)))), int(np.ceil(F(F(y0 + s) - F(0.5)))) - 1
        k = (np.uint64(np.array(zw, F).view(np.uint32)) << np.uint64(32)) | np.uint64(i)
        for fy in range(max(fy0, 0), min(fy1, H - 1) + 1):
            for fx in range(max(fx0, 0), min(fx1, W - 1) + 1):
                ox = F(F(F(F(F(fx) + F(0.5)) - x0) / s) - F(0.5))
                oy = F(F(F(F(F(fy) + F(0.5)) - y0) / s) - F(0.5))
                if F(F(ox * ox) + F(oy * oy)) > F(0.25):
                    continue
                p = (H - 1 - fy) * W + fx
                key[p] = min(key[p], k)
    img = np.zeros((H, W, 3), np.uint8)
    # the basis the device uses for the floor rays
    y, p_ = np.radians(cam.yaw), np.radians(cam.pitch)
    f = np.array([np.cos(y) * np.cos(p_), np.sin(p_), np.sin(y) * np.cos(p_)]); f /= np.linalg.norm(f)
    r = np.cross(f, [0, 1, 0]); r /= np.linalg.norm(r)
    u = np.cross(r, f); u /= np.linalg.norm(u)
    s_ = np.cross(f, u); s_ /= np.linalg.norm(s_)
    uu = np.cross(s_, f)
    f, s_, uu = f.astype(F), s_.astype(F), uu.astype(F)
    th, asp = F(np.tan(np.radians(cam.fovY) / 2)), F(W / H)
    for row in range(H):
        fy = H - 1 - row
        for fx in range(W):
            rgb = [F(0.15), F(0.15), F(0.18)]
            hit, zf = False, F(1.0)
            if floor:
                nx = F(F(F(F(fx) + F(0.5)) / F(W)) * F(2.0)) - F(1.0)
                ny = F(F(F(F(fy) + F(0.5)) / F(H)) * F(2.0)) - F(1.0)
                a, c = F(F(nx * th) * asp), F(ny * th)
                dv = [F(F(f[k] + F(a * s_[k])) + F(c * uu[k])) for k in range(3)]
                if dv[1] != 0:
                    t = F(F(F(-0.02) - eye[1]) / dv[1])
                    hx, hz = F(eye[0] + F(t * dv[0])), F(eye[2] + F(t * dv[2]))
                    if t > 0 and 0 <= hx <= 100 and 0 <= hz <= 100:
                        cz = F(F(F(F(m[2, 0] * hx) + F(m[2, 1] * F(-0.02))) + F(m[2, 2] * hz)) + m[2, 3])
                        cw = F(F(F(F(m[3, 0] * hx) + F(m[3, 1] * F(-0.02))) + F(m[3, 2] * hz)) + m[3, 3])
                        if cw > 0 and abs(cz) <= cw:
                            hit, zf = True, F(F(F(cz / cw) * F(0.5)) + F(0.5))
                            chk = (int(np.floor(hx)) + int(np.floor(hz))) & 1
                            rgb = [F(0.5 if chk else 0.8)] * 3
            kk = key[row * W + fx]
            if kk != np.iinfo(np.uint64).max:
                zp = np.array(np.uint32(kk >> np.uint64(32))).view(F)
                if not hit or zp < zf:
                    i = int(kk & np.uint64(0xFFFFFFFF))
                    x0, y0, s = info[i]
                    ox = F(F(F(F(F(fx) + F(0.5)) - x0) / s) - F(0.5))
                    oy = F(F(F(F(F(fy) + F(0.5)) - y0) / s) - F(0.5))
                    al = F(np.exp(F(F(-3.0) * F(F(ox * ox) + F(oy * oy)))))
                    rgb = [F(col[i][k] * al) for k in range(3)]
            img[row, fx] = [int(np.rint(min(max(v, 0.0), 1.0) * 255.0)) for v in rgb]
    return img


def _solver_with(S, pos, col, size):
    from akuaengine_amd import _lib
    from akuaengine_amd.config import LambdaCorrParams, PBFConfig
    n = len(pos)
    recs = (_lib.AkuaParticle * n)()
    for i in range(n):
        recs[i].position[:] = [float(v) for v in pos[i]]
        recs[i].mass = 1.0
        recs[i].color[:] = [float(v) for v in col[i]]
        recs[i].size = float(size[i])
    s = S.PBFSolver(n, PBFConfig(), LambdaCorrParams())
    s.set_particles(recs)
    return s


def test_matches_gl_restatement(gpu):
    S = gpu
    rng = np.random.default_rng(3)
    n = 300
    pos = np.stack([rng.uniform(1.5, 4.5, n), rng.uniform(0.0, 3.0, n), rng.uniform(-2.0, 4.5, n)], 1).astype(F)
    col = np.concatenate([rng.uniform(0, 1, (n, 3)), np.ones((n, 1))], 1).astype(F)
    size = rng.uniform(20, 80, n).astype(F)
    W, H = 160, 90
    s = _solver_with(S, pos, col, size)
    cam = S.default_camera()
    img = s.render(cam, W, H, point_size=0.0)
    # the solver keeps the particles in upload (draw) order before its first step
    ref = _reference_render(S, cam, pos, col, size, W, H)
    diff = np.abs(img.astype(int) - ref.astype(int))
    assert diff.max() <= 1
    assert (diff.max(axis=2) == 0).mean() >= 0.99
    assert (img != [38, 38, 46]).any(axis=2).mean() > 0.2          # something was drawn


def test_single_sprite_known_answer(gpu):
    S = gpu
    # a particle on the gaze line 5 away: a 10-pixel sprite (size 50 / distance 5) in the middle
    s = _solver_with(S, np.array([[3.0, 2.0, 1.0]], F), np.array([[0, 0, 1, 1]], F), np.array([50.0], F))
    W, H = 1280, 720
    img = s.render(S.default_camera(), W, H, point_size=0.0, floor=False)
    blue = (img[..., 2] > 46) & (img[..., 0] == 0)
    ys, xs = np.nonzero(blue)
    assert 60 <= blue.sum() <= 90                                   # about pi * 5^2
    assert xs.min() >= 634 and xs.max() <= 645 and ys.min() >= 354 and ys.max() <= 365
    assert tuple(img[0, 0]) == (38, 38, 46)                          # the clear colour
    assert img[360, 640, 2] >= 250                                   # near the centre alpha ~ 1


def test_nearer_sprite_and_draw_order_win(gpu):
    S = gpu
    pos = np.array([[3.0, 2.0, 0.0], [3.0, 2.0, 1.0], [3.5, 2.0, 1.0], [3.5, 2.0, 1.0]], F)
    col = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1], [1, 1, 0, 1]], F)
    s = _solver_with(S, pos, col, np.full(4, 50.0, F))
    W, H = 1280, 720
    img = s.render(S.default_camera(), W, H, point_size=0.0, floor=False)
    # the centre: the green particle at distance 5 hides the red one at 6
    assert img[360, 640, 1] >= 250 and img[360, 640, 0] == 0
    # two particles at the same point: the first in buffer order (blue) is kept
    m = S.camera_matrix(S.default_camera(), W, H)
    c = m @ np.array([3.5, 2.0, 1.0, 1.0])
    fx, fy = int((c[0] / c[3] * 0.5 + 0.5) * W), int((c[1] / c[3] * 0.5 + 0.5) * H)
    px = img[H - 1 - fy, fx]
    assert px[2] > 200 and px[0] == 0 and px[1] == 0


def test_floor_checker(gpu):
    S = gpu
    # one particle behind the camera (clipped): only the floor and the clear colour remain
    s = _solver_with(S, np.array([[3.0, 2.0, 9.0]], F), np.array([[0, 0, 1, 1]], F), np.array([50.0], F))
    W, H = 64, 36
    cam = S.default_camera()
    img = s.render(cam, W, H, point_size=0.0)
    ref = _reference_render(S, cam, np.array([[3.0, 2.0, 9.0]], F), np.array([[0, 0, 1, 1]], F),
                            np.array([50.0], F), W, H)
    np.testing.assert_array_equal(img, ref)
    assert (img[-1] == 204).all(axis=1).any() or (img[-1] == 128).all(axis=1).any()   # floor greys
    assert tuple(img[0, 0]) == (38, 38, 46)                                          # sky
