"""Scene generators: deterministic, the reference's lattice order and FP32 arithmetic
(Application.cpp:170-182), jitter inside its bound."""
import numpy as np

from akuaengine_amd import scenes


def test_splitmix64_known_answer():
    # SplitMix64 seeded with 0: first outputs (Steele, Lea, Flood 2014 reference implementation)
    z = scenes.splitmix64(0, 3)
    assert [int(x) for x in z] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_lattice_order_and_values():
    p = scenes.lattice(3, 4, 5, 0.05, (2.0, 1.0, 2.0))
    assert p.dtype == np.float32 and p.shape == (60, 3)
    # x-major, then y, then z innermost (Application.cpp:170-172)
    i = (2 * 4 + 3) * 5 + 4
    want = np.float32(2.0) + np.float32(2) * np.float32(0.05), np.float32(1.0) + np.float32(3) * np.float32(0.05), \
        np.float32(2.0) + np.float32(4) * np.float32(0.05)
    assert tuple(p[i]) == want


def test_scene_sizes_and_boxes():
    s0 = scenes.dam_break_27k()
    assert s0.n == 27000 and s0.box_max == (2.0, 1.5, 1.0)
    assert s0.pos[:, 2].max() > 1.0  # the 30-layer lattice is deeper than the 1.0-deep box
    counts = scenes.batched_counts()
    assert len(counts) == 256 and min(counts) >= 2000 and max(counts) <= 200000
    s = scenes.batched_scene(counts[0], 0)
    assert s.n == counts[0]


def test_jitter_bounds_and_determinism():
    a = scenes.lattice(10, 10, 10, 0.05, (0.025, 0.025, 0.025), jitter=0.005, seed=1)
    b = scenes.lattice(10, 10, 10, 0.05, (0.025, 0.025, 0.025), jitter=0.005, seed=1)
    c = scenes.lattice(10, 10, 10, 0.05, (0.025, 0.025, 0.025))
    assert np.array_equal(a, b)
    d = np.abs(a - c)
    assert d.max() <= 0.005 + 1e-6 and d.max() > 0.004
