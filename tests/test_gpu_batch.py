"""Batched scenes: one solver steps many independent scenes, each with its own box and hash table;
every scene's result equals the same scene stepped alone, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
DT = 0.0083


@pytest.fixture(scope="module")
def solvers():
    import torch
    if torch.cuda.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    from akuaengine_amd.solver import BatchSolver, PBFSolver
    return BatchSolver, PBFSolver


def test_batch_equals_independent_scenes(solvers, dam_params):
    BatchSolver, PBFSolver = solvers
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    counts = [100, 4097, 2000, 513, 8000, 1, 777]
    scs = [scenes.batched_scene(n, i, seed=5) for i, n in enumerate(counts)]
    b = BatchSolver(counts, cfg, corr)
    b.set_state(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]))
    for i, s in enumerate(scs):
        b.set_box(i, s.box_min, s.box_max)
    b.step_batch(DT, 3)
    got = b.get_state()
    off = 0
    for i, s in enumerate(scs):
        one = PBFSolver(s.n, cfg, corr)
        one.set_state(s.pos, s.vel)
        one.step_n(DT, s.box_min, s.box_max, 3)
        want = one.get_state()
        sl = slice(off, off + s.n)
        for k in ("pos", "vel", "density", "lam"):
            np.testing.assert_array_equal(got[k][sl], want[k], err_msg=f"scene {i} {k}")
        np.testing.assert_array_equal(got["ids"][sl] - off, want["ids"], err_msg=f"scene {i} ids")
        off += s.n
    mx, mean = b.density_errors()
    for i, s in enumerate(scs):
        one = PBFSolver(s.n, cfg, corr)
        one.set_state(s.pos, s.vel)
        one.step_n(DT, s.box_min, s.box_max, 3)
        m1, a1 = one.density_error()
        assert mx[i] == m1 and mean[i] == pytest.approx(a1, rel=1e-9, abs=1e-12)


def test_256_scene_workload_properties(solvers, dam_params):
    """The 256-scene batch (log-uniform 2,000..200,000 particles): a few steps stay finite and
    every scene keeps its own particles."""
    BatchSolver, _ = solvers
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    scs = scenes.batched_scenes(256, seed=3)
    counts = [s.n for s in scs]
    b = BatchSolver(counts, cfg, corr)
    b.set_state(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]))
    for i, s in enumerate(scs):
        b.set_box(i, s.box_min, s.box_max)
    b.step_batch(DT, 2)
    st = b.get_state()
    assert np.isfinite(st["pos"]).all() and np.isfinite(st["vel"]).all()
    off = 0
    for n in counts:
        ids = st["ids"][off:off + n]
        assert ids.min() == off and ids.max() == off + n - 1 and len(np.unique(ids)) == n
        off += n


def _oracle_state(O, s, cfg, corr, steps):
    o = O.Oracle(s.n, cfg, corr, 32)
    o.set_state(s.pos, s.vel)
    o.step(DT, s.box_min, s.box_max, steps)
    return o.state()


def _assert_scene_equals_oracle(got, off, want, i):
    sl = slice(off, off + len(want["ids"]))
    for k in ("pos", "vel", "density", "lam"):
        np.testing.assert_array_equal(got[k][sl], want[k].astype(np.float32), err_msg=f"scene {i} {k}")
    np.testing.assert_array_equal(got["ids"][sl] - off, want["ids"], err_msg=f"scene {i} ids")


def test_batch_scenes_equal_the_oracle(solvers, oracle_mod, dam_params):
    """Every scene of a 7-scene batch (1 to 8,000 particles, sort-tile edges included) equals the
    FP32 oracle stepping that scene alone, bit for bit."""
    BatchSolver, _ = solvers
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    counts = [100, 4097, 2000, 513, 8000, 1, 777]
    scs = [scenes.batched_scene(n, i, seed=5) for i, n in enumerate(counts)]
    b = BatchSolver(counts, cfg, corr)
    b.set_state(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]))
    for i, s in enumerate(scs):
        b.set_box(i, s.box_min, s.box_max)
    b.step_batch(DT, 4)
    got = b.get_state()
    off = 0
    for i, s in enumerate(scs):
        _assert_scene_equals_oracle(got, off, _oracle_state(oracle_mod, s, cfg, corr, 4), i)
        off += s.n


@pytest.mark.parametrize("mode,level", [(1, 1), (2, 1), (2, 0)])
def test_work_queue_batch_equals_the_oracle(solvers, oracle_mod, dam_params, mode, level):
    """With the sweeps' work queue (tuning workQueue 1 and 2), every scene of the 7-scene batch equals
    the FP32 oracle stepping that scene alone, bit for bit, with multi-cell and single-cell
    super-cells."""
    BatchSolver, _ = solvers
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    counts = [100, 4097, 2000, 513, 8000, 1, 777]
    scs = [scenes.batched_scene(n, i, seed=5) for i, n in enumerate(counts)]
    b = BatchSolver(counts, cfg, corr)
    b.set_tuning(workQueue=mode, layoutLevel=level)
    b.set_state(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]))
    for i, s in enumerate(scs):
        b.set_box(i, s.box_min, s.box_max)
    b.step_batch(DT, 4)
    got = b.get_state()
    off = 0
    for i, s in enumerate(scs):
        _assert_scene_equals_oracle(got, off, _oracle_state(oracle_mod, s, cfg, corr, 4), i)
        off += s.n


def test_256_scene_workload_sample_equals_the_oracle(solvers, oracle_mod, dam_params):
    """The 256-scene workload at full size, two batched steps: the smallest, the largest and six
    seeded random scenes each equal the FP32 oracle stepping that scene alone, bit for bit."""
    BatchSolver, _ = solvers
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    scs = scenes.batched_scenes(256, seed=3)
    counts = [s.n for s in scs]
    b = BatchSolver(counts, cfg, corr)
    b.set_state(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]))
    for i, s in enumerate(scs):
        b.set_box(i, s.box_min, s.box_max)
    b.step_batch(DT, 2)
    got = b.get_state()
    offs = np.concatenate([[0], np.cumsum(counts)[:-1]])
    pick = {int(np.argmin(counts)), int(np.argmax(counts))}
    pick |= set(int(i) for i in np.random.default_rng(17).choice(256, 6, replace=False))
    for i in sorted(pick):
        print(f"[batch 256] scene {i}: {counts[i]} particles", flush=True)
        _assert_scene_equals_oracle(got, int(offs[i]), _oracle_state(oracle_mod, scs[i], cfg, corr, 2), i)
