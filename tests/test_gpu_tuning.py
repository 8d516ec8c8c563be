"""Tuning choices (workgroup size of the neighbour kernels, physical super-cell size, radix digit
widths, layout interval, the balanced work order of the neighbour sweeps, the slab walk, the
first sweep fused into the list build, lane groups, the sweeps' work queue) change only speed: every
setting yields bit-identical trajectories, and the autotuner leaves the particle state as it found
it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
DT = 0.0083
SETTINGS = [
    dict(block=256, layoutLevel=1, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=0, laneGroup=0, workQueue=0),
    dict(block=128, layoutLevel=0, hashDigitBits=5, layoutDigitBits=5, layoutInterval=3,
         balanceTile=1024, slabBatch=8, fuseSweep=1, laneGroup=0, workQueue=0),
    dict(block=512, layoutLevel=2, hashDigitBits=8, layoutDigitBits=8, layoutInterval=2,
         balanceTile=256, slabBatch=4, fuseSweep=0, laneGroup=0, workQueue=0),
    dict(block=256, layoutLevel=3, hashDigitBits=6, layoutDigitBits=0, layoutInterval=8,
         balanceTile=2048, slabBatch=0, fuseSweep=0, laneGroup=0, workQueue=0),
    dict(block=128, layoutLevel=1, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=768, slabBatch=8, fuseSweep=0, laneGroup=0, workQueue=0),
    dict(block=128, layoutLevel=0, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=1, laneGroup=0, workQueue=0),
    dict(block=512, layoutLevel=1, hashDigitBits=6, layoutDigitBits=8, layoutInterval=2,
         balanceTile=0, slabBatch=0, fuseSweep=1, laneGroup=0, workQueue=0),
    dict(block=128, layoutLevel=1, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=1, laneGroup=8, workQueue=0),
    dict(block=256, layoutLevel=0, hashDigitBits=0, layoutDigitBits=6, layoutInterval=3,
         balanceTile=0, slabBatch=0, fuseSweep=0, laneGroup=16, workQueue=0),
    dict(block=512, layoutLevel=2, hashDigitBits=5, layoutDigitBits=0, layoutInterval=1,
         balanceTile=512, slabBatch=0, fuseSweep=0, laneGroup=4, workQueue=0),
    dict(block=256, layoutLevel=1, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=0, laneGroup=0, workQueue=1),
    dict(block=128, layoutLevel=0, hashDigitBits=5, layoutDigitBits=0, layoutInterval=2,
         balanceTile=1024, slabBatch=8, fuseSweep=0, laneGroup=0, workQueue=2),
    dict(block=512, layoutLevel=1, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=1, laneGroup=0, workQueue=1),
    dict(block=256, layoutLevel=1, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=0, laneGroup=8, workQueue=2),
    dict(block=128, layoutLevel=0, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=1, laneGroup=0, workQueue=2),
    dict(block=128, layoutLevel=0, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=1, laneGroup=4, workQueue=1),
    dict(block=256, layoutLevel=1, hashDigitBits=0, layoutDigitBits=0, layoutInterval=1,
         balanceTile=0, slabBatch=0, fuseSweep=0, laneGroup=2, workQueue=0),
]


@pytest.fixture(scope="module")
def classes():
    import torch
    if torch.cuda.device_count() < 1:
        pytest.fail("no GPU visible: the gpu tests need an MI355X")
    from akuaengine_amd.solver import BatchSolver, PBFSolver
    return BatchSolver, PBFSolver


def _run(PBFSolver, sc, cfg, corr, steps, tuning=None):
    s = PBFSolver(sc.n, cfg, corr)
    if tuning:
        s.set_tuning(**tuning)
    s.set_state(sc.pos, sc.vel)
    s.step_n(DT, sc.box_min, sc.box_max, steps)
    return s


@pytest.mark.parametrize("tuning", SETTINGS[1:], ids=lambda t: "b{block}-L{layoutLevel}-h{hashDigitBits}-l{layoutDigitBits}-i{layoutInterval}-t{balanceTile}-s{slabBatch}-f{fuseSweep}-g{laneGroup}-q{workQueue}".format(**t))
def test_every_setting_gives_identical_trajectories(classes, dam_params, tuning):
    _, PBFSolver = classes
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    sc = scenes.dam_break_27k()
    want = _run(PBFSolver, sc, cfg, corr, 20).get_state()
    s = _run(PBFSolver, sc, cfg, corr, 20, tuning)
    assert s.get_tuning() == tuning
    got = s.get_state()
    for k in ("pos", "vel", "density", "lam", "ids"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


@pytest.mark.parametrize("scene", ["scaled_16", "scaled_60"])
def test_paired_sorts_keep_the_layout(classes, scene):
    """Single-cell super-cells (layout level 0) with equal digit widths sort the layout keys in
    the previous logical order, in the hash sort's launches; with widths that differ, in the new
    logical order after it.  Within one cell both orders agree, so the physical order itself (not
    only the results) must be the same while no key is clamped.  The reference application's
    parameters (PBFConfig defaults, 2 ms steps) keep every predicted position within the key
    range's margin beyond the walls; the benchmark parameters fling single particles far past it
    (hundreds of m/s), and there a clamped key's particles may differ in order.  4,096 particles
    take the look-back sort, 216,000 the three-launch one."""
    _, PBFSolver = classes
    from akuaengine_amd import scenes
    from akuaengine_amd.config import LambdaCorrParams, PBFConfig
    cfg, corr = PBFConfig(), LambdaCorrParams()
    sc = scenes.dam_break_scaled(16 if scene == "scaled_16" else 60, seed=1)
    out = []
    for digits in (0, 5):  # 0: the hash sort's width (paired); 5: never the hash sort's here
        s = PBFSolver(sc.n, cfg, corr)
        s.set_tuning(block=256, layoutLevel=0, hashDigitBits=8, layoutDigitBits=digits)
        s.set_state(sc.pos, sc.vel)
        s.step_n(0.002, sc.box_min, sc.box_max, 20)
        out.append((s.debug_physical_order(), s.get_state()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    for k in ("pos", "vel", "ids"):
        np.testing.assert_array_equal(out[0][1][k], out[1][1][k], err_msg=k)


def test_setting_changes_between_steps(classes, dam_params):
    _, PBFSolver = classes
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    sc = scenes.dam_break_scaled(12)
    seq = SETTINGS + SETTINGS[:2]
    want = _run(PBFSolver, sc, cfg, corr, len(seq)).get_state()
    s = PBFSolver(sc.n, cfg, corr)
    s.set_state(sc.pos, sc.vel)
    for t in seq:
        s.set_tuning(**t)
        s.step(DT, sc.box_min, sc.box_max)
    got = s.get_state()
    for k in ("pos", "vel", "ids"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)


def test_slot_limit_is_refused(classes, dam_params):
    """A solver holds at most 2^28 particle slots (per-slot arrays are addressed by 32-bit byte
    offsets): one scene or a batch past that is refused before any allocation."""
    import dataclasses
    BatchSolver, PBFSolver = classes
    from akuaengine_amd._lib import AkuaError
    cfg, corr = dam_params
    cfg = dataclasses.replace(cfg, maxNeighbours=1)  # keeps maxNeighbours * N within 32 bits
    with pytest.raises(AkuaError) as e:
        PBFSolver((1 << 28) + 1, cfg, corr)
    assert e.value.status == 6  # AKUA_ERR_UNSUPPORTED
    with pytest.raises(AkuaError) as e:
        BatchSolver([1 << 27, (1 << 27) + 1], cfg, corr)
    assert e.value.status == 6


def test_invalid_settings_are_rejected(classes, dam_params):
    _, PBFSolver = classes
    from akuaengine_amd._lib import AkuaError
    cfg, corr = dam_params
    s = PBFSolver(1000, cfg, corr)
    for bad in (dict(block=64), dict(layoutLevel=4), dict(hashDigitBits=4), dict(layoutDigitBits=9),
                dict(layoutInterval=0), dict(layoutInterval=9), dict(balanceTile=100),
                dict(balanceTile=4096), dict(balanceTile=-256), dict(slabBatch=2),
                dict(slabBatch=16), dict(fuseSweep=2), dict(fuseSweep=-1), dict(laneGroup=3),
                dict(laneGroup=32), dict(laneGroup=1), dict(workQueue=3), dict(workQueue=-1)):
        with pytest.raises(AkuaError):
            s.set_tuning(**bad)
    assert s.get_tuning() == SETTINGS[0]


def test_autotune_restores_state_and_uses_cache(classes, dam_params, tmp_path):
    _, PBFSolver = classes
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    sc = scenes.dam_break_27k()
    s = _run(PBFSolver, sc, cfg, corr, 3)
    before = s.get_state()
    cache = str(tmp_path / "tuning.txt")
    best = s.autotune(DT, sc.box_min, sc.box_max, trialSteps=2, cachePath=cache)
    assert s.get_tuning() == best
    after = s.get_state()
    for k in ("pos", "vel", "density", "lam", "ids"):
        np.testing.assert_array_equal(after[k], before[k], err_msg=k)
    s.step_n(DT, sc.box_min, sc.box_max, 4)
    want = _run(PBFSolver, sc, cfg, corr, 7).get_state()
    got = s.get_state()
    for k in ("pos", "vel", "ids"):
        np.testing.assert_array_equal(got[k], want[k], err_msg=k)
    lines = open(cache).read().splitlines()
    assert len(lines) == 1 and " | " in lines[0]
    s2 = _run(PBFSolver, sc, cfg, corr, 1)
    assert s2.autotune(DT, trialSteps=2, cachePath=cache) == best
    assert open(cache).read().splitlines() == lines  # a hit appends nothing


def test_batch_settings_are_invariant(classes, dam_params):
    BatchSolver, _ = classes
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    counts = [300, 4097, 2000, 9000]
    scs = [scenes.batched_scene(n, i, seed=9) for i, n in enumerate(counts)]
    out = []
    for t in (SETTINGS[1], SETTINGS[0], SETTINGS[2], SETTINGS[3], SETTINGS[4], SETTINGS[5],
              SETTINGS[7], SETTINGS[8], SETTINGS[10], SETTINGS[11], SETTINGS[12], SETTINGS[14]):
        b = BatchSolver(counts, cfg, corr)
        b.set_tuning(**t)
        b.set_state(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]))
        for i, s in enumerate(scs):
            b.set_box(i, s.box_min, s.box_max)
        b.step_batch(DT, 4)
        out.append(b.get_state())
    for o in out[1:]:
        for k in ("pos", "vel", "ids"):
            np.testing.assert_array_equal(o[k], out[0][k], err_msg=k)


def test_batch_paired_sorts_with_three_launches(classes, dam_params):
    """A batch of more than 64 sort tiles (150,778 particles in five scenes) at layout level 0:
    the hash and layout sorts share three launches per pass, scene-segmented.  Bit-identical to
    multi-cell super-cells (sequential sorts)."""
    BatchSolver, _ = classes
    from akuaengine_amd import scenes
    cfg, corr = dam_params
    counts = [60000, 50000, 40000, 1, 777]
    scs = [scenes.batched_scene(n, i, seed=3) for i, n in enumerate(counts)]
    out = []
    for level in (1, 0):
        b = BatchSolver(counts, cfg, corr)
        b.set_tuning(block=256, layoutLevel=level)
        b.set_state(np.concatenate([s.pos for s in scs]), np.concatenate([s.vel for s in scs]))
        for i, s in enumerate(scs):
            b.set_box(i, s.box_min, s.box_max)
        b.step_batch(DT, 3)
        out.append(b.get_state())
    for k in ("pos", "vel", "ids", "density", "lam"):
        np.testing.assert_array_equal(out[1][k], out[0][k], err_msg=k)
