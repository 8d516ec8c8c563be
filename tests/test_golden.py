"""The FP32 oracle reproduces the committed golden fixture bit for bit (tests/golden/)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dam_break_512.npz")


def test_oracle_reproduces_golden(oracle_mod, dam_params):
    g = np.load(GOLDEN)
    cfg, corr = dam_params
    n = g["pos0"].shape[0]
    o = oracle_mod.Oracle(n, cfg, corr, 32)
    o.set_state(g["pos0"], g["vel0"])
    dt, bmin, bmax = float(g["dt"]), g["box_min"], g["box_max"]
    o.step(dt, bmin, bmax, 1)
    t, s = o.taps(), o.state()
    np.testing.assert_array_equal(t["sorted_hash"], g["step1_sorted_hash"])
    np.testing.assert_array_equal(t["perm"], g["step1_perm"])
    np.testing.assert_array_equal(t["cnt"], g["step1_cnt"])
    for k in ("pos", "vel", "density", "lam"):
        np.testing.assert_array_equal(s[k].astype(np.float32), g["step1_" + k])
    o.step(dt, bmin, bmax, 9)
    s = o.state()
    np.testing.assert_array_equal(s["pos"].astype(np.float32), g["step10_pos"])
    np.testing.assert_array_equal(s["vel"].astype(np.float32), g["step10_vel"])
    np.testing.assert_array_equal(s["ids"], g["step10_ids"])
