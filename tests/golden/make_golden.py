"""Writes tests/golden/dam_break_512.npz: a small dam break stepped by the FP32 CPU oracle.

The reference has no golden vectors of its own (it ships no tests), so this fixture records the
oracle's output once; tests/test_golden.py checks that the oracle still reproduces it bit for bit
and tests/test_gpu_parity.py checks that the HIP solver does.

Contents: the input state (an 8x8x8 jittered lattice), the dam-break parameters, and after step 1
the sorted hashes, permutation, neighbour counts and state; after step 10 the state.
Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from akuaengine_amd.config import dam_break_params  # noqa: E402
from akuaengine_amd.scenes import dam_break_scaled  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dam_break_512.npz")
DT = 0.0083


def main():
    cfg, corr = dam_break_params(4)
    sc = dam_break_scaled(8, seed=7)
    o = O.Oracle(sc.n, cfg, corr, 32)
    o.set_state(sc.pos, sc.vel, sc.mass)
    o.step(DT, sc.box_min, sc.box_max, 1)
    t = o.taps()
    s1 = o.state()
    o.step(DT, sc.box_min, sc.box_max, 9)
    s10 = o.state()
    f32 = lambda a: np.asarray(a, np.float32)
    np.savez_compressed(
        OUT, pos0=sc.pos, vel0=sc.vel, box_min=np.array(sc.box_min, np.float32),
        box_max=np.array(sc.box_max, np.float32), dt=np.float32(DT),
        step1_sorted_hash=t["sorted_hash"], step1_perm=t["perm"], step1_cnt=t["cnt"],
        step1_pos=f32(s1["pos"]), step1_vel=f32(s1["vel"]), step1_density=f32(s1["density"]),
        step1_lam=f32(s1["lam"]), step1_ids=s1["ids"],
        step10_pos=f32(s10["pos"]), step10_vel=f32(s10["vel"]), step10_ids=s10["ids"])
    print("wrote", OUT)


if __name__ == "__main__":
    main()
