"""Trains the warm-start network on the GPU with the library's own kernels.

Recipe (the learned-workload definition): dam-break scenes shaped like the 1M workload (100^3
lattice, spacing 0.05, jitter +-0.005, box [0,10]x[0,6]x[0,5]) with seeds other than the
evaluation seed 1, stepped by the plain 4-iteration solver while the network trains online on
their features and targets: 8 scenes x 500 steps = 4,000 solver steps, 65,536 sampled particles
per step, minibatches of 8,192, 2 epochs per step, Adam with the learning rate lowered per scene.
The plain solver's own residual reduction over the same steps calibrates the gate ratio (its 90th
percentile).

    python -m akuaengine_amd.train_warm_start --iterations 1 --out akuaengine_amd/networks/x.npy
"""
from __future__ import annotations

import argparse
import json
import time
from dataclasses import asdict, replace

import numpy as np

from . import learned, scenes
from .config import LambdaCorrParams, PBFConfig, dam_break_params
from .solver import PBFSolver

PHYSICS = {
    # the reference application's solver parameters (PBFConfig.h:10-29) at a 2 ms substep
    "reference": lambda: (PBFConfig(), LambdaCorrParams(), 0.002),
    # the dam-break workload's parameter set at its 8.3 ms step
    "dam_break": lambda: (*dam_break_params(4), 0.0083),
}


def train(physics: str = "reference", iterations: int = 1, seeds=tuple(range(11, 19)),
          steps: int = 500, lrs=(1e-3, 5e-4, 2.5e-4, 1.25e-4, 1e-4, 8e-5, 6e-5, 5e-5),
          n_side: int = 100, log=print):
    cfg, corr, dt = PHYSICS[physics]()
    solver = None
    ratios, losses = [], []
    t0 = time.perf_counter()
    for k, seed in enumerate(seeds):
        sc = scenes.dam_break_1m(seed=seed, n_side=n_side)
        if solver is None:
            solver = PBFSolver(sc.n, cfg, corr)
        solver.set_state(sc.pos, sc.vel)
        loss, res = solver.train(dt, sc.box_min, sc.box_max, steps, samplesPerStep=65536,
                                 minibatch=8192, epochs=2, learningRate=lrs[min(k, len(lrs) - 1)],
                                 seed=1000 + seed, normSteps=20, targetIterations=iterations)
        ok = res[:, 0] > 0
        ratios.append(res[ok, 2] / res[ok, 0])
        losses.append(loss)
        log(f"scene seed {seed}: loss {np.nanmean(loss[:50]):.4f} -> {np.nanmean(loss[-50:]):.4f}")
    ratio = np.concatenate(ratios)
    loss = np.concatenate(losses)
    meta = dict(physics=physics, dt=dt, iterations=iterations, seeds=list(seeds),
                steps_per_scene=steps, particles=int(sc.n), learning_rates=list(lrs[:len(seeds)]),
                samples_per_step=65536, minibatch=8192, epochs=2,
                loss_first50=float(np.nanmean(loss[:50])), loss_last50=float(np.nanmean(loss[-50:])),
                gate_ratio_p50=float(np.percentile(ratio, 50)),
                gate_ratio_p90=float(np.percentile(ratio, 90)),
                seconds=time.perf_counter() - t0,
                config=dict(asdict(cfg), **{"corr_" + k: v for k, v in asdict(corr).items()}))
    return solver.get_network(), meta


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--physics", default="reference", choices=sorted(PHYSICS))
    ap.add_argument("--iterations", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--seeds", default="11,12,13,14,15,16,17,18")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    net, meta = train(a.physics, a.iterations, tuple(int(s) for s in a.seeds.split(",")), a.steps)
    learned.save_network(a.out, net, meta)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()
