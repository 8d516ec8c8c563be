"""Deterministic dam-break scene generators (the headless replacement of
Application::prepareDamBreak, Application.cpp:126-204).

The reference fills a lattice in x-major, then y, then z order (Application.cpp:170-176) with
``position = minPos + index * spacing`` in FP32; the generators below do the same, then add an
optional per-axis uniform jitter in [-j, +j] drawn from SplitMix64 (seeded, so every scene is
reproducible bit-for-bit).  Velocities start at zero and every mass is 1 (Application.cpp:184-185).

Workloads:
  * ``app_dam_break``   -- the reference application's own scene (Application.cpp:12-16, 159-190):
                           a lattice from (2, 1, 2) to (3.5, 2.5, 3.5) at the particle spacing
                           (30^3 at 0.05), box [1.5,4.5]x[0,4]x[1.5,4.5], time step 0.016, the
                           reference's default parameters.
  * ``dam_break_27k``   -- 30^3 lattice, spacing 0.05, origin 0.025, box [0,2]x[0,1.5]x[0,1]
                           (the lattice reaches z = 1.475, so the top 10 z-layers start outside
                           the box and are pushed in by the wall clamp).
  * ``dam_break_1m``    -- 100^3 lattice, jitter +-0.005 (seed 1), box [0,10]x[0,6]x[0,5].
  * ``dam_break_8m``    -- two 200x200x100 columns at opposite x ends, jitter +-0.005 (seed 2),
                           box [0,30]x[0,12]x[0,5].
  * ``batched_scenes``  -- 256 scenes with log-uniform counts in [2000, 200000] (seed 3), each a
                           near-cubic lattice in a box of 2 x 1.5 x 1 times the lattice extent.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int) -> np.ndarray:
    """The first ``count`` outputs of SplitMix64 seeded with ``seed`` (uint64)."""
    k = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, count: int) -> np.ndarray:
    """``count`` FP32 values uniform in [0, 1): the top 24 bits of each SplitMix64 output."""
    z = splitmix64(seed, count)
    return (z >> np.uint64(40)).astype(np.float32) * np.float32(1.0 / (1 << 24))


@dataclass
class Scene:
    name: str
    pos: np.ndarray            # (N, 3) float32
    vel: np.ndarray            # (N, 3) float32
    mass: np.ndarray           # (N,) float32
    box_min: Tuple[float, float, float]
    box_max: Tuple[float, float, float]

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])


def lattice(nx: int, ny: int, nz: int, spacing: float, origin, jitter: float = 0.0,
            seed: int = 0, count: int | None = None) -> np.ndarray:
    """Positions of an nx x ny x nz lattice in x-major order, FP32 (Application.cpp:170-182)."""
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1).astype(np.float32)
    if count is not None:
        idx = idx[:count]
    s = np.float32(spacing)
    org = np.asarray(origin, dtype=np.float32)
    pos = org[None, :] + idx * s
    if jitter > 0.0:
        u = uniform01(seed, pos.size).reshape(pos.shape)
        pos = pos + (u * np.float32(2.0) - np.float32(1.0)) * np.float32(jitter)
    return pos.astype(np.float32)


def _scene(name, pos, box_min, box_max) -> Scene:
    n = pos.shape[0]
    return Scene(name, np.ascontiguousarray(pos, dtype=np.float32),
                 np.zeros((n, 3), np.float32), np.ones(n, np.float32),
                 tuple(float(b) for b in box_min), tuple(float(b) for b in box_max))


APP_DT = 0.016  # Application::_deltaTime (Application.h:32)


def app_dam_break(spacing: float = 0.05) -> Scene:
    """Application::prepareDamBreak (Application.cpp:159-190): numX = int((maxPos - minPos).x /
    particle_spacing) in FP32, position = minPos + index * spacing, box BOX_MIN..BOX_MAX
    (Application.cpp:14-15)."""
    s = np.float32(spacing)
    lo = np.array([2.0, 1.0, 2.0], np.float32)
    hi = np.array([3.5, 2.5, 3.5], np.float32)
    dims = hi - lo
    nx, ny, nz = (int(np.float32(dims[a] / s)) for a in range(3))
    pos = lattice(nx, ny, nz, float(s), lo)
    return _scene("app_dam_break", pos, (1.5, 0.0, 1.5), (4.5, 4.0, 4.5))


def dam_break_27k() -> Scene:
    pos = lattice(30, 30, 30, 0.05, (0.025, 0.025, 0.025))
    return _scene("dam_break_27k", pos, (0.0, 0.0, 0.0), (2.0, 1.5, 1.0))


def dam_break_1m(seed: int = 1, n_side: int = 100) -> Scene:
    pos = lattice(n_side, n_side, n_side, 0.05, (0.025, 0.025, 0.025), jitter=0.005, seed=seed)
    return _scene("dam_break_1m", pos, (0.0, 0.0, 0.0), (10.0, 6.0, 5.0))


def dam_break_8m(seed: int = 2) -> Scene:
    a = lattice(200, 200, 100, 0.05, (0.025, 0.025, 0.025), jitter=0.005, seed=seed)
    b = lattice(200, 200, 100, 0.05, (20.025, 0.025, 0.025), jitter=0.005, seed=seed + 1000)
    return _scene("dam_break_8m", np.concatenate([a, b]), (0.0, 0.0, 0.0), (30.0, 12.0, 5.0))


def dam_break_scaled(n_side: int, seed: int = 1, jitter: float = 0.005) -> Scene:
    """A cube-lattice dam break of n_side^3 particles, box scaled like the 1M scene."""
    ext = n_side * 0.05
    pos = lattice(n_side, n_side, n_side, 0.05, (0.025, 0.025, 0.025), jitter=jitter, seed=seed)
    return _scene(f"dam_break_{n_side}^3", pos, (0.0, 0.0, 0.0), (2.0 * ext, 1.2 * ext, ext))


def batched_counts(num_scenes: int = 256, seed: int = 3, lo: int = 2000, hi: int = 200000) -> List[int]:
    u = uniform01(seed, num_scenes).astype(np.float64)
    return [int(round(math.exp(math.log(lo) + x * (math.log(hi) - math.log(lo))))) for x in u]


def batched_scene(n: int, index: int = 0, seed: int = 3) -> Scene:
    nx = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
    ny = nx
    nz = int(math.ceil(n / (nx * ny)))
    pos = lattice(nx, ny, nz, 0.05, (0.025, 0.025, 0.025), jitter=0.005,
                  seed=seed * 1000003 + index, count=n)
    ext = np.array([nx, ny, nz], np.float64) * 0.05
    box_max = (2.0 * ext[0], 1.5 * ext[1], 1.0 * ext[2])
    return _scene(f"batched_{index}", pos, (0.0, 0.0, 0.0), box_max)


def batched_scenes(num_scenes: int = 256, seed: int = 3) -> List[Scene]:
    return [batched_scene(n, i, seed) for i, n in enumerate(batched_counts(num_scenes, seed))]
