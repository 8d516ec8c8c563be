"""PBFSolver: the reference's Simulation-layer class over the C ABI.

Mirrors AkuaEngine::PBFSolver (PBFSolver.h:12-31): constructed from a particle count, a
``PBFConfig`` and ``LambdaCorrParams``; ``step(deltaTime, boxMin, boxMax)`` advances the particle
buffer by one PBF step (PBFSolver.cpp:22-78); ``setGravity`` replaces the gravity vector.  The
reference's step takes the OpenGL interop resource that holds the particles; here the solver owns
the particle buffer (``set_particles`` / ``get_particles``), or ``step_device`` steps a caller-owned
device buffer of 108-byte records in place.
"""
from __future__ import annotations

import ctypes
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from .config import LambdaCorrParams, PBFConfig, to_c_config, to_c_corr


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f3(v: Sequence[float]) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(3))


class PBFSolver:
    """One scene on one GPU.  Particle arrays are returned in the solver's current order, which
    the spatial sort changes every step (as the reference's buffer is physically reordered);
    ``ids`` maps each slot back to the particle's original index."""

    def __init__(self, numParticles: int, config: PBFConfig, corrParams: LambdaCorrParams):
        self._L = _lib.load()
        self.numParticles = int(numParticles)
        self.config = config
        self.corrParams = corrParams
        self._c_cfg = to_c_config(config)
        self._c_corr = to_c_corr(corrParams)
        h = ctypes.c_void_p()
        _lib.check(self._L.akua_solver_create(self.numParticles, ctypes.byref(self._c_cfg),
                                              ctypes.byref(self._c_corr), ctypes.byref(h)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self._L.akua_solver_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state --------------------------------------------------------------------------------
    def set_state(self, pos: np.ndarray, vel: np.ndarray, mass: Optional[np.ndarray] = None,
                  density: Optional[np.ndarray] = None, lam: Optional[np.ndarray] = None) -> None:
        n = self.numParticles
        pos = np.ascontiguousarray(pos, np.float32).reshape(n, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(n, 3)
        opt = [None if a is None else np.ascontiguousarray(a, np.float32).reshape(n)
               for a in (mass, density, lam)]
        _lib.check(self._L.akua_solver_set_state(self._h, _p(pos), _p(vel), *[_p(a) for a in opt], n))

    def get_state(self) -> Dict[str, np.ndarray]:
        n = self.numParticles
        out = dict(pos=np.zeros((n, 3), np.float32), vel=np.zeros((n, 3), np.float32),
                   mass=np.zeros(n, np.float32), density=np.zeros(n, np.float32),
                   lam=np.zeros(n, np.float32), ids=np.zeros(n, np.uint32))
        _lib.check(self._L.akua_solver_get_state(self._h, _p(out["pos"]), _p(out["vel"]),
                                                 _p(out["mass"]), _p(out["density"]),
                                                 _p(out["lam"]), _p(out["ids"]), n))
        return out

    def set_particles(self, particles) -> None:
        """Uploads an array of ``_lib.AkuaParticle`` records (Renderer::uploadParticles)."""
        _lib.check(self._L.akua_solver_set_particles(self._h, ctypes.cast(particles, ctypes.c_void_p),
                                                     len(particles)))

    def get_particles(self):
        arr = (_lib.AkuaParticle * self.numParticles)()
        _lib.check(self._L.akua_solver_get_particles(self._h, ctypes.cast(arr, ctypes.c_void_p),
                                                     self.numParticles))
        return arr

    # -- stepping -----------------------------------------------------------------------------
    def step(self, deltaTime: float, boxMin: Sequence[float], boxMax: Sequence[float]) -> None:
        _lib.check(self._L.akua_solver_step(self._h, float(deltaTime), _p(_f3(boxMin)), _p(_f3(boxMax))))

    def step_n(self, deltaTime: float, boxMin, boxMax, nsteps: int) -> None:
        _lib.check(self._L.akua_solver_step_n(self._h, float(deltaTime), _p(_f3(boxMin)),
                                              _p(_f3(boxMax)), int(nsteps)))

    def step_device(self, d_particles: int, deltaTime: float, boxMin, boxMax, stream: int = 0) -> None:
        """Steps a device buffer of numParticles 108-byte records (e.g. a torch uint8 tensor's
        data_ptr()), ordered on `stream` (0: the null stream, torch's default stream): the step
        waits for the work already on that stream, and later work on it waits for the step."""
        _lib.check(self._L.akua_solver_step_device(self._h, ctypes.c_void_p(d_particles),
                                                   self.numParticles, float(deltaTime),
                                                   _p(_f3(boxMin)), _p(_f3(boxMax)),
                                                   ctypes.c_void_p(stream) if stream else None))

    def setGravity(self, gravity: Sequence[float]) -> None:
        _lib.check(self._L.akua_solver_set_gravity(self._h, _p(_f3(gravity))))
        self.config.gravity = tuple(float(g) for g in gravity)

    def synchronize(self) -> None:
        _lib.check(self._L.akua_solver_synchronize(self._h))

    def set_graph(self, enable: bool) -> None:
        _lib.check(self._L.akua_solver_set_graph(self._h, 1 if enable else 0))

    # -- tuning -------------------------------------------------------------------------------
    def set_tuning(self, block: int = 256, layoutLevel: int = 1, hashDigitBits: int = 0,
                   layoutDigitBits: int = 0, layoutInterval: int = 1, balanceTile: int = 0,
                   slabBatch: int = 0, fuseSweep: int = 0, laneGroup: int = 0,
                   workQueue: int = 0) -> None:
        t = _lib.AkuaTuning(block, layoutLevel, hashDigitBits, layoutDigitBits, layoutInterval,
                            balanceTile, slabBatch, fuseSweep, laneGroup, workQueue)
        _lib.check(self._L.akua_solver_set_tuning(self._h, ctypes.byref(t)))

    def set_search(self, cellsPerH: int) -> None:
        """1: the reference's neighbour search (bit-exact); 2: cells of h / 2 (results equal to
        FP32 rounding, not bit for bit)."""
        _lib.check(self._L.akua_solver_set_search(self._h, int(cellsPerH)))

    def get_search(self) -> int:
        c = ctypes.c_int32()
        _lib.check(self._L.akua_solver_get_search(self._h, ctypes.byref(c)))
        return int(c.value)

    def set_exact(self, exact: bool) -> None:
        """False lets autotune() also try the h / 2 search (which is not bit-exact)."""
        _lib.check(self._L.akua_solver_set_exact(self._h, 1 if exact else 0))

    def get_tuning(self) -> Dict[str, int]:
        t = _lib.AkuaTuning()
        _lib.check(self._L.akua_solver_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def autotune(self, deltaTime: float, boxMin=None, boxMax=None, trialSteps: int = 5,
                 cachePath: Optional[str] = None) -> Dict[str, int]:
        """Times candidate settings on the current state (restored afterwards), keeps the fastest."""
        t = _lib.AkuaTuning()
        bmin = _f3(boxMin) if boxMin is not None else None
        bmax = _f3(boxMax) if boxMax is not None else None
        path = cachePath.encode() if cachePath else None
        _lib.check(self._L.akua_solver_autotune(self._h, float(deltaTime), _p(bmin), _p(bmax),
                                                int(trialSteps), path, ctypes.byref(t)))
        return t.as_dict()

    # -- headless rendering -------------------------------------------------------------------
    def render(self, camera=None, width: int = 1280, height: int = 720, point_size: float = 50.0,
               floor: bool = True) -> np.ndarray:
        """(height, width, 3) uint8 image of the particles as the reference draws them (point
        sprites over the checkered floor); camera: an akua_camera (default: the reference's)."""
        cam = camera if camera is not None else default_camera()
        img = np.zeros((int(height), int(width), 3), np.uint8)
        _lib.check(self._L.akua_render(self._h, ctypes.byref(cam), int(width), int(height),
                                       float(point_size), 1 if floor else 0, _p(img)))
        return img

    # -- boundary samples on the walls --------------------------------------------------------
    def set_boundary(self, layers: int, spacing: float = 0.0) -> None:
        """Wall samples with Akinci volume weights (layers 0: off; spacing 0: particle_spacing).
        No reference counterpart: the reference plans them (ConstraintSolverCUDA.cu:132-135)."""
        _lib.check(self._L.akua_solver_set_boundary(self._h, int(layers), float(spacing)))

    def debug_boundary(self):
        """(positions (m, 3), psi (m,)) of the samples for the last box, in index order."""
        m = ctypes.c_int32()
        _lib.check(self._L.akua_debug_boundary(self._h, None, None, 0, ctypes.byref(m)))
        pos = np.zeros((m.value, 3), np.float32)
        psi = np.zeros(m.value, np.float32)
        _lib.check(self._L.akua_debug_boundary(self._h, _p(pos), _p(psi), m.value, ctypes.byref(m)))
        return pos, psi

    # -- learned warm start -------------------------------------------------------------------
    def set_network(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32).reshape(_lib.MLP_PARAMS)
        _lib.check(self._L.akua_solver_set_network(self._h, _p(p), _lib.MLP_PARAMS))

    def get_network(self) -> np.ndarray:
        p = np.zeros(_lib.MLP_PARAMS, np.float32)
        _lib.check(self._L.akua_solver_get_network(self._h, _p(p), _lib.MLP_PARAMS))
        return p

    def set_learned(self, enable: bool, ratio: float = 1.0, tolerance: float = 0.01,
                    statistic: str = "max", iterations: int = 1) -> None:
        """Learned steps: warm start + the last `iterations` Jacobi iterations, per-scene fallback
        to the plain solve when r(x*_w) > ratio * r(x*_pred) + tolerance, r the max or mean of
        max(0, C)."""
        st = {"max": 0, "mean": 1}[statistic]
        _lib.check(self._L.akua_solver_set_learned(self._h, 1 if enable else 0, int(iterations), st,
                                                   float(ratio), float(tolerance)))

    def learned_stats(self) -> Dict[str, object]:
        steps, fb = ctypes.c_int64(), ctypes.c_int64()
        r = np.zeros(6, np.float32)
        _lib.check(self._L.akua_solver_learned_stats(self._h, ctypes.byref(steps), ctypes.byref(fb), _p(r)))
        return dict(steps=int(steps.value), fallbacks=int(fb.value),
                    max_pred=float(r[0]), max_warm=float(r[1]), max_plain=float(r[2]),
                    mean_pred=float(r[3]), mean_warm=float(r[4]), mean_plain=float(r[5]))

    def train(self, deltaTime: float, boxMin, boxMax, steps: int, samplesPerStep: int = 65536,
              minibatch: int = 8192, epochs: int = 1, learningRate: float = 1e-3, seed: int = 1,
              lossWeights=(1.0, 1.0, 1.0, 0.25), normSteps: int = 20, targetIterations: int = 1):
        """Online training on this scene's plain trajectory.  Returns (loss per step, residuals
        (steps, 4): max and mean of max(0, C) at x*_pred, then at the last plain density pass)."""
        c = _lib.AkuaTrainConfig()
        c.samplesPerStep, c.minibatch, c.epochs = int(samplesPerStep), int(minibatch), int(epochs)
        c.normSteps = int(normSteps)
        c.targetIterations = int(targetIterations)
        c.learningRate, c.seed = float(learningRate), int(seed)
        for k in range(4):
            c.lossWeights[k] = float(lossWeights[k])
        loss = np.zeros(int(steps), np.float32)
        resid = np.zeros((int(steps), 4), np.float32)
        _lib.check(self._L.akua_solver_train(self._h, float(deltaTime), _p(_f3(boxMin)), _p(_f3(boxMax)),
                                             int(steps), ctypes.byref(c), _p(loss), _p(resid)))
        return loss, resid

    def debug_features(self, targets: bool = False):
        F = np.zeros((self.numParticles, 16), np.float32)
        T = np.zeros((self.numParticles, 4), np.float32) if targets else None
        _lib.check(self._L.akua_debug_features(self._h, _p(F), _p(T), self.numParticles))
        return (F, T) if targets else F

    def debug_warm_start(self):
        n = self.numParticles
        xp, xw, lh = np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros(n, np.float32)
        _lib.check(self._L.akua_debug_warm_start(self._h, _p(xp), _p(xw), _p(lh), n))
        return xp, xw, lh

    # -- parity taps --------------------------------------------------------------------------
    def debug_hashes(self) -> np.ndarray:
        out = np.zeros(self.numParticles, np.uint32)
        _lib.check(self._L.akua_debug_hashes(self._h, _p(out), self.numParticles))
        return out

    def debug_permutation(self) -> np.ndarray:
        out = np.zeros(self.numParticles, np.uint32)
        _lib.check(self._L.akua_debug_permutation(self._h, _p(out), self.numParticles))
        return out

    def debug_physical_order(self) -> np.ndarray:
        """Logical index of the particle in each physical slot (the library's memory layout)."""
        out = np.zeros(self.numParticles, np.uint32)
        _lib.check(self._L.akua_debug_physical_order(self._h, _p(out), self.numParticles))
        return out

    def debug_cell_ranges(self):
        n = ctypes.c_int32(0)
        st = self._L.akua_debug_cell_ranges(self._h, None, None, None, 0, ctypes.byref(n))
        if st not in (_lib.AKUA_OK, 1):
            _lib.check(st)
        cap = n.value
        h = np.zeros(cap, np.uint32)
        s = np.zeros(cap, np.uint32)
        e = np.zeros(cap, np.uint32)
        _lib.check(self._L.akua_debug_cell_ranges(self._h, _p(h), _p(s), _p(e), cap, ctypes.byref(n)))
        return h, s, e

    def debug_neighbours(self):
        m = int(self.config.maxNeighbours)
        nbr = np.zeros((self.numParticles, m), np.uint32)
        cnt = np.zeros(self.numParticles, np.uint32)
        _lib.check(self._L.akua_debug_neighbours(self._h, _p(nbr), _p(cnt), self.numParticles))
        return nbr, cnt

    # -- measurement --------------------------------------------------------------------------
    def set_timing(self, enable: bool) -> None:
        _lib.check(self._L.akua_solver_set_timing(self._h, 1 if enable else 0))

    def reset_timing(self) -> None:
        _lib.check(self._L.akua_solver_reset_timing(self._h))

    def get_timing(self) -> Dict[str, tuple]:
        cap, nl = 64, 64
        names = ctypes.create_string_buffer(cap * nl)
        ms = np.zeros(cap, np.float64)
        launches = np.zeros(cap, np.int64)
        count = ctypes.c_int32(0)
        _lib.check(self._L.akua_solver_get_timing(self._h, names, nl, _p(ms), _p(launches), cap,
                                                  ctypes.byref(count)))
        out = {}
        for i in range(min(cap, count.value)):
            nm = names.raw[i * nl:(i + 1) * nl].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(launches[i]))
        return out

    def density_error(self):
        """(max, mean) over particles of max(0, rho / restDensity - 1), rho the exact SPH density
        at the current positions (computed on the device)."""
        m, a = ctypes.c_float(0.0), ctypes.c_double(0.0)
        _lib.check(self._L.akua_solver_density_error(self._h, ctypes.byref(m), ctypes.byref(a)))
        return float(m.value), float(a.value)

    def debug_probe_density(self) -> np.ndarray:
        out = np.zeros(self.numParticles, np.float32)
        _lib.check(self._L.akua_debug_probe_density(self._h, _p(out), self.numParticles))
        return out

    def mean_neighbours(self) -> float:
        v = ctypes.c_double(0.0)
        _lib.check(self._L.akua_solver_mean_neighbours(self._h, ctypes.byref(v)))
        return float(v.value)


def default_camera():
    """The reference's camera (CameraConfig.h:8-16) and clip planes (Application.cpp:17-18)."""
    c = _lib.AkuaCamera()
    _lib.load().akua_default_camera(ctypes.byref(c))
    return c


def camera_matrix(camera, width: int, height: int) -> np.ndarray:
    """The renderer's row-major view-projection matrix (4, 4) for this image size."""
    m = np.zeros(16, np.float32)
    _lib.check(_lib.load().akua_camera_matrix(ctypes.byref(camera), int(width), int(height), _p(m)))
    return m.reshape(4, 4)


def boundary_samples(config: PBFConfig, box_min, box_max, layers: int, spacing: float = 0.0):
    """(positions (m, 3), psi (m,)) of the wall samples a solver with this configuration builds for
    the box: host only, no GPU (akua_boundary_samples)."""
    L = _lib.load()
    c = to_c_config(config)
    lo, hi = _f3(box_min), _f3(box_max)
    m = ctypes.c_int32()
    _lib.check(L.akua_boundary_samples(ctypes.byref(c), _p(lo), _p(hi), int(layers), float(spacing),
                                       None, None, 0, ctypes.byref(m)))
    pos = np.zeros((m.value, 3), np.float32)
    psi = np.zeros(m.value, np.float32)
    _lib.check(L.akua_boundary_samples(ctypes.byref(c), _p(lo), _p(hi), int(layers), float(spacing),
                                       _p(pos), _p(psi), m.value, ctypes.byref(m)))
    return pos, psi


class BatchSolver(PBFSolver):
    """Many independent scenes stepped together: one kernel launch per stage covers every scene
    and each step is one hipGraph.  Particles are passed scene after scene (concatenated)."""

    def __init__(self, counts: Sequence[int], config: PBFConfig, corrParams: LambdaCorrParams):
        self._L = _lib.load()
        self.counts = [int(c) for c in counts]
        self.numParticles = int(sum(self.counts))
        self.config = config
        self.corrParams = corrParams
        self._c_cfg = to_c_config(config)
        self._c_corr = to_c_corr(corrParams)
        arr = np.ascontiguousarray(self.counts, np.int32)
        h = ctypes.c_void_p()
        _lib.check(self._L.akua_batch_create(len(self.counts), _p(arr), ctypes.byref(self._c_cfg),
                                             ctypes.byref(self._c_corr), ctypes.byref(h)))
        self._h = h

    @property
    def num_scenes(self) -> int:
        return len(self.counts)

    def set_box(self, scene: int, boxMin, boxMax) -> None:
        _lib.check(self._L.akua_batch_set_box(self._h, int(scene), _p(_f3(boxMin)), _p(_f3(boxMax))))

    def step_batch(self, deltaTime: float, nsteps: int = 1) -> None:
        _lib.check(self._L.akua_batch_step(self._h, float(deltaTime), int(nsteps)))

    def density_errors(self):
        """Per-scene (max, mean) density-constraint error at the current positions."""
        mx = np.zeros(self.num_scenes, np.float32)
        mean = np.zeros(self.num_scenes, np.float64)
        _lib.check(self._L.akua_batch_density_error(self._h, _p(mx), _p(mean)))
        return mx, mean
