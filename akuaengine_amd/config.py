"""Solver parameters, mirroring the reference's Simulation-layer structs.

``PBFConfig`` follows AkuaEngine::PBFConfig (PBFConfig.h:18-29) and ``LambdaCorrParams`` follows
AkuaEngine::LambdaCorrParams (PBFConfig.h:10-15): same field names, same defaults.  The C ABI
carries them as ``akua_pbf_config`` / ``akua_lambda_corr`` (include/akua.h) with the same field
order.

``dam_break_params()`` returns the parameter set of the dam-break workloads (rest density 6378,
artificial pressure k = 0.1, n = 4, delta_q = 0.2 h, dt = 0.0083), which override several of the
reference's defaults.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field, replace
from typing import Tuple


@dataclass
class LambdaCorrParams:
    """Artificial-pressure (s_corr) parameters; PBFConfig.h:10-15.

    ``enabled = False`` drops the term.  The reference never reads the flag
    (ConstraintSolverCUDA.cu:99-130 always applies it), so results are the reference's whenever
    it is set.
    """

    enabled: bool = True
    k: float = 0.0001
    n: float = 4.0
    delta_q: float = 0.03


@dataclass
class PBFConfig:
    """PBF solver parameters; PBFConfig.h:18-29."""

    restDensity: float = 7600.0
    particle_spacing: float = 0.05
    smoothRadius: float = 0.1
    spatialHashCellSize: float = 0.1
    relaxation: float = 600.0
    vorticityEpsilon: float = 0.00001
    viscosity: float = 0.01
    maxNeighbours: int = 128
    solverIterations: int = 4
    gravity: Tuple[float, float, float] = field(default_factory=lambda: (0.0, -9.8, 0.0))


class CPBFConfig(ctypes.Structure):
    """``akua_pbf_config`` (include/akua.h)."""

    _fields_ = [
        ("restDensity", ctypes.c_float),
        ("particle_spacing", ctypes.c_float),
        ("smoothRadius", ctypes.c_float),
        ("spatialHashCellSize", ctypes.c_float),
        ("relaxation", ctypes.c_float),
        ("vorticityEpsilon", ctypes.c_float),
        ("viscosity", ctypes.c_float),
        ("maxNeighbours", ctypes.c_int32),
        ("solverIterations", ctypes.c_int32),
        ("gravity", ctypes.c_float * 3),
    ]


class CLambdaCorr(ctypes.Structure):
    """``akua_lambda_corr`` (include/akua.h)."""

    _fields_ = [
        ("enabled", ctypes.c_int32),
        ("k", ctypes.c_float),
        ("n", ctypes.c_float),
        ("delta_q", ctypes.c_float),
    ]


def to_c_config(cfg: PBFConfig) -> CPBFConfig:
    c = CPBFConfig()
    c.restDensity = cfg.restDensity
    c.particle_spacing = cfg.particle_spacing
    c.smoothRadius = cfg.smoothRadius
    c.spatialHashCellSize = cfg.spatialHashCellSize
    c.relaxation = cfg.relaxation
    c.vorticityEpsilon = cfg.vorticityEpsilon
    c.viscosity = cfg.viscosity
    c.maxNeighbours = int(cfg.maxNeighbours)
    c.solverIterations = int(cfg.solverIterations)
    for k in range(3):
        c.gravity[k] = float(cfg.gravity[k])
    return c


def to_c_corr(corr: LambdaCorrParams) -> CLambdaCorr:
    c = CLambdaCorr()
    c.enabled = 1 if corr.enabled else 0
    c.k = corr.k
    c.n = corr.n
    c.delta_q = corr.delta_q
    return c


DAM_BREAK_DT = 0.0083


def dam_break_params(iterations: int = 4) -> Tuple[PBFConfig, LambdaCorrParams]:
    """Parameters shared by every dam-break workload (27k, 1M, 8M, batched, learned)."""
    cfg = PBFConfig(restDensity=6378.0, smoothRadius=0.1, spatialHashCellSize=0.1,
                    relaxation=600.0, viscosity=0.01, vorticityEpsilon=0.00001,
                    maxNeighbours=128, solverIterations=iterations, gravity=(0.0, -9.8, 0.0))
    corr = LambdaCorrParams(enabled=True, k=0.1, n=4.0, delta_q=0.2 * 0.1)
    return cfg, corr


def with_iterations(cfg: PBFConfig, iterations: int) -> PBFConfig:
    return replace(cfg, solverIterations=iterations)
