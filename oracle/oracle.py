"""ctypes wrapper over oracle/liboracle.so -- the CPU restatement of the reference PBF step.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, where it is the checker or the timed CPU baseline; the product package never
imports it.  ``prec=32`` is the FP32 parity oracle, ``prec=64`` the FP64 error reference.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_PLANT_PATH = os.path.join(_HERE, "liboracle_plant.so")
_libs: Dict[bool, ctypes.CDLL] = {}


class _Cfg(ctypes.Structure):
    _fields_ = [("restDensity", ctypes.c_float), ("particle_spacing", ctypes.c_float),
                ("smoothRadius", ctypes.c_float), ("spatialHashCellSize", ctypes.c_float),
                ("relaxation", ctypes.c_float), ("vorticityEpsilon", ctypes.c_float),
                ("viscosity", ctypes.c_float), ("maxNeighbours", ctypes.c_int32),
                ("solverIterations", ctypes.c_int32), ("gravity", ctypes.c_float * 3)]


class _Corr(ctypes.Structure):
    _fields_ = [("enabled", ctypes.c_int32), ("k", ctypes.c_float), ("n", ctypes.c_float),
                ("delta_q", ctypes.c_float)]


def build(force: bool = False) -> str:
    """Compile liboracle.so (and liboracle_plant.so) with the committed Makefile."""
    if force or not (os.path.exists(_LIB_PATH) and os.path.exists(_PLANT_PATH)):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB_PATH


def lib(plants: bool = False):
    """The oracle library; plants=True: its twin with the planted-mistake switch (set_plant)."""
    if plants not in _libs:
        path = _PLANT_PATH if plants else _LIB_PATH
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        vp, f, d, u32, i32, lng = (ctypes.c_void_p, ctypes.c_float, ctypes.c_double,
                                   ctypes.c_uint32, ctypes.c_int32, ctypes.c_long)
        P = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.c_int, lng, ctypes.POINTER(_Cfg), ctypes.POINTER(_Corr), u32]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_gravity.argtypes = [vp, P]
        L.orc_set_state.argtypes = [vp, lng, P, P, P, P, P, P]
        L.orc_get_state.argtypes = [vp, P, P, P, P, P, P]
        L.orc_get_stage.argtypes = [vp, P, P, P, P]
        L.orc_step.argtypes = [vp, f, P, P, ctypes.c_int]
        L.orc_stage.argtypes = [vp, ctypes.c_int, f, P, P]
        L.orc_ncells.restype = lng
        L.orc_ncells.argtypes = [vp]
        L.orc_get_taps.argtypes = [vp, P, P, P, P, P, P, P]
        L.orc_true_density.argtypes = [vp, P, P]
        L.orc_get_features.argtypes = [vp, P]
        L.orc_set_xs.argtypes = [vp, P]
        L.orc_poly6.restype = d
        L.orc_poly6.argtypes = [ctypes.c_int, d, d]
        L.orc_grad_spiky.argtypes = [ctypes.c_int, P, d, P]
        L.orc_cell_hash.restype = u32
        L.orc_cell_hash.argtypes = [i32, i32, i32, u32]
        L.orc_set_boundary.argtypes = [vp, ctypes.c_int, f]
        L.orc_boundary_samples.restype = lng
        L.orc_boundary_samples.argtypes = [vp, P, P, P, P, lng]
        L.orc_boundary_sums.argtypes = [vp, P, P, P, P]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_set_vort.argtypes = [vp, P]
        if plants:
            L.orc_set_plant.argtypes = [ctypes.c_int]
        _libs[plants] = L
    return _libs[plants]


def set_plant(k: int) -> None:
    """Turns planted mistake k on in the planted-mistake library (0: none)."""
    lib(True).orc_set_plant(int(k))


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f3(v: Sequence[float]) -> np.ndarray:
    return np.asarray(v, dtype=np.float32).reshape(3).copy()


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def max_threads() -> int:
    return int(lib().orc_max_threads())


def poly6(d2: float, h: float, prec: int = 32) -> float:
    return float(lib().orc_poly6(prec, d2, h))


def grad_spiky(r: Sequence[float], h: float, prec: int = 32) -> np.ndarray:
    rr = np.asarray(r, dtype=np.float64).reshape(3).copy()
    out = np.zeros(3, np.float64)
    lib().orc_grad_spiky(prec, _p(rr), h, _p(out))
    return out


def cell_hash(cx: int, cy: int, cz: int, table_size: int) -> int:
    return int(lib().orc_cell_hash(cx, cy, cz, table_size))


STAGES = {"predict": 0, "hashes": 1, "sort": 2, "neighbours": 3, "densities": 4, "lambdas": 5,
          "deltas": 6, "correct": 7, "update": 8, "damping": 9, "vorticity": 10,
          "confinement": 11, "xsph": 12, "features": 13}


class Oracle:
    """One scene stepped by the CPU oracle; state is kept in the reference's sorted order."""

    def __init__(self, n: int, cfg, corr, prec: int = 32, table_size: int = 0,
                 plants: bool = False):
        c = _Cfg()
        for name, _ in _Cfg._fields_:
            if name == "gravity":
                for k in range(3):
                    c.gravity[k] = float(cfg.gravity[k])
            else:
                setattr(c, name, getattr(cfg, name))
        k = _Corr()
        k.enabled = 1 if corr.enabled else 0
        k.k, k.n, k.delta_q = corr.k, corr.n, corr.delta_q
        self.n = int(n)
        self.prec = prec
        self.cfg = cfg
        self.max_neighbours = int(cfg.maxNeighbours)
        self._L = lib(plants)
        self._h = self._L.orc_create(prec, self.n, ctypes.byref(c), ctypes.byref(k), int(table_size))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._L.orc_destroy(h)
            self._h = None

    def set_gravity(self, g):
        self._L.orc_set_gravity(self._h, _p(_f3(g)))

    def set_state(self, pos, vel, mass=None, density=None, lam=None, ids=None):
        pos = np.ascontiguousarray(pos, np.float32).reshape(self.n, 3)
        vel = np.ascontiguousarray(vel, np.float32).reshape(self.n, 3)
        m = None if mass is None else np.ascontiguousarray(mass, np.float32)
        d = None if density is None else np.ascontiguousarray(density, np.float32)
        l = None if lam is None else np.ascontiguousarray(lam, np.float32)
        i = None if ids is None else np.ascontiguousarray(ids, np.uint32)
        self._L.orc_set_state(self._h, self.n, _p(pos), _p(vel), _p(m), _p(d), _p(l), _p(i))

    def step(self, dt, box_min, box_max, nsteps: int = 1):
        self._L.orc_step(self._h, float(dt), _p(_f3(box_min)), _p(_f3(box_max)), int(nsteps))

    def stage(self, name: str, dt=0.0, box_min=(0, 0, 0), box_max=(1, 1, 1)):
        self._L.orc_stage(self._h, STAGES[name], float(dt), _p(_f3(box_min)), _p(_f3(box_max)))

    def state(self) -> Dict[str, np.ndarray]:
        n = self.n
        out = dict(pos=np.zeros((n, 3)), vel=np.zeros((n, 3)), density=np.zeros(n),
                   lam=np.zeros(n), ids=np.zeros(n, np.uint32), mass=np.zeros(n))
        self._L.orc_get_state(self._h, _p(out["pos"]), _p(out["vel"]), _p(out["density"]),
                            _p(out["lam"]), _p(out["ids"]), _p(out["mass"]))
        return out

    def stage_arrays(self) -> Dict[str, np.ndarray]:
        n = self.n
        out = dict(xs=np.zeros((n, 3)), vs=np.zeros((n, 3)), dp=np.zeros((n, 3)),
                   vort=np.zeros((n, 3)))
        self._L.orc_get_stage(self._h, _p(out["xs"]), _p(out["vs"]), _p(out["dp"]), _p(out["vort"]))
        return out

    def taps(self) -> Dict[str, np.ndarray]:
        n = self.n
        nc = int(self._L.orc_ncells(self._h))
        out = dict(sorted_hash=np.zeros(n, np.uint32), perm=np.zeros(n, np.uint32),
                   nbr=np.zeros(n * self.max_neighbours, np.uint32), cnt=np.zeros(n, np.uint32),
                   cell_hash=np.zeros(nc, np.uint32), cell_start=np.zeros(nc, np.uint32),
                   cell_end=np.zeros(nc, np.uint32))
        self._L.orc_get_taps(self._h, *[_p(out[k]) for k in ("sorted_hash", "perm", "nbr", "cnt",
                                                           "cell_hash", "cell_start", "cell_end")])
        out["nbr"] = out["nbr"].reshape(n, self.max_neighbours)
        return out

    def features(self) -> np.ndarray:
        """Features of the last "features" stage: (n, 16), sorted order."""
        out = np.zeros((self.n, 16))
        self._L.orc_get_features(self._h, _p(out))
        return out

    def set_xs(self, xs) -> None:
        """Replaces x* (sorted order): continues a step from a warm start."""
        a = np.ascontiguousarray(xs, np.float32).reshape(self.n, 3)
        self._L.orc_set_xs(self._h, _p(a))

    def set_vort(self, w) -> None:
        """Replaces the vorticity (sorted order): a confinement pass from a given field."""
        a = np.ascontiguousarray(w, np.float32).reshape(self.n, 3)
        self._L.orc_set_vort(self._h, _p(a))

    def set_boundary(self, layers: int, spacing: float = 0.0) -> None:
        """Boundary samples on the box walls (layers 0: off; spacing <= 0: particle_spacing)."""
        self._L.orc_set_boundary(self._h, int(layers), float(spacing))

    def boundary_samples(self, box_min, box_max):
        """(positions (m, 3), psi (m,)) of the samples for this box, in index order."""
        lo, hi = _f3(box_min), _f3(box_max)
        m = self._L.orc_boundary_samples(self._h, _p(lo), _p(hi), None, None, 0)
        if m < 0:
            raise ValueError("no boundary for this box and spacing")
        pos, psi = np.zeros((m, 3)), np.zeros(m)
        self._L.orc_boundary_samples(self._h, _p(lo), _p(hi), _p(pos), _p(psi), m)
        return pos, psi

    def boundary_sums(self, box_min, box_max):
        """(rhoB, gB) at x* of the current state, sorted order."""
        rb, gb = np.zeros(self.n), np.zeros((self.n, 3))
        self._L.orc_boundary_sums(self._h, _p(_f3(box_min)), _p(_f3(box_max)), _p(rb), _p(gb))
        return rb, gb

    def true_density(self):
        """Exact SPH density at the current positions: (by sorted slot, by particle id)."""
        a, b = np.zeros(self.n), np.zeros(self.n)
        self._L.orc_true_density(self._h, _p(a), _p(b))
        return a, b

    def density_error(self, rest_density: float):
        """(max, mean) over particles of max(0, rho / rest_density - 1) at the current positions."""
        _, rho = self.true_density()
        e = np.maximum(0.0, rho / rest_density - 1.0)
        return float(e.max()), float(e.mean())
