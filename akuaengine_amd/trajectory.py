"""Trajectory files: a versioned binary structure-of-arrays record of particle frames.

The format (version 1) is specified in include/akua_io.h.  ``read_trajectory`` maps a file with
numpy alone (no library, no GPU), so fixtures and training data load anywhere.  ``TrajectoryWriter``
and ``TrajectoryReader`` go through libakua.so's own writer and reader (akua_traj_*), the code the
headless runner ``akua_run`` uses.

Frames hold the solver's logical order (the reference's hash-sorted order); ``ids`` maps each slot
to the particle's original index, and ``Frames.by_id`` reorders a field back to that index.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass
from typing import Dict, Optional, Sequence

import numpy as np

MAGIC = b"AKUATRJ\0"
VERSION = 1
HEADER_BYTES = 128

POSITION, VELOCITY, MASS, DENSITY, LAMBDA, ID = 1, 2, 4, 8, 16, 32
ALL = 63
# (bit, name, dtype, components) in file order
FIELDS = [(POSITION, "pos", "<f4", 3), (VELOCITY, "vel", "<f4", 3), (MASS, "mass", "<f4", 1),
          (DENSITY, "density", "<f4", 1), (LAMBDA, "lam", "<f4", 1), (ID, "ids", "<u4", 1)]

_HEADER = np.dtype([("magic", "S8"), ("version", "<u4"), ("header_bytes", "<u4"), ("n", "<i4"),
                    ("fields", "<u4"), ("dt", "<f4"), ("box_min", "<f4", 3), ("box_max", "<f4", 3),
                    ("gravity", "<f4", 3), ("frame_bytes", "<u8"), ("zero", "V56")])
assert _HEADER.itemsize == HEADER_BYTES


class TrajectoryFormatError(ValueError):
    pass


@dataclass
class TrajInfo:
    version: int
    n: int
    fields: int
    dt: float
    box_min: tuple
    box_max: tuple
    gravity: tuple
    frame_bytes: int
    frames: int
    partial_frame: bool


def frame_dtype(n: int, fields: int) -> np.dtype:
    """The numpy record of one frame: step, time, then each present field."""
    cols = [("step", "<i8"), ("time", "<f8")]
    for bit, name, dt, comp in FIELDS:
        if fields & bit:
            cols.append((name, dt, (n, comp) if comp > 1 else (n,)))
    return np.dtype(cols)


def read_info(path: str) -> TrajInfo:
    with open(path, "rb") as f:
        raw = f.read(HEADER_BYTES)
        if len(raw) < HEADER_BYTES:
            raise TrajectoryFormatError(f"{path}: shorter than a trajectory header")
        size = os.fstat(f.fileno()).st_size
    h = np.frombuffer(raw, _HEADER)[0]
    if bytes(h["magic"]).ljust(8, b"\0") != MAGIC:
        raise TrajectoryFormatError(f"{path}: not a trajectory file (bad magic)")
    if int(h["version"]) != VERSION:
        raise TrajectoryFormatError(f"{path}: trajectory version {int(h['version'])}, this reader reads {VERSION}")
    n, fields = int(h["n"]), int(h["fields"])
    if (int(h["header_bytes"]) != HEADER_BYTES or n < 1 or fields == 0 or fields & ~ALL
            or int(h["frame_bytes"]) != frame_dtype(n, fields).itemsize):
        raise TrajectoryFormatError(f"{path}: inconsistent trajectory header")
    fb = int(h["frame_bytes"])
    body = size - HEADER_BYTES
    return TrajInfo(VERSION, n, fields, float(h["dt"]), tuple(float(x) for x in h["box_min"]),
                    tuple(float(x) for x in h["box_max"]), tuple(float(x) for x in h["gravity"]), fb,
                    body // fb, body % fb != 0)


class Frames:
    """The whole frames of a trajectory file, memory-mapped (read only)."""

    def __init__(self, path: str):
        self.info = read_info(path)
        i = self.info
        self.records = np.memmap(path, dtype=frame_dtype(i.n, i.fields), mode="r", offset=HEADER_BYTES,
                                 shape=(i.frames,)) if i.frames else np.zeros(0, frame_dtype(i.n, i.fields))

    def __len__(self) -> int:
        return self.info.frames

    def __getitem__(self, k: int) -> Dict[str, np.ndarray]:
        r = self.records[k]
        out = {"step": int(r["step"]), "time": float(r["time"])}
        for bit, name, _, _ in FIELDS:
            if self.info.fields & bit:
                out[name] = np.array(r[name])
        return out

    def by_id(self, k: int, name: str) -> np.ndarray:
        """Field `name` of frame k indexed by original particle index."""
        if not self.info.fields & ID:
            raise TrajectoryFormatError("the file holds no ids")
        r = self.records[k]
        v = np.array(r[name])
        out = np.empty_like(v)
        out[np.asarray(r["ids"], dtype=np.int64)] = v
        return out


def read_trajectory(path: str) -> Frames:
    return Frames(path)


def _field_bits(names: Optional[Sequence[str]]) -> int:
    if names is None:
        return POSITION | VELOCITY | DENSITY | ID
    table = {name: bit for bit, name, _, _ in FIELDS}
    table.update({"position": POSITION, "velocity": VELOCITY, "lambda": LAMBDA, "id": ID})
    bits = 0
    for nm in names:
        if nm == "all":
            return ALL
        bits |= table[nm]
    return bits


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class TrajectoryWriter:
    """akua_traj_create / akua_traj_append / akua_traj_close."""

    def __init__(self, path: str, n: int, fields: Optional[Sequence[str]] = None, dt: float = 0.0,
                 box_min=(0.0, 0.0, 0.0), box_max=(0.0, 0.0, 0.0), gravity=(0.0, -9.8, 0.0)):
        from . import _lib
        self._L = _lib.load()
        self.n = int(n)
        self.fields = _field_bits(fields)
        h = ctypes.c_void_p()
        f3 = lambda v: np.asarray(v, np.float32).reshape(3)
        self._boxes = (f3(box_min), f3(box_max), f3(gravity))
        _lib.check(self._L.akua_traj_create(path.encode(), self.n, self.fields, float(dt),
                                            *(_ptr(b) for b in self._boxes), ctypes.byref(h)))
        self._h = h

    def append(self, step: int, time: float, pos=None, vel=None, mass=None, density=None, lam=None,
               ids=None) -> None:
        from . import _lib
        n = self.n
        arrs = [None if pos is None else np.ascontiguousarray(pos, np.float32).reshape(n, 3),
                None if vel is None else np.ascontiguousarray(vel, np.float32).reshape(n, 3),
                None if mass is None else np.ascontiguousarray(mass, np.float32).reshape(n),
                None if density is None else np.ascontiguousarray(density, np.float32).reshape(n),
                None if lam is None else np.ascontiguousarray(lam, np.float32).reshape(n),
                None if ids is None else np.ascontiguousarray(ids, np.uint32).reshape(n)]
        _lib.check(self._L.akua_traj_append(self._h, int(step), float(time), *(_ptr(a) for a in arrs)))

    def append_solver(self, solver, step: int, time: float, id_map=None) -> None:
        """The solver's current state (synchronises it)."""
        from . import _lib
        m = None if id_map is None else np.ascontiguousarray(id_map, np.uint32)
        _lib.check(self._L.akua_traj_append_solver(self._h, solver._h, int(step), float(time), _ptr(m)))

    def close(self) -> None:
        from . import _lib
        if self._h:
            h, self._h = self._h, None
            _lib.check(self._L.akua_traj_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.akua_traj_close(self._h)


class TrajectoryReader:
    """akua_traj_open / akua_traj_read / akua_traj_close."""

    def __init__(self, path: str):
        from . import _lib
        self._L = _lib.load()
        info = _lib.AkuaTrajInfo()
        h = ctypes.c_void_p()
        _lib.check(self._L.akua_traj_open(path.encode(), ctypes.byref(info), ctypes.byref(h)))
        self._h = h
        self.n = int(info.numParticles)
        self.fields = int(info.fields)
        self.frames = int(info.frames)
        self.partial_frame = bool(info.partialFrame)
        self.dt = float(info.deltaTime)
        self.box_min = tuple(info.boxMin)
        self.box_max = tuple(info.boxMax)
        self.gravity = tuple(info.gravity)

    def read(self, frame: int) -> Dict[str, np.ndarray]:
        from . import _lib
        n = self.n
        out = {}
        arrs = []
        for bit, name, dt, comp in FIELDS:
            a = np.empty((n, comp) if comp > 1 else (n,), dtype=np.dtype(dt)) if self.fields & bit else None
            arrs.append(a)
            if a is not None:
                out[name] = a
        step, time = ctypes.c_int64(), ctypes.c_double()
        _lib.check(self._L.akua_traj_read(self._h, int(frame), ctypes.byref(step), ctypes.byref(time),
                                          *(_ptr(a) for a in arrs)))
        out["step"], out["time"] = int(step.value), float(time.value)
        return out

    def close(self) -> None:
        from . import _lib
        if self._h:
            h, self._h = self._h, None
            _lib.check(self._L.akua_traj_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.akua_traj_close(self._h)


def native_scene(name: str, spacing: float = 0.0, seed: int = -1):
    """(positions, akua_scene_info) of a scene generated by libakua.so (akua_scene_generate)."""
    from . import _lib
    L = _lib.load()
    info = _lib.AkuaSceneInfo()
    _lib.check(L.akua_scene_describe(name.encode(), float(spacing), int(seed), ctypes.byref(info)))
    pos = np.empty((info.numParticles, 3), np.float32)
    _lib.check(L.akua_scene_generate(name.encode(), float(spacing), int(seed), _ptr(pos), info.numParticles))
    return pos, info


RUNNER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "akua_run")
