"""ctypes binding of libakua.so (include/akua.h).

The library is loaded from this package directory only.  If it is missing the import of any
solver class raises: there is no CPU fallback for the product path.
"""
from __future__ import annotations

import ctypes
import os

from .config import CLambdaCorr, CPBFConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libakua.so")

AKUA_OK = 0
STATUS_NAMES = {
    0: "AKUA_OK", 1: "AKUA_ERR_INVALID_ARG", 2: "AKUA_ERR_OUT_OF_MEMORY", 3: "AKUA_ERR_HIP",
    4: "AKUA_ERR_NON_FINITE", 5: "AKUA_ERR_STATE", 6: "AKUA_ERR_UNSUPPORTED", 7: "AKUA_ERR_NO_DEVICE",
}

# Every symbol include/akua.h declares (checked by tests/test_abi.py).
EXPORTED = [
    "akua_abi_version", "akua_last_error", "akua_default_config",
    "akua_solver_create", "akua_solver_destroy", "akua_solver_set_particles",
    "akua_solver_get_particles", "akua_solver_set_state", "akua_solver_get_state",
    "akua_solver_step", "akua_solver_step_n", "akua_solver_step_device",
    "akua_solver_set_gravity", "akua_solver_synchronize", "akua_solver_num_particles",
    "akua_solver_set_graph", "akua_debug_hashes", "akua_debug_permutation",
    "akua_debug_physical_order",
    "akua_debug_cell_ranges", "akua_debug_neighbours", "akua_solver_set_timing",
    "akua_solver_reset_timing", "akua_solver_get_timing", "akua_solver_mean_neighbours",
    "akua_solver_density_error", "akua_debug_probe_density", "akua_stream_copy_bandwidth",
    "akua_selftest_arith", "akua_batch_create", "akua_batch_num_scenes", "akua_batch_set_box",
    "akua_batch_step", "akua_batch_density_error", "akua_solver_set_tuning",
    "akua_solver_get_tuning", "akua_solver_autotune", "akua_solver_set_network",
    "akua_solver_set_search", "akua_solver_get_search", "akua_solver_set_exact",
    "akua_solver_get_network", "akua_solver_set_learned", "akua_solver_learned_stats",
    "akua_solver_train", "akua_debug_features", "akua_debug_warm_start", "akua_mlp_forward",
    "akua_mlp_gradient", "akua_mlp_adam_step", "akua_mlp_adam_step_device",
    "akua_solver_set_boundary", "akua_debug_boundary",
    "akua_boundary_samples",
]

# Every symbol include/akua_io.h declares (scenes and trajectory files; no GPU needed except by
# akua_traj_append_solver).
EXPORTED_IO = [
    "akua_scene_describe", "akua_scene_generate", "akua_traj_create", "akua_traj_append",
    "akua_traj_append_solver", "akua_traj_open", "akua_traj_read", "akua_traj_close",
    "akua_default_camera", "akua_camera_matrix", "akua_render",
]

MLP_PARAMS = 5544
MLP_TRAINABLE = 5508


class AkuaError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class AkuaParticle(ctypes.Structure):
    """``akua_particle``: the reference's 108-byte record (Particle.h:8-31)."""

    _fields_ = [
        ("position", ctypes.c_float * 3), ("velocity", ctypes.c_float * 3),
        ("new_position", ctypes.c_float * 3), ("new_velocity", ctypes.c_float * 3),
        ("position_delta", ctypes.c_float * 3), ("vorticity", ctypes.c_float * 3),
        ("mass", ctypes.c_float), ("density", ctypes.c_float), ("lambda_", ctypes.c_float),
        ("hash", ctypes.c_uint32), ("color", ctypes.c_float * 4), ("size", ctypes.c_float),
    ]


assert ctypes.sizeof(AkuaParticle) == 108


class AkuaTuning(ctypes.Structure):
    """``akua_tuning``: launch and layout choices that never change a result."""

    _fields_ = [("block", ctypes.c_int32), ("layoutLevel", ctypes.c_int32),
                ("hashDigitBits", ctypes.c_int32), ("layoutDigitBits", ctypes.c_int32),
                ("layoutInterval", ctypes.c_int32), ("balanceTile", ctypes.c_int32),
                ("slabBatch", ctypes.c_int32), ("fuseSweep", ctypes.c_int32),
                ("laneGroup", ctypes.c_int32), ("workQueue", ctypes.c_int32)]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class AkuaTrainConfig(ctypes.Structure):
    """``akua_train_config``: online training of the warm-start network."""

    _fields_ = [("samplesPerStep", ctypes.c_int32), ("minibatch", ctypes.c_int32),
                ("epochs", ctypes.c_int32), ("normSteps", ctypes.c_int32),
                ("targetIterations", ctypes.c_int32),
                ("learningRate", ctypes.c_float),
                ("seed", ctypes.c_uint64), ("lossWeights", ctypes.c_float * 4)]

class AkuaSceneInfo(ctypes.Structure):
    """``akua_scene_info`` (include/akua_io.h)."""

    _fields_ = [("numParticles", ctypes.c_int32), ("boxMin", ctypes.c_float * 3),
                ("boxMax", ctypes.c_float * 3), ("deltaTime", ctypes.c_float),
                ("referenceParams", ctypes.c_int32)]


class AkuaCamera(ctypes.Structure):
    """``akua_camera`` (include/akua_io.h): the reference's CameraConfig and clip planes."""

    _fields_ = [("position", ctypes.c_float * 3), ("yaw", ctypes.c_float), ("pitch", ctypes.c_float),
                ("fovY", ctypes.c_float), ("nearZ", ctypes.c_float), ("farZ", ctypes.c_float)]


class AkuaTrajInfo(ctypes.Structure):
    """``akua_traj_info`` (include/akua_io.h)."""

    _fields_ = [("version", ctypes.c_uint32), ("numParticles", ctypes.c_int32),
                ("fields", ctypes.c_uint32), ("deltaTime", ctypes.c_float),
                ("boxMin", ctypes.c_float * 3), ("boxMax", ctypes.c_float * 3),
                ("gravity", ctypes.c_float * 3), ("frameBytes", ctypes.c_uint64),
                ("frames", ctypes.c_int64), ("partialFrame", ctypes.c_int32)]


_lib = None


def load() -> ctypes.CDLL:
    """Loads libakua.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m akuaengine_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    i32, f32 = ctypes.c_int32, ctypes.c_float
    L.akua_abi_version.restype = i32
    L.akua_last_error.restype = ctypes.c_char_p
    L.akua_default_config.argtypes = [ctypes.POINTER(CPBFConfig), ctypes.POINTER(CLambdaCorr)]
    sigs = {
        "akua_solver_create": [i32, ctypes.POINTER(CPBFConfig), ctypes.POINTER(CLambdaCorr), ctypes.POINTER(P)],
        "akua_solver_destroy": [P],
        "akua_solver_set_particles": [P, P, i32],
        "akua_solver_get_particles": [P, P, i32],
        "akua_solver_set_state": [P, P, P, P, P, P, i32],
        "akua_solver_get_state": [P, P, P, P, P, P, P, i32],
        "akua_solver_step": [P, f32, P, P],
        "akua_solver_step_n": [P, f32, P, P, i32],
        "akua_solver_step_device": [P, P, i32, f32, P, P, P],
        "akua_solver_set_gravity": [P, P],
        "akua_solver_synchronize": [P],
        "akua_solver_num_particles": [P, ctypes.POINTER(i32)],
        "akua_solver_set_graph": [P, i32],
        "akua_debug_hashes": [P, P, i32],
        "akua_debug_permutation": [P, P, i32],
        "akua_debug_physical_order": [P, P, i32],
        "akua_debug_cell_ranges": [P, P, P, P, i32, ctypes.POINTER(i32)],
        "akua_debug_neighbours": [P, P, P, i32],
        "akua_solver_set_timing": [P, i32],
        "akua_solver_reset_timing": [P],
        "akua_solver_get_timing": [P, P, i32, P, P, i32, ctypes.POINTER(i32)],
        "akua_solver_mean_neighbours": [P, ctypes.POINTER(ctypes.c_double)],
        "akua_solver_density_error": [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_double)],
        "akua_debug_probe_density": [P, P, i32],
        "akua_stream_copy_bandwidth": [ctypes.c_int64, i32, ctypes.POINTER(ctypes.c_double)],
        "akua_selftest_arith": [ctypes.c_uint64, ctypes.c_uint64, P],
        "akua_batch_create": [i32, P, ctypes.POINTER(CPBFConfig), ctypes.POINTER(CLambdaCorr), ctypes.POINTER(P)],
        "akua_batch_num_scenes": [P, ctypes.POINTER(i32)],
        "akua_batch_set_box": [P, i32, P, P],
        "akua_batch_step": [P, f32, i32],
        "akua_batch_density_error": [P, P, P],
        "akua_solver_set_tuning": [P, ctypes.POINTER(AkuaTuning)],
        "akua_solver_get_tuning": [P, ctypes.POINTER(AkuaTuning)],
        "akua_solver_set_search": [P, i32],
        "akua_solver_get_search": [P, P],
        "akua_solver_set_exact": [P, i32],
        "akua_solver_autotune": [P, f32, P, P, i32, ctypes.c_char_p, ctypes.POINTER(AkuaTuning)],
        "akua_solver_set_network": [P, P, i32],
        "akua_solver_get_network": [P, P, i32],
        "akua_solver_set_learned": [P, i32, i32, i32, f32, f32],
        "akua_solver_learned_stats": [P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), P],
        "akua_solver_train": [P, f32, P, P, i32, ctypes.POINTER(AkuaTrainConfig), P, P],
        "akua_debug_features": [P, P, P, i32],
        "akua_debug_warm_start": [P, P, P, P, i32],
        "akua_mlp_forward": [P, P, i32, P],
        "akua_mlp_gradient": [P, P, P, i32, P, P, P],
        "akua_mlp_adam_step": [P, P, P, ctypes.c_int64, f32, P, P, i32, P, P],
        "akua_mlp_adam_step_device": [P, P, P, ctypes.c_int64, f32, P, P, i32, P, P, P],
        "akua_solver_set_boundary": [P, i32, f32],
        "akua_debug_boundary": [P, P, P, i32, ctypes.POINTER(i32)],
        "akua_boundary_samples": [ctypes.POINTER(CPBFConfig), P, P, i32, f32, P, P, i32, ctypes.POINTER(i32)],
        "akua_scene_describe": [ctypes.c_char_p, f32, ctypes.c_int64, ctypes.POINTER(AkuaSceneInfo)],
        "akua_scene_generate": [ctypes.c_char_p, f32, ctypes.c_int64, P, i32],
        "akua_traj_create": [ctypes.c_char_p, i32, ctypes.c_uint32, f32, P, P, P, ctypes.POINTER(P)],
        "akua_traj_append": [P, ctypes.c_int64, ctypes.c_double, P, P, P, P, P, P],
        "akua_traj_append_solver": [P, P, ctypes.c_int64, ctypes.c_double, P],
        "akua_traj_open": [ctypes.c_char_p, ctypes.POINTER(AkuaTrajInfo), ctypes.POINTER(P)],
        "akua_traj_read": [P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_double),
                           P, P, P, P, P, P],
        "akua_traj_close": [P],
        "akua_camera_matrix": [ctypes.POINTER(AkuaCamera), i32, i32, P],
        "akua_render": [P, ctypes.POINTER(AkuaCamera), i32, i32, f32, i32, P],
    }
    L.akua_default_camera.argtypes = [ctypes.POINTER(AkuaCamera)]
    L.akua_default_camera.restype = None
    for name, args in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = i32
    _lib = L
    return L


def check(status: int) -> None:
    if status != AKUA_OK:
        msg = load().akua_last_error()
        raise AkuaError(status, msg.decode() if msg else "")
