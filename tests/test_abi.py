"""The C ABI library (libakua.so) loads and exports every function include/akua.h declares; the
error paths that need no GPU return the documented status codes."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "akua.h")
HEADERS = sorted(os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
                 if f.endswith(".h"))


def _declared_functions(headers=(HEADER,)):
    names = set()
    for h in headers:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(akua_[a-z_0-9]+)\s*\(", text,
                                flags=re.M))
    return sorted(names)


@pytest.fixture(scope="module")
def lib():
    from akuaengine_amd import _lib
    return _lib.load()


def test_header_declares_the_abi():
    names = _declared_functions()
    assert "akua_solver_create" in names and "akua_solver_step" in names
    from akuaengine_amd import _lib
    assert sorted(_lib.EXPORTED) == names
    io = os.path.join(ROOT, "include", "akua_io.h")
    assert sorted(_lib.EXPORTED_IO) == _declared_functions((io,))
    assert sorted(_lib.EXPORTED + _lib.EXPORTED_IO) == _declared_functions(HEADERS)


def test_every_declared_symbol_is_exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", os.path.join(ROOT, "akuaengine_amd", "libakua.so")],
                         capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    for name in _declared_functions(HEADERS):
        assert name in exported, name
        assert getattr(lib, name) is not None


def test_abi_version_and_struct_sizes(lib):
    from akuaengine_amd import _lib
    from akuaengine_amd.config import CLambdaCorr, CPBFConfig
    assert lib.akua_abi_version() == 1
    assert ctypes.sizeof(CPBFConfig) == 48     # sizeof(AkuaEngine::PBFConfig)
    assert ctypes.sizeof(CLambdaCorr) == 16    # sizeof(AkuaEngine::LambdaCorrParams)
    assert ctypes.sizeof(_lib.AkuaParticle) == 108  # sizeof(AkuaEngine::Particle)
    assert _lib.AkuaParticle.hash.offset == 84 and _lib.AkuaParticle.color.offset == 88


def test_default_config_matches_reference_defaults(lib):
    from akuaengine_amd.config import CLambdaCorr, CPBFConfig, LambdaCorrParams, PBFConfig
    c, k = CPBFConfig(), CLambdaCorr()
    lib.akua_default_config(ctypes.byref(c), ctypes.byref(k))
    d, dk = PBFConfig(), LambdaCorrParams()
    for f, _ in CPBFConfig._fields_:
        if f == "gravity":
            assert list(c.gravity) == pytest.approx(list(d.gravity))
        else:
            assert getattr(c, f) == pytest.approx(getattr(d, f))
    assert (k.enabled, k.k, k.n, k.delta_q) == (1, pytest.approx(dk.k), dk.n, pytest.approx(dk.delta_q))


def _create(lib, n, **over):
    from akuaengine_amd.config import CLambdaCorr, CPBFConfig
    c, k = CPBFConfig(), CLambdaCorr()
    lib.akua_default_config(ctypes.byref(c), ctypes.byref(k))
    for key, v in over.items():
        setattr(c, key, v)
    h = ctypes.c_void_p()
    st = lib.akua_solver_create(n, ctypes.byref(c), ctypes.byref(k), ctypes.byref(h))
    return st, h


@pytest.mark.parametrize("n,over", [(0, {}), (-5, {}), (10, {"maxNeighbours": 0}),
                                    (10, {"smoothRadius": 0.0}), (10, {"solverIterations": -1}),
                                    (10, {"restDensity": float("nan")})])
def test_invalid_arguments_are_rejected_before_any_device_call(lib, n, over):
    st, h = _create(lib, n, **over)
    assert st == 1  # AKUA_ERR_INVALID_ARG
    assert h.value is None
    assert lib.akua_last_error()


def test_null_pointers_are_rejected(lib):
    assert lib.akua_solver_create(10, None, None, None) == 1
    assert lib.akua_solver_step(None, ctypes.c_float(0.01), None, None) == 1
    assert lib.akua_solver_destroy(None) == 0


def test_tuning_struct_and_null_arguments(lib):
    from akuaengine_amd import _lib
    assert ctypes.sizeof(_lib.AkuaTuning) == 40
    t = _lib.AkuaTuning(256, 1, 0, 0, 1, 0, 0)
    assert lib.akua_solver_set_tuning(None, ctypes.byref(t)) == 1
    assert lib.akua_solver_get_tuning(None, ctypes.byref(t)) == 1
    assert lib.akua_solver_autotune(None, ctypes.c_float(0.01), None, None, 5, None, None) == 1


def test_no_device_is_reported(lib):
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    st, h = _create(lib, 100)
    assert st == 7  # AKUA_ERR_NO_DEVICE
    assert b"device" in lib.akua_last_error()


def test_solver_class_raises_without_device():
    from akuaengine_amd import PBFSolver, dam_break_params
    from akuaengine_amd._lib import AkuaError
    import torch
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    cfg, corr = dam_break_params()
    with pytest.raises(AkuaError):
        PBFSolver(64, cfg, corr)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    from akuaengine_amd import _lib
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "libakua.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(ImportError):
        _lib.load()
