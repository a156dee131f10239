"""CPU-side checks of the C ABI library: it loads, it exports exactly what
include/parm_hip.h declares, and without a GPU it fails loudly instead of
computing anything on the CPU."""
import ctypes
import os
import re

import pytest

from parm_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "parm_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(parm_[a-z0-9_]+)\s*\(", txt))


def test_header_matches_binding_list():
    assert _declared() == set(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [s for s in _declared() if not hasattr(lib, s)]
    assert not missing, missing


def test_version_and_last_error():
    maj, mnr = ctypes.c_int(), ctypes.c_int()
    assert capi.lib().parm_version(ctypes.byref(maj), ctypes.byref(mnr)) == 0
    assert maj.value == 1
    assert isinstance(capi.lib().parm_last_error(), bytes)


def test_no_gpu_means_loud_failure():
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible here")
    with pytest.raises(RuntimeError, match="no HIP device"):
        capi.System(3, [1.0, 1.0])


def test_invalid_arguments_rejected_before_device():
    with pytest.raises(ValueError):
        capi.System(4, [1.0])


def test_batch_no_gpu_means_loud_failure():
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible here")
    with pytest.raises(RuntimeError, match="no HIP device"):
        capi.Batch(3, [4, 8], [[2.0] * 3, [3.0] * 3], 0.2)
    with pytest.raises(ValueError):
        capi.Batch(3, [4, 0], [[2.0] * 3, [3.0] * 3], 0.2)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message before
    touching any device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "exactly one GPU" in r.stderr
    assert r.stdout == ""
