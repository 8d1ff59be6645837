import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sop-cpu_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libsopg.so on the device)")


def gpu_available():
    try:
        import sopg
        n = sopg.C.c_int()
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


_cache = {}


def chain_inputs(tmp_root, n, seed, kT=0.6):
    """Generated native chain (+ Maxwell velocities) written in the reference's formats."""
    import sopg
    key = (n, seed)
    if key not in _cache:
        x = sopg.gen_native(n, seed)
        v = sopg.gen_maxwell(n, kT, seed)
        d = os.path.join(str(tmp_root), f"chain_{n}_{seed}")
        sopg.write_inputs(x, d, vel=v)
        _cache[key] = (x, v, d)
    return _cache[key]


@pytest.fixture(scope="session")
def inputs_root(tmp_path_factory):
    return tmp_path_factory.mktemp("inputs")
