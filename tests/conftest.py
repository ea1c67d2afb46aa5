import ctypes
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    return oracle


class DevBuf:
    """Device memory from the HIP runtime the library itself uses (the
    buffers jse's Java side would own)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            h = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
            h.hipFree.argtypes = [ctypes.c_void_p]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
            cls._hip = h
        return cls._hip

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        p = ctypes.c_void_p()
        assert self.hip().hipMalloc(ctypes.byref(p), max(1, self.host.nbytes)) == 0
        self.ptr = p.value
        assert self.hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0  # host to device

    def numpy(self):
        out = np.empty_like(self.host)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0  # device to host
        return out

    def __del__(self):
        try:
            self.hip().hipFree(self.ptr)
        except Exception:
            pass
