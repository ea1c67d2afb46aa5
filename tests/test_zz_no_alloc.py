"""jse_nnap_computeGPUMD_noalloc, the GPUMD hook's variant that allocates
nothing (every buffer beyond GPUMD's inputs and outputs comes from the
library's static per-GPU scratch array), against jse_nnap_computeGPUMD on
the same packed arguments: two calls in a row, the second with a smaller
system, equal to the allocating entry point's results: bit for bit over all
particles (both run the same kernels, and the gather sums in a fixed order),
and within FP64 summation-order differences (1e-9) over parts of the range,
where both scatter with FP64 atomics whose order varies from run to run.  A
system too large for the scratch array is refused with status 6 and a
message naming the limit.  Collected last (the suite runs with -x)."""
import ctypes

import numpy as np
import pytest

from jse_amd import nnap
from jse_amd import workloads as W
from test_gpu_gpumd import gpumd_list
from test_gpu_lammps import DevBuf, addr, pack

pytestmark = pytest.mark.gpu


def _lib():
    lib = nnap.lib()
    for name in ("jse_nnap_gpumd_bind", "jse_nnap_computeGPUMD", "jse_nnap_computeGPUMD_noalloc"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    return lib


def _block(N, N1, N2, dev, out):
    return pack(("i", N), ("i", N1), ("i", N2), ("p", addr(dev["num"])), ("p", addr(dev["nbr"])),
                ("p", addr(dev["dx"])), ("p", addr(dev["dy"])), ("p", addr(dev["dz"])), ("p", addr(dev["type"])),
                ("p", 0), ("p", 0), ("p", 0), ("p", 0), ("p", 0), ("p", 0), ("p", 0),
                ("p", addr(out["fx"])), ("p", addr(out["fy"])), ("p", addr(out["fz"])), ("p", addr(out["vir"])),
                ("p", addr(out["pot"])), *([("p", 0)] * 9))


def _run(entry, pot, st, ranges, seed):
    lib = _lib()
    num, nbr, d = gpumd_list(st, pot.rcutMax() + 0.5, seed)
    n = st.natoms
    dev = {"num": DevBuf(num), "nbr": DevBuf(nbr.ravel()), "dx": DevBuf(d[0].ravel()), "dy": DevBuf(d[1].ravel()),
           "dz": DevBuf(d[2].ravel()), "type": DevBuf((st.types - 1).astype(np.int32))}
    base = {"fx": np.full(n, 0.25), "fy": np.full(n, -0.5), "fz": np.zeros(n), "vir": np.full(9 * n, 0.125),
            "pot": np.full(n, 1.0)}
    out = {k: DevBuf(v) for k, v in base.items()}
    for N1, N2 in ranges:
        rc = getattr(lib, entry)(_block(n, N1, N2, dev, out))
        assert rc == 0, lib.jse_nnap_last_error()
    return {k: out[k].numpy() - base[k] for k in base}


@pytest.mark.parametrize("precision", [nnap.FP32, nnap.BF16_MLP])
def test_noalloc_gpumd_equals_allocating_entry(precision):
    md = W.ternary_model()
    big = W.ternary_liquid(n=1500, seed=41)
    small = W.ternary_liquid(n=600, seed=42)
    lib = _lib()
    with nnap.NNAP(md) as pot:
        assert lib.jse_nnap_gpumd_bind(pack(("p", pot._h.value), ("i", precision))) == 0
        for st, seed in ((big, 1), (small, 2)):   # the second call is smaller
            n = st.natoms
            print(f"noalloc vs allocating: {n} atoms, precision {precision}", flush=True)
            for ranges in ([(0, n)], [(0, n // 3), (n // 3, n)]):
                want = _run("jse_nnap_computeGPUMD", pot, st, ranges, seed)
                got = _run("jse_nnap_computeGPUMD_noalloc", pot, st, ranges, seed)
                for k in want:
                    if len(ranges) == 1:
                        assert np.array_equal(got[k], want[k]), (n, ranges, k)
                    else:
                        assert np.abs(got[k] - want[k]).max() <= 1e-9, (n, ranges, k)


def test_noalloc_gpumd_rejects_a_system_beyond_its_scratch():
    md = W.ternary_model()
    st = W.ternary_liquid(n=12000, seed=43)
    lib = _lib()
    with nnap.NNAP(md) as pot:
        assert lib.jse_nnap_gpumd_bind(pack(("p", pot._h.value), ("i", nnap.FP32))) == 0
        num, nbr, d = gpumd_list(st, pot.rcutMax() + 0.5, 3)
        n = st.natoms
        dev = {"num": DevBuf(num), "nbr": DevBuf(nbr.ravel()), "dx": DevBuf(d[0].ravel()),
               "dy": DevBuf(d[1].ravel()), "dz": DevBuf(d[2].ravel()),
               "type": DevBuf((st.types - 1).astype(np.int32))}
        out = {k: DevBuf(v) for k, v in {"fx": np.zeros(n), "fy": np.zeros(n), "fz": np.zeros(n),
                                          "vir": np.zeros(9 * n), "pot": np.zeros(n)}.items()}
        rc = lib.jse_nnap_computeGPUMD_noalloc(_block(n, 0, n, dev, out))
        assert rc == 6
        assert "scratch" in lib.jse_nnap_last_error().decode()
