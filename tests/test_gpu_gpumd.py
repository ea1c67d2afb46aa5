"""The GPUMD hook through its jse-compatible packed-argument entry point
(jse_nnap_computeGPUMD; argument block packed as jse's SourceScanner wrapper
reads it: declaration order, no padding).

GPUMD's inputs are built from a periodic liquid: a full neighbour list on the
device, column-major [jj * N + i], with FP32 minimum-image displacements
r_j - r_i (here within rc + 0.5 A, in a shuffled order per atom) and 0-based
types.  The hook adds forces, per-atom virials (on the neighbour) and
per-atom energies into GPUMD's FP64 arrays; they must meet the FP32 bar
against the FP64 oracle and agree with the plain path.  A call over all
particles with a symmetric list goes through the atomic-free gather, calls
over part of the range through the reference's atomic scatter: two calls
over [0, N/2) and [N/2, N) must add up to one call over [0, N), and a list
with a missing reverse pair must fall back to the scatter."""
import ctypes

import numpy as np
import pytest

from jse_amd import model as M
from jse_amd import nnap
from jse_amd import workloads as W
from test_gpu_lammps import DevBuf, addr, pack

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def gpumd_list(st, rg, seed):
    """GPUMD's device layout of a full minimum-image neighbour list."""
    from scipy.spatial import cKDTree
    L = np.asarray(st.box, float)
    assert np.all(L > 2 * rg)
    x = st.xyz - np.floor(st.xyz / L) * L
    n = st.natoms
    tree = cKDTree(x, boxsize=L)
    rng = np.random.default_rng(seed)
    lists = [rng.permutation([j for j in tree.query_ball_point(x[i], rg) if j != i]) for i in range(n)]
    mn = max(len(l) for l in lists)
    num = np.array([len(l) for l in lists], dtype=np.int32)
    nbr = np.zeros((mn, n), dtype=np.int32)
    d = np.zeros((3, mn, n), dtype=np.float32)
    for i, l in enumerate(lists):
        dv = x[l] - x[i]
        dv -= np.round(dv / L) * L
        nbr[: len(l), i] = l
        d[:, : len(l), i] = dv.T.astype(np.float32)
    return num, nbr, d


def call(lib, N, N1, N2, dev, out):
    scratch = [("p", 0)] * 9
    rc_ = lib.jse_nnap_computeGPUMD(pack(
        ("i", N), ("i", N1), ("i", N2), ("p", addr(dev["num"])), ("p", addr(dev["nbr"])),
        ("p", addr(dev["dx"])), ("p", addr(dev["dy"])), ("p", addr(dev["dz"])), ("p", addr(dev["type"])),
        ("p", 0), ("p", 0), ("p", 0), ("p", 0), ("p", 0), ("p", 0), ("p", 0),
        ("p", addr(out["fx"])), ("p", addr(out["fy"])), ("p", addr(out["fz"])), ("p", addr(out["vir"])),
        ("p", addr(out["pot"])), *scratch))
    assert rc_ == 0, lib.jse_nnap_last_error()


def run_hook(pot, st, precision, ranges, seed=0, drop=None):
    """drop = (i, k): remove entry k of atom i's list (its pair then lacks the reverse)"""
    lib = nnap.lib()
    for name in ("jse_nnap_gpumd_bind", "jse_nnap_computeGPUMD"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    num, nbr, d = gpumd_list(st, pot.rcutMax() + 0.5, seed)
    if drop is not None:
        i, k = drop
        last = num[i] - 1
        nbr[k, i], d[:, k, i] = nbr[last, i], d[:, last, i]
        num[i] -= 1
    n = st.natoms
    dev = {"num": DevBuf(num), "nbr": DevBuf(nbr.ravel()), "dx": DevBuf(d[0].ravel()), "dy": DevBuf(d[1].ravel()),
           "dz": DevBuf(d[2].ravel()), "type": DevBuf((st.types - 1).astype(np.int32))}
    # GPUMD's outputs start from other terms' values: the hook must add to them
    base = {"fx": np.full(n, 0.25), "fy": np.full(n, -0.5), "fz": np.zeros(n), "vir": np.full(9 * n, 0.125),
            "pot": np.full(n, 1.0)}
    out = {k: DevBuf(v) for k, v in base.items()}
    assert lib.jse_nnap_gpumd_bind(pack(("p", pot._h.value), ("i", precision))) == 0
    for N1, N2 in ranges:
        call(lib, n, N1, N2, dev, out)
    got = {k: out[k].numpy() - base[k] for k in base}
    return {"energies": got["pot"], "forces": np.stack([got["fx"], got["fy"], got["fz"]], 1),
            "virials": got["vir"].reshape(9, n).T}


def test_gpumd_hook_matches_oracle_and_plain_path(O):
    st = W.ternary_liquid(n=1000, seed=5)
    md = W.ternary_model()
    n = st.natoms
    with nnap.NNAP(md) as pot:
        got = run_hook(pot, st, nnap.FP32, [(0, n)], seed=2)
        split = run_hook(pot, st, nnap.FP32, [(0, n // 2), (n // 2, n)], seed=2)
        plain = pot.calEnergyForceVirials(nnap.AtomData(st.box, st.xyz, st.types))
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    # the FP32 bar against the FP64 oracle (the displacements arrive in FP32)
    df = got["forces"] - ref["forces"]
    assert abs(got["energies"].sum() - ref["energies"].sum()) / n <= 1e-6
    assert np.abs(got["energies"] - ref["energies"]).max() <= 1e-5
    assert np.sqrt((df ** 2).mean()) <= 1e-5
    assert np.abs(got["virials"].sum(0) - ref["virials"].sum(0)).max() <= 1e-5 * n
    assert np.abs(got["virials"] - ref["virials"]).max() <= 1e-4
    # the plain path (FP64 displacements, gathered forces)
    assert np.abs(got["forces"] - plain["forces"]).max() <= 1e-4
    assert np.abs(got["virials"] - plain["virials"]).max() <= 1e-4
    # split centre ranges (the atomic scatter) add up to the whole (the gather):
    # summation-order differences only
    for k in ("energies", "forces", "virials"):
        assert np.abs(split[k] - got[k]).max() <= 1e-9


def test_gpumd_hook_asymmetric_list_falls_back_to_scatter():
    """A list in which one pair lacks its reverse cannot go through the
    gather: the whole-range call must detect it and scatter, matching the
    split-range calls (which always scatter) on the same list."""
    st = W.ternary_liquid(n=1000, seed=6)
    md = W.ternary_model()
    n = st.natoms
    with nnap.NNAP(md) as pot:
        whole = run_hook(pot, st, nnap.FP32, [(0, n)], seed=4, drop=(17, 0))
        split = run_hook(pot, st, nnap.FP32, [(0, n // 3), (n // 3, n)], seed=4, drop=(17, 0))
        sym = run_hook(pot, st, nnap.FP32, [(0, n)], seed=4)
    for k in ("energies", "forces", "virials"):
        assert np.abs(whole[k] - split[k]).max() <= 1e-9
    assert np.abs(whole["forces"] - sym["forces"]).max() > 1e-6  # the dropped pair mattered
