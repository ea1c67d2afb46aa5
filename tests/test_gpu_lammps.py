"""The LAMMPS GPU pair hook through its jse-compatible packed-argument entry
points (jse_nnap_statNeiNumLammps / lammps2cuda / computeLammpsCuda /
cuda2lammps; argument blocks packed exactly as jse's SourceScanner wrapper
reads them: declaration order, no padding).

A LAMMPS-style domain is built from a periodic liquid: the local atoms, the
ghost atoms (periodic images within rc + skin of the box) and, per local
atom, the full neighbour list of local and ghost atoms within rc + skin.
After ghost forces are folded onto their owners, energies, forces and
virials equal the plain path's (FP32 rounding of a different neighbour
order) and meet the FP32 bar against the FP64 oracle."""
import ctypes
import struct

import numpy as np
import pytest

from jse_amd import model as M
from jse_amd import nnap
from jse_amd import workloads as W
from conftest import DevBuf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def pack(*items):
    """('i', int) -> 4 bytes, ('p', address) -> 8 bytes, back to back."""
    fmt = "<" + "".join("i" if k == "i" else "Q" for k, _ in items)
    buf = ctypes.create_string_buffer(struct.pack(fmt, *[int(v) for _, v in items]))
    return buf


def addr(a):
    return a.ctypes.data if isinstance(a, np.ndarray) else a.ptr


def rows_ptrs(a):
    """double** / int** view of a 2-D C-contiguous array"""
    return np.array([a.ctypes.data + r * a.strides[0] for r in range(a.shape[0])], dtype=np.uint64)


def lammps_domain(st, rg, seed):
    from scipy.spatial import cKDTree
    L = np.asarray(st.box, float)
    x = st.xyz - np.floor(st.xyz / L) * L
    n = st.natoms
    gx, own = [], []
    for sx in (-1, 0, 1):
        for sy in (-1, 0, 1):
            for sz in (-1, 0, 1):
                if sx == sy == sz == 0:
                    continue
                img = x + np.array([sx, sy, sz]) * L
                m = np.all((img > -rg) & (img < L + rg), axis=1)
                gx.append(img[m])
                own.append(np.nonzero(m)[0])
    allx = np.ascontiguousarray(np.concatenate([x] + gx))
    owner = np.concatenate([np.arange(n)] + own)
    tree = cKDTree(allx)
    lists = [sorted(j for j in tree.query_ball_point(x[i], rg) if j != i) for i in range(n)]
    rng = np.random.default_rng(seed)
    ilist = rng.permutation(n).astype(np.int32)
    return allx, owner, lists, ilist


def run_hook(pot, st, precision, cv=False, seed=0):
    lib = nnap.lib()
    for name in ("jse_nnap_lammps_bind", "jse_nnap_statNeiNumLammps", "jse_nnap_lammps2cuda",
                 "jse_nnap_computeLammpsCuda", "jse_nnap_cuda2lammps"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    rc = pot.rcutMax()
    allx, owner, lists, ilist = lammps_domain(st, rc + 1.0, seed)
    n, nall = st.natoms, allx.shape[0]
    # LAMMPS types 1..3 map onto the model's types through a permutation
    lmp_of = np.array([0, 3, 1, 2], dtype=np.int32)       # model type -> LAMMPS type
    tmap = np.zeros(4, dtype=np.int32)
    tmap[lmp_of[1:]] = np.arange(1, 4)                     # LAMMPS type -> model type
    ltype = lmp_of[st.types[owner]].astype(np.int32)
    numneigh = np.array([len(l) for l in lists], dtype=np.int32)
    neigh_rows = [np.array(l, dtype=np.int32) for l in lists]
    firstneigh = np.array([r.ctypes.data for r in neigh_rows], dtype=np.uint64)
    assert lib.jse_nnap_lammps_bind(pack(("p", pot._h.value), ("i", precision))) == 0
    out = np.zeros(1, dtype=np.int32)
    assert lib.jse_nnap_statNeiNumLammps(pack(("p", addr(ilist)), ("p", addr(numneigh)), ("i", n),
                                              ("p", addr(out)))) == 0
    nmax = int(out[0])
    assert nmax == numneigh.max()
    cudaX = DevBuf(np.zeros(3 * nall, np.float32))
    cudaType = DevBuf(np.zeros(nall, np.int32))
    cudaIlist = DevBuf(np.zeros(n, np.int32))
    cudaNumneigh = DevBuf(np.zeros(n, np.int32))
    cudaFirst = DevBuf(np.zeros(n * nmax, np.int32))
    cudaMap = DevBuf(tmap)
    f32 = DevBuf(np.zeros(3 * nall, np.float32))
    e32 = DevBuf(np.zeros(n, np.float32))
    v0 = DevBuf(np.zeros(6 * n, np.float32))
    v1 = DevBuf(np.zeros(9 * nall, np.float32))
    fltBuf = np.zeros(9 * nall, dtype=np.float32)
    intBuf = np.zeros(n * nmax, dtype=np.int32)
    xptr = rows_ptrs(allx)
    rc_ = lib.jse_nnap_lammps2cuda(pack(
        ("i", n), ("i", nall), ("i", 1), ("i", nmax), ("p", addr(xptr)), ("p", addr(ltype)), ("p", addr(ilist)),
        ("p", addr(numneigh)), ("p", addr(firstneigh)), ("p", addr(fltBuf)), ("p", addr(intBuf)),
        ("p", addr(cudaX)), ("p", addr(cudaType)), ("p", addr(cudaIlist)), ("p", addr(cudaNumneigh)),
        ("p", addr(cudaFirst))))
    assert rc_ == 0, lib.jse_nnap_last_error()
    # the caller's device copies are the reference's layouts
    assert np.array_equal(cudaX.numpy().reshape(3, nall), allx.T.astype(np.float32))
    assert np.array_equal(cudaNumneigh.numpy(), numneigh[ilist])
    vflag_atom = 1
    scratch = [("p", 0)] * 10
    rc_ = lib.jse_nnap_computeLammpsCuda(pack(
        ("i", n), ("i", nall), ("i", 1), ("i", 1), ("i", 1), ("i", vflag_atom), ("i", 1 if cv else 0),
        ("p", addr(cudaX)), ("p", addr(cudaType)), ("p", addr(cudaIlist)), ("p", 0), ("p", 0), ("p", 0),
        ("p", addr(cudaNumneigh)), ("p", addr(cudaFirst)), ("p", addr(cudaMap)),
        ("p", 0), ("p", 0), ("p", 0), ("p", 0),
        ("p", addr(f32)), ("p", addr(e32)), ("p", addr(v0)), ("p", addr(v1)), *scratch))
    assert rc_ == 0, lib.jse_nnap_last_error()
    f = np.zeros((nall, 3))
    eng = np.zeros(1)
    eatom = np.zeros(nall)
    virial = np.zeros(6)
    vatom = np.zeros((nall, 6))
    cvatom = np.zeros((nall, 9))
    fptr, vptr, cvptr = rows_ptrs(f), rows_ptrs(vatom), rows_ptrs(cvatom)   # kept alive across the call
    rc_ = lib.jse_nnap_cuda2lammps(pack(
        ("i", n), ("i", nall), ("i", 1), ("i", 1), ("i", 1), ("i", vflag_atom), ("i", 1 if cv else 0),
        ("p", addr(fptr)), ("p", addr(eng)), ("p", addr(eatom)), ("p", addr(virial)),
        ("p", addr(vptr)), ("p", addr(cvptr)), ("p", addr(fltBuf)), ("p", addr(ilist)),
        ("p", addr(f32)), ("p", addr(e32)), ("p", addr(v0)), ("p", addr(v1))))
    assert rc_ == 0, lib.jse_nnap_last_error()
    # fold ghost contributions onto their owners
    F = np.zeros((n, 3))
    np.add.at(F, owner, f)
    Va = np.zeros((n, 9 if cv else 6))
    np.add.at(Va, owner, cvatom if cv else vatom)
    return {"energy": float(eng[0]), "energies": eatom[:n], "forces": F, "virial": virial, "virials": Va,
            "f32": f32.numpy().reshape(3, nall).T, "f64": f}


@pytest.mark.parametrize("cv", [False, True])
def test_lammps_hook_matches_plain_path_and_oracle(O, cv):
    st = W.ternary_liquid(n=2000, seed=3)
    md = W.ternary_model()
    with nnap.NNAP(md) as pot:
        got = run_hook(pot, st, nnap.FP32, cv=cv, seed=1)
        plain = pot.calEnergyForceVirials(nnap.AtomData(st.box, st.xyz, st.types))
    n = st.natoms
    assert abs(got["energy"] - plain["energy"]) <= 1e-6 * n
    assert np.abs(got["energies"] - plain["energies"]).max() <= 1e-5
    assert np.abs(got["forces"] - plain["forces"]).max() <= 1e-4
    assert np.sqrt(((got["forces"] - plain["forces"]) ** 2).mean()) <= 1e-5
    # LAMMPS' virial order xx yy zz xy xz yz
    assert np.abs(got["virial"] - plain["virial"][:6]).max() <= 1e-5 * n
    ncomp = 9 if cv else 6
    assert np.abs(got["virials"] - plain["virials"][:, :ncomp]).max() <= 1e-4
    assert np.abs(got["f32"] - got["f64"]).max() <= 1e-5 * (1 + np.abs(got["f64"]).max())
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    df = got["forces"] - ref["forces"]
    assert abs(got["energy"] - ref["energies"].sum()) / n <= 1e-6
    assert np.sqrt((df ** 2).mean()) <= 1e-5
    assert np.abs(got["virial"] - ref["virials"].sum(0)[:6]).max() / n <= 1e-5
