"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and its host-side model parser agrees with the Python loader and
reports errors with the reference's wording.  No GPU calls."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from jse_amd import model as M
from jse_amd import nnap
from jse_amd import workloads as W


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"\b(jse_nnap_\w+)\s*\(", txt))
    return names


def test_library_exports_header_symbols():
    L = nnap.lib()
    declared = _declared()
    assert declared, "no declarations found"
    for name in sorted(declared):
        assert hasattr(L, name), f"{name} declared in include/ but not exported"
    assert set(nnap.EXPORTS) <= declared


def _parse(text):
    L = nnap.lib()
    nt = ctypes.c_int()
    P = (ctypes.c_int * 8)()
    S = (ctypes.c_int * 8)()
    rc = L.jse_nnap_model_parse(text.encode(), ctypes.byref(nt), P, S)
    return rc, nt.value, list(P)[:nt.value], list(S)[:nt.value]


@pytest.mark.parametrize("wtype", ["default", "none", "full", "exfull"])
def test_cpp_parser_matches_python_loader(wtype):
    md = M.random_model(["Cu", "Zr", "Al"], 4, 5, 5.0, (8, 8), seed=2, wtype=wtype)
    if wtype in ("full", "exfull", "none"):
        # random_model sizes the default layout; rebuild the NN for the requested wtype
        mdl0 = M.load_model({"models": [dict(m, nn=None) for m in md["models"]]})
        md = M.random_model(["Cu", "Zr", "Al"], 4, 5, 5.0, (8, 8), seed=2, wtype="default")
        for m, sp in zip(md["models"], mdl0.species):
            m["basis"]["wtype"] = wtype
            m["nn"]["input_dim"] = sp.S
            rng = np.random.default_rng(1)
            m["nn"]["hidden_weights"][0] = rng.normal(size=(8, sp.S)).tolist()
            m["norm_mu"] = [0.0] * sp.S
            m["norm_sigma"] = [1.0] * sp.S
    rc, nt, P, S = _parse(json.dumps(md))
    assert rc == 0, nnap.lib().jse_nnap_last_error()
    mdl = M.load_model(md)
    assert nt == 3 and P == [s.P for s in mdl.species] and S == [s.S for s in mdl.species]


@pytest.mark.parametrize("kind", ["share", "mirror"])
def test_cpp_parser_share_mirror(kind):
    """mirror / share entries take another species' basis (Basis.java:22-67):
    both loaders agree on the sizes, and a mirror or share of a mirror or share
    is rejected by both."""
    md = M.random_model(["Cu", "Zr", "Al"], 4, 5, 5.0, (8, 8), seed=3)
    md["models"][0]["basis"]["rcut"] = 4.5
    md["models"][2]["basis"] = {"type": kind, kind: 1}
    rc, nt, P, S = _parse(json.dumps(md))
    assert rc == 0, nnap.lib().jse_nnap_last_error()
    mdl = M.load_model(md)
    assert P == [s.P for s in mdl.species] and S == [s.S for s in mdl.species]
    assert mdl.species[2].rcut == 4.5 and mdl.species[2].mirror == (1 if kind == "mirror" else 0)
    md["models"][1]["basis"] = {"type": kind, kind: 3}
    rc, _, _, _ = _parse(json.dumps(md))
    assert rc != 0
    with pytest.raises(ValueError):
        M.load_model(md)
    md["models"][1]["basis"] = {"type": kind}
    assert _parse(json.dumps(md))[0] != 0
    with pytest.raises(ValueError, match="required"):
        M.load_model(md)


def test_cpp_parser_post_fuse():
    """A post-fuse weight (P rows over the base mixing's sizeN) gives P mixed
    channels in both loaders; with wtype rfuse its keys are rejected."""
    md = M.random_model(["Cu", "Zr", "Al"], 4, 5, 5.0, (8, 8), seed=4)
    sn = M.load_model(md).species[0].P
    P = 7
    rng = np.random.default_rng(5)
    for m in md["models"]:
        m["basis"]["post_fuse_weight"] = rng.normal(size=P * sn).tolist()
        m["basis"]["post_fuse_scale"] = 0.5
        m["nn"]["input_dim"] = P * 6
        m["nn"]["hidden_weights"][0] = rng.normal(size=(8, P * 6)).tolist()
        m["norm_mu"] = [0.0] * (P * 6)
        m["norm_sigma"] = [1.0] * (P * 6)
    rc, nt, Ps, S = _parse(json.dumps(md))
    assert rc == 0, nnap.lib().jse_nnap_last_error()
    mdl = M.load_model(md)
    assert Ps == [P] * 3 == [s.P for s in mdl.species] and S == [s.S for s in mdl.species]
    md["models"][0]["basis"]["wtype"] = "rfuse"
    assert _parse(json.dumps(md))[0] != 0
    with pytest.raises(ValueError):
        M.load_model(md)


def test_cpp_parser_radial_basis():
    """The radial-only `chebyshev` basis (no lmax; S = P) in both loaders."""
    md = M.random_model(["Cu", "Zr", "Al"], 4, 0, 5.0, (8, 8), seed=6)
    for m in md["models"]:
        b = m["basis"]
        m["basis"] = {"type": "chebyshev", "nmax": b["nmax"], "rcut": b["rcut"], "wtype": b["wtype"]}
    rc, nt, P, S = _parse(json.dumps(md))
    assert rc == 0, nnap.lib().jse_nnap_last_error()
    mdl = M.load_model(md)
    assert P == S == [s.P for s in mdl.species] == [10] * 3 and all(s.radial for s in mdl.species)


def test_parser_errors():
    L = nnap.lib()
    rc, *_ = _parse('{"models": [{"basis": {}}]}')
    assert rc == 1 and b"No symbol in model" in L.jse_nnap_last_error()
    rc, *_ = _parse('{"version": 7, "models": []}')
    assert rc == 1 and b"Unsupported version" in L.jse_nnap_last_error()
    rc, *_ = _parse('{"models": [{"symbol": "Cu", "basis": {"lmax": 13}}]}')
    assert rc == 1 and b"lmax MUST be in [0, 12]" in L.jse_nnap_last_error()
    rc, *_ = _parse('{"models": [{"symbol": "Cu", "basis": {"nmax": 21}}]}')
    assert rc == 1 and b"nmax MUST be in [0, 20]" in L.jse_nnap_last_error()
    rc, *_ = _parse("/nonexistent/model.json")
    assert rc == 1 and b"cannot open model file" in L.jse_nnap_last_error()
    rc, *_ = _parse('{"models": [')
    assert rc == 1


def test_python_loader_packing_matches_reference_layout():
    md = M.random_model(["Cu", "Zr"], 2, 1, 4.0, (3,), seed=0)
    md["models"][0]["norm_mu_eng"] = 0.5
    md["models"][0]["norm_sigma_eng"] = 2.0
    md["models"][1]["ref_eng"] = -1.25
    mdl = M.load_model(md)
    sp = mdl.species[1]
    nn = md["models"][1]["nn"]
    want = np.concatenate([np.ravel(nn["hidden_weights"][0]), nn["output_weight"], nn["hidden_biases"][0],
                           [nn["output_bias"]]])
    assert np.array_equal(sp.nn_param, want)
    assert sp.norm_param[0] == 0.5 - 1.25 and sp.norm_param[1] == 2.0
    assert sp.S == 2 * 3 * 2


def test_workload_geometry():
    st = W.ternary_liquid(n=2000, seed=3)
    assert st.natoms == 2000 and np.isclose(st.box[0] ** 3 * 0.058, 2000)
    d = st.xyz[:, None, :] - st.xyz[None, :, :]
    d -= st.box * np.round(d / st.box)
    r = np.sqrt((d ** 2).sum(-1)) + np.eye(2000) * 10
    assert r.min() >= 2.0
    assert np.bincount(st.types).tolist() == [0, 1000, 800, 200]


def _pack(*items):
    import ctypes
    import struct
    fmt = "<" + "".join("i" if k == "i" else "Q" for k, _ in items)
    return ctypes.create_string_buffer(struct.pack(fmt, *[int(v) for _, v in items]))


def test_lammps_packed_host_entries():
    """Host-only parts of the LAMMPS hook: the neighbour-count statistic on
    a packed (int *ilist, int *numneigh, int inum, int *numneighMax) block,
    and the model binding's argument checks (no device is touched)."""
    import ctypes
    lib = nnap.lib()
    lib.jse_nnap_statNeiNumLammps.argtypes = [ctypes.c_void_p]
    lib.jse_nnap_lammps_bind.argtypes = [ctypes.c_void_p]
    ilist = np.array([2, 0, 1], dtype=np.int32)
    numneigh = np.array([5, 9, 7, 40], dtype=np.int32)   # atom 3 is not listed
    out = np.zeros(1, dtype=np.int32)
    blk = _pack(("p", ilist.ctypes.data), ("p", numneigh.ctypes.data), ("i", 3), ("p", out.ctypes.data))
    assert lib.jse_nnap_statNeiNumLammps(blk) == 0 and out[0] == 9
    assert lib.jse_nnap_lammps_bind(_pack(("p", 0), ("i", 0))) == 1
    assert b"null model" in lib.jse_nnap_last_error()
