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
    if kind == "mirror":  # a mirror entry has no network or normalisation of its own
        md["models"][2] = {"symbol": "Al", "basis": md["models"][2]["basis"]}
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


def _blocks(text, ctype, sp):
    """The C++ loader's packed blocks of species `ctype`, sized from the
    Python loader's species `sp`."""
    L = nnap.lib()
    nt = 0
    fp = np.zeros(sp.W.size)
    nn = np.zeros(sp.nn_param.size)
    norm = np.zeros(sp.norm_param.size)
    nl = ctypes.c_int()
    hid = (ctypes.c_int * 8)()
    mir = ctypes.c_int()
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    rc = L.jse_nnap_model_parse_blocks(text.encode(), ctype, dp(fp), dp(nn), dp(norm), ctypes.byref(nl), hid,
                                       ctypes.byref(mir))
    assert rc == 0, nnap.lib().jse_nnap_last_error()
    return fp, nn, norm, list(hid)[:nl.value], mir.value


def _both_equal(md):
    """Both loaders build identical blocks for every species of `md`."""
    mdl = M.load_model(md)
    text = json.dumps(md)
    for t, sp in enumerate(mdl.species):
        fp, nn, norm, hid, mir = _blocks(text, t + 1, sp)
        assert np.array_equal(fp, sp.W.ravel()), f"species {t + 1} fp_param"
        assert np.array_equal(nn, sp.nn_param), f"species {t + 1} nn_param"
        assert np.array_equal(norm, sp.norm_param), f"species {t + 1} norm_param"
        assert hid == list(sp.hidden) and mir == sp.mirror
    return mdl


def _mirror_model():
    """Cu/Zr/Al with Zr a mirror of Cu in the reference's own shape: the mirror
    entry carries no nn, ref_eng or normalisation (NNAP.java:281-288,
    NeuralNetwork.java:53-56)."""
    md = M.random_model(["Cu", "Zr", "Al"], 4, 5, 5.0, (8, 8), seed=5)
    rng = np.random.default_rng(5)
    for m in md["models"]:
        S = m["nn"]["input_dim"]
        m["norm_mu"] = rng.normal(size=S).tolist()
        m["norm_sigma"] = (1.0 + rng.random(S)).tolist()
        m["ref_eng"] = float(rng.normal())
    md["models"][1] = {"symbol": "Zr", "basis": {"type": "mirror", "mirror": 1}}
    return md


def test_mirror_species_takes_the_mirrored_network_and_normalisation():
    """A mirror species evaluates with the full network (SharedFeedForward.full)
    and the ref_eng / norm_mu / norm_sigma of the species it mirrors, in both
    loaders."""
    md = _mirror_model()
    mdl = _both_equal(md)
    a, b = mdl.species[0], mdl.species[1]
    assert b.mirror == 1 and b.hidden == a.hidden
    assert np.array_equal(b.nn_param, a.nn_param) and np.array_equal(b.norm_param, a.norm_param)
    assert not np.array_equal(mdl.species[2].nn_param[:10], a.nn_param[:10])


@pytest.mark.parametrize("key,msg", [("nn", "nn data in mirror"), ("ref_eng", "ref_eng in mirror"),
                                     ("norm_mu", "norm_vec/norm_sigma/norm_mu in mirror"),
                                     ("norm_vec", "norm_vec/norm_sigma/norm_mu in mirror")])
def test_mirror_species_rejects_its_own_network_or_normalisation(key, msg):
    md = _mirror_model()
    S = md["models"][0]["nn"]["input_dim"]
    md["models"][1][key] = md["models"][0]["nn"] if key == "nn" else ([0.5] * S if key != "ref_eng" else 0.1)
    rc = _parse(json.dumps(md))[0]
    assert rc == 1 and msg in nnap.lib().jse_nnap_last_error().decode()
    with pytest.raises(ValueError, match=msg.split()[0]):
        M.load_model(md)


def test_shared_feed_forward_network():
    """shared_feed_forward (SharedFeedForward.java:163-198,235-353): species 3
    takes species 1's network; per hidden layer its first (hidden - shared)
    rows and biases are its own, the rest are species 1's; the output layer is
    its own unless shared.  Fully shared it equals species 1's network."""
    md = M.random_model(["Cu", "Zr", "Al"], 4, 5, 5.0, (8, 6), seed=6)
    S = md["models"][0]["nn"]["input_dim"]
    rng = np.random.default_rng(6)
    own_w = [rng.normal(size=(3, S)).tolist(), rng.normal(size=(6, 8)).tolist()]
    own_b = [rng.normal(size=3).tolist(), rng.normal(size=6).tolist()]
    md["models"][2]["nn"] = {"type": "shared_feed_forward", "share": 1, "shared_hidden_dims": [5, 0],
                             "hidden_weights": own_w, "hidden_biases": own_b,
                             "output_weight": rng.normal(size=6).tolist(), "output_bias": 0.25}
    mdl = _both_equal(md)
    base, sh = mdl.species[0].nn_param, mdl.species[2].nn_param
    w1 = sh[:8 * S].reshape(8, S)
    assert np.array_equal(w1[:3], np.asarray(own_w[0])) and np.array_equal(w1[3:], base[:8 * S].reshape(8, S)[3:])
    w2 = sh[8 * S:8 * S + 48].reshape(6, 8)
    assert np.array_equal(w2, np.asarray(own_w[1]))
    hb = 8 * S + 48 + 6
    assert np.array_equal(sh[hb:hb + 3], own_b[0]) and np.array_equal(sh[hb + 3:hb + 8], base[hb + 3:hb + 8])
    assert sh[-1] == 0.25
    md["models"][2]["nn"] = {"type": "shared_feed_forward", "share": 1, "shared_hidden_dims": [8, 6, 1]}
    mdl = _both_equal(md)
    assert np.array_equal(mdl.species[2].nn_param, mdl.species[0].nn_param)
    md["models"][2]["nn"] = {"type": "shared_feed_forward", "share": 1}
    assert _parse(json.dumps(md))[0] == 1
    with pytest.raises(ValueError, match="shared_hidden_dims"):
        M.load_model(md)


def test_rfuse_weight_is_ignored_for_other_wtypes():
    """The reference reads rfuse_weight only for wtype rfuse and fuse_weight only
    for fuse/exfuse (WTypeBasis.java:365-410): a stray key of the wrong size
    changes nothing, in either loader."""
    md = M.random_model(["Cu", "Zr"], 4, 5, 5.0, (8, 8), seed=7, wtype="default")
    for m in md["models"]:
        m["basis"]["wtype"] = "exfull"
        m["nn"] = None
        m.pop("norm_mu", None)
        m.pop("norm_sigma", None)
    plain = _both_equal(md)
    for m in md["models"]:
        m["basis"]["rfuse_weight"] = [1.0] * (3 * 15 * 5)
        m["basis"]["fuse_weight"] = [2.0] * 7
    stray = _both_equal(md)
    for a, b in zip(plain.species, stray.species):
        assert a.P == b.P == 15 and np.array_equal(a.W, b.W)


def test_weight_standardization_epsilon():
    """Weight standardisation divides each rfuse row by (WS_EPS + sum |w|) / sizeN
    with WS_EPS = 1e-5 (WTypeBasis.java:20,210-226): a row whose only non-zero
    weight is w = 1e-5 becomes scale * w * sizeN / (1e-5 + w) = scale * sizeN / 2."""
    md = M.random_model(["Cu", "Zr"], 3, 2, 5.0, (8, 8), seed=8, wtype="default")
    K, P, sn = 4, 3, 2 * 4
    rw = np.zeros((2, P, K))
    rw[1, 0, 2] = 1e-5                       # row 0: one small weight
    rw[:, 1, :] = np.arange(2 * K).reshape(2, K) - 3.0   # row 1: ordinary weights
    rw[0, 2, 0] = 4.0                        # row 2: one large weight
    for m in md["models"]:
        m["basis"].update(wtype="rfuse", rfuse_weight=rw.ravel().tolist(), rfuse_scale=0.5,
                          weight_standardization=True)
        m["nn"] = None
        m.pop("norm_mu", None)
        m.pop("norm_sigma", None)
    mdl = _both_equal(md)
    W = mdl.species[0].W
    assert W.shape == (2, P, K)
    assert abs(W[1, 0, 2] - 0.5 * sn / 2) <= 1e-12 * sn
    assert abs(W[0, 2, 0] - 0.5 * 4.0 * sn / (1e-5 + 4.0)) <= 1e-12
    norm1 = (1e-5 + np.abs(rw[:, 1, :]).sum()) / sn
    assert np.allclose(W[:, 1, :], 0.5 * rw[:, 1, :] / norm1, rtol=1e-14, atol=0)


@pytest.mark.parametrize("form", ["flat", "rows", "old_columns", "size_only"])
def test_fuse_weight_formats(form):
    """fuse_weight as a flat [ntypes*fuse] list (with fuse_size), as rows of a
    [ntypes][fuse] matrix, as an older file's columns (no fuse_size), or zeros
    from fuse_size alone (WTypeBasis.getFuseWeight_, :365-389); both loaders
    give the same mixing."""
    md = M.random_model(["Cu", "Zr", "Al"], 3, 2, 5.0, (8, 8), seed=9, wtype="default")
    fw = np.arange(1, 3 * 2 + 1, dtype=float).reshape(3, 2) / 4     # [ntypes][fuse]
    for m in md["models"]:
        b = m["basis"]
        b["wtype"] = "fuse"
        if form == "flat":
            b.update(fuse_size=2, fuse_weight=fw.ravel().tolist())
        elif form == "rows":
            b.update(fuse_size=2, fuse_weight=fw.tolist())
        elif form == "old_columns":
            b.update(fuse_weight=fw.T.tolist())
        else:
            b.update(fuse_size=2)
        m["nn"] = None
        m.pop("norm_mu", None)
        m.pop("norm_sigma", None)
    mdl = _both_equal(md)
    W = mdl.species[0].W                    # [t][k*K + n][n] = fw[t][k]
    want = fw if form != "size_only" else np.zeros_like(fw)
    for t in range(3):
        for k in range(2):
            assert W[t, k * 4 + 1, 1] == want[t, k]


def test_cpp_parser_merge_basis():
    """merge (MergedBasis.java:204-229): both loaders give each species the
    parts' total descriptor size, the network over that size and part 0's
    mixing in the packed blocks; both refuse a merge without `basis`, with an
    empty list, or with a part of another type."""
    parts = [{"type": "spherical_chebyshev", "nmax": 4, "lmax": 4, "rcut": 5.5, "wtype": "default"},
             {"type": "chebyshev", "nmax": 6, "rcut": 4.0, "wtype": "full"},
             {"type": "spherical_chebyshev", "nmax": 3, "lmax": 2, "rcut": 3.5, "wtype": "none"}]
    md = M.random_merge_model(["Cu", "Zr", "Al"], parts, (8, 8), 3)
    rc, nt, P, S = _parse(json.dumps(md))
    assert rc == 0, nnap.lib().jse_nnap_last_error()
    mdl = _both_equal(md)
    assert S == [s.S for s in mdl.species] == [83, 83, 83] and P == [10, 10, 10]
    assert mdl.rcut_max == 5.5
    for bad, msg in (({"type": "merge"}, "Key `basis`"), ({"type": "merge", "basis": []}, "empty"),
                     ({"type": "merge", "basis": [{"type": "mirror", "mirror": 1}]}, "Unsupported basis type")):
        m2 = json.loads(json.dumps(md))
        m2["models"][1]["basis"] = bad
        assert _parse(json.dumps(m2))[0] != 0 and msg in nnap.lib().jse_nnap_last_error().decode()
        with pytest.raises(ValueError, match=msg):
            M.load_model(m2)


def test_coupling_terms_match_the_reference_tables():
    """The L3 / L4 coupling terms the descriptor kernels read, evaluated on the
    host, against values of the reference's own coefficient tables
    (basis_SphericalUtil0.hpp, tests/golden/make_l34_vectors.py)."""
    from conftest import GOLDEN
    g = np.load(os.path.join(GOLDEN, "l34_invariants.npz"))
    L = nnap.lib()
    for A, l3, l4 in zip(g["A"], g["l3"], g["l4"]):
        A = np.ascontiguousarray(A)
        out = np.zeros(32)
        dp = ctypes.POINTER(ctypes.c_double)
        assert L.jse_nnap_sh_invariants(6, 3, A.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
        want = np.concatenate([l3, l4])
        assert np.allclose(out, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
        small = np.zeros(12)
        A5 = np.ascontiguousarray(A[:25])
        assert L.jse_nnap_sh_invariants(4, 2, A5.ctypes.data_as(dp), small.ctypes.data_as(dp)) == 0
        assert np.allclose(small, np.concatenate([out[:9], out[23:26]]), rtol=1e-13, atol=1e-12)
    out = np.zeros(32)
    dp = ctypes.POINTER(ctypes.c_double)
    assert L.jse_nnap_sh_invariants(7, 0, out.ctypes.data_as(dp), out.ctypes.data_as(dp)) != 0


def test_cpp_parser_l34():
    """l3max / l4max (SphericalChebyshev.java:32-42): both loaders size the
    descriptor as P (lmax + 1 + L3 columns + L4 columns) and refuse values out
    of range with the reference's messages."""
    md = M.random_model(["Cu", "Zr"], 3, 2, 5.0, (8,), seed=2, l3max=5, l4max=2)
    rc, nt, P, S = _parse(json.dumps(md))
    assert rc == 0, nnap.lib().jse_nnap_last_error()
    mdl = _both_equal(md)
    assert S == [s.S for s in mdl.species] == [P[0] * (3 + 14 + 3)] * 2
    for key, val, msg in (("l3max", 7, "l3max MUST be in [0, 6]"), ("l4max", 4, "l4max MUST be in [0, 3]")):
        m2 = json.loads(json.dumps(md))
        m2["models"][0]["basis"][key] = val
        assert _parse(json.dumps(m2))[0] != 0 and msg in nnap.lib().jse_nnap_last_error().decode()
        with pytest.raises(ValueError, match=msg.split(" MUST")[0]):
            M.load_model(m2)


# ------------------------------------------------------------- YAML model files
def _jackson_style(obj, ind=0):
    """A YAML text in the layout of a Jackson / Groovy YamlBuilder writer: a
    '---' line, double-quoted strings, sequences at their key's column and
    nested sequences opened on the item's line ('- - - 0.1')."""
    pad = " " * ind
    if isinstance(obj, dict):
        out = []
        for k, v in obj.items():
            if isinstance(v, (dict, list)) and v:
                out.append(f"{pad}{k}:")
                out.append(_jackson_style(v, ind if isinstance(v, list) else ind + 2))
            else:
                out.append(f"{pad}{k}: {_jackson_scalar(v)}")
        return "\n".join(out)
    out = []
    for v in obj:
        if isinstance(v, (dict, list)) and v:
            inner = _jackson_style(v, ind + 2)
            out.append(pad + "- " + inner[ind + 2:])
        else:
            out.append(f"{pad}- {_jackson_scalar(v)}")
    return "\n".join(out)


def _jackson_scalar(v):
    if isinstance(v, str):
        return json.dumps(v)
    if isinstance(v, (list, dict)):
        return "[]" if isinstance(v, list) else "{}"
    if isinstance(v, bool):
        return "true" if v else "false"
    return repr(v)


def _yaml_texts(md):
    import yaml
    return {
        "block": yaml.safe_dump(md, default_flow_style=False, sort_keys=False),
        "flow": yaml.safe_dump(md, default_flow_style=True, width=60),   # flow collections over many lines
        "json": json.dumps(md, indent=2),                                 # JSON is YAML
        "jackson": "%YAML 1.2\n---\n# written in the Jackson layout\n" + _jackson_style(md) + "\n...\n",
    }


@pytest.mark.parametrize("style", ["block", "flow", "json", "jackson"])
def test_yaml_model_file_gives_the_json_blocks(tmp_path, style):
    """A model file ending in .yaml / .yml is read as YAML (NNAP(String),
    NNAP.java:407-409): the C++ loader builds bit-identical parameter blocks
    from four YAML layouts of the same model (merge bases with L3 invariants,
    a mirror species, FP64 weights written with 17 digits) as from its JSON."""
    parts = [{"type": "spherical_chebyshev", "nmax": 3, "lmax": 3, "l3max": 2, "rcut": 5.0, "wtype": "full"},
             {"type": "chebyshev", "nmax": 4, "rcut": 4.5}]
    md = M.random_merge_model(["Cu", "Zr", "Al"], parts, (8, 8), 9)
    rng = np.random.default_rng(4)
    for m in md["models"]:
        S = m["nn"]["input_dim"]
        m["norm_mu"] = rng.normal(size=S).tolist()
        m["norm_sigma"] = (1.0 + rng.random(S)).tolist()
        m["ref_eng"] = float(rng.normal())
    md["models"][1] = {"symbol": "Zr", "basis": {"type": "mirror", "mirror": 1}}
    mdl = _both_equal(md)
    text = _yaml_texts(md)[style]
    for ext in (".yaml", ".yml"):
        path = tmp_path / ("model" + ext)
        path.write_text(text)
        for t, sp in enumerate(mdl.species):
            fp, nn, norm, hid, mir = _blocks(str(path), t + 1, sp)
            assert np.array_equal(fp, sp.W.ravel()) and np.array_equal(nn, sp.nn_param)
            assert np.array_equal(norm, sp.norm_param) and hid == list(sp.hidden) and mir == sp.mirror
        assert _parse(str(path)) == _parse(json.dumps(md))


@pytest.mark.parametrize("text,what", [
    ("models: &a\n- symbol: Cu\n", "anchors"),
    ("models:\n\t- symbol: Cu\n", "tabs"),
    ("units: metal\nunits: real\n", "duplicate key"),
    ("units: |\n  metal\n", "block scalars"),
    ("units: met\n  al\n", "multi-line plain"),
    ("models:\n  - a: 1\n - b: 2\n", "indentation"),
    ("models: [1, 2\n", "unterminated flow"),
    ("w: .inf\n", "non-finite"),
    ("a: 1\n---\nb: 2\n", "more than one document"),
])
def test_yaml_outside_the_subset_names_the_line(tmp_path, text, what):
    path = tmp_path / "bad.yaml"
    path.write_text(text)
    rc, *_ = _parse(str(path))
    assert rc == 1
    msg = nnap.lib().jse_nnap_last_error().decode()
    assert "YAML model line" in msg and what in msg, msg


@pytest.mark.parametrize("ext,text", [(".yaml", "a: " + "[" * 5000 + "]" * 5000 + "\n"),
                                      (".yaml", "".join(" " * k + "- \n" for k in range(200))),
                                      (".json", "[" * 100000 + "]" * 100000)])
def test_deep_nesting_is_refused(tmp_path, ext, text):
    """A file nested past 64 levels is refused with status 1 (no recursion
    without bound in either parser)."""
    path = tmp_path / ("deep" + ext)
    path.write_text(text)
    rc, *_ = _parse(str(path))
    assert rc == 1
    assert "nesting deeper than 64" in nnap.lib().jse_nnap_last_error().decode()
