"""GPU parity of the other descriptor shapes against the FP64 oracle
(through the C-ABI): dense radial mixing (a two-species model with wtype
"full", P = 2K, each species' channels in its own block) and lmax = 4 with
nmax = 5 (one species: no mixing; three species: the default mixing),
mirror / share bases, a post-fuse weight, the radial-only chebyshev basis,
shapes without a tiled kernel (the untiled kernels: lmax 0-12, nmax up to
20, any wtype), merge bases, L3 / L4 invariants, and a network without
matrix-core tiles (the vector MLP).
Plain path bar: 1e-6 eV/atom, every force component within 1e-5 eV/A,
1e-5 eV/atom virial."""
import numpy as np
import pytest

from jse_amd import model as M
from jse_amd import nnap
from jse_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _liquid(symbols, n, seed):
    st = W.ternary_liquid(n=n, seed=seed)
    types = ((st.types - 1) % len(symbols)) + 1
    return W.Structure(st.box, st.xyz, types.astype(np.int32))


def _check(O, md, st):
    mdl = M.load_model(md)
    with nnap.NNAP(md) as pot:
        res = pot.calEnergyForceVirials(nnap.AtomData(st.box, st.xyz, st.types))
    ref = O.Oracle(mdl).system_efv(st)
    n = st.natoms
    e = abs(res["energy"] - ref["energies"].sum()) / n
    df = res["forces"] - ref["forces"]
    f_rms = float(np.sqrt((df ** 2).mean()))
    f_max = float(np.abs(df).max())
    v = float(np.abs(res["virial"] - ref["virials"].sum(0)).max()) / n
    print(f"e {e:.2e} f_rms {f_rms:.2e} f_max {f_max:.2e} v {v:.2e}")
    assert e <= 1e-6 and f_max <= 1e-5 and v <= 1e-5


def test_dense_mixing_two_species_full(O):
    st = _liquid(["Cu", "Zr"], 1500, 21)
    md = M.random_model(["Cu", "Zr"], nmax=8, lmax=6, rcut=6.0, hidden=(64, 64), seed=31, wtype="full")
    for m in md["models"]:
        assert m["basis"]["wtype"] == "full"
    mdl = M.load_model(md)
    assert mdl.species[0].P == 18          # ntypes * (nmax + 1)
    md = W._normalized(md, st, 300, 31)
    _check(O, md, st)


@pytest.mark.parametrize("symbols", [["Cu"], ["Cu", "Zr", "Al"]])
def test_lmax4_nmax5(O, symbols):
    st = _liquid(symbols, 1500, 22)
    md = M.random_model(symbols, nmax=5, lmax=4, rcut=6.0, hidden=(32, 32), seed=32)
    md = W._normalized(md, st, 300, 32)
    _check(O, md, st)


def test_mirror_and_share_bases(O):
    """mirror / share bases (MirrorBasis.java, SharedBasis.java): species 1
    with its own rcut, species 2 mirroring it (neighbours of species 1 and 2
    swap roles for species-2 centres; its entry has no network or
    normalisation and uses species 1's, NNAP.java:281-288), species 3 sharing
    it with its own network (a shared_feed_forward over species 1's: three
    rows per layer of its own, SharedFeedForward.java:163-198) and
    normalisation."""
    st = _liquid(["Cu", "Zr", "Al"], 1500, 23)
    md = M.random_model(["Cu", "Zr", "Al"], nmax=8, lmax=6, rcut=6.0, hidden=(64, 64), seed=33)
    md["models"][0]["basis"]["rcut"] = 5.5
    md["models"][2]["basis"] = {"type": "share", "share": 1}
    md = W._normalized(md, st, 300, 33)
    md["models"][1] = {"symbol": "Zr", "basis": {"type": "mirror", "mirror": 1}}
    own = md["models"][2]["nn"]
    md["models"][2]["nn"] = {"type": "shared_feed_forward", "share": 1, "shared_hidden_dims": [61, 61],
                             "hidden_weights": [w[:3] for w in own["hidden_weights"]],
                             "hidden_biases": [b[:3] for b in own["hidden_biases"]],
                             "output_weight": own["output_weight"], "output_bias": own["output_bias"]}
    mdl = M.load_model(md)
    assert [s.rcut for s in mdl.species] == [5.5, 5.5, 5.5] and [s.mirror for s in mdl.species] == [0, 1, 0]
    assert np.array_equal(mdl.species[1].nn_param, mdl.species[0].nn_param)
    _check(O, md, st)


def test_post_fuse_basis(O):
    """An older file's post-fuse weight (18 x 18 over the default mixing):
    the C++ loader's rfuse weight, through the dense-mixing kernels, against
    the oracle built from the Python loader's."""
    st = _liquid(["Cu", "Zr", "Al"], 1500, 24)
    md = M.random_model(["Cu", "Zr", "Al"], nmax=8, lmax=6, rcut=6.0, hidden=(64, 64), seed=34)
    rng = np.random.default_rng(34)
    for m in md["models"]:
        m["basis"]["post_fuse_weight"] = (rng.normal(size=18 * 18) / 4).tolist()
        m["basis"]["post_fuse_scale"] = 0.75
    mdl = M.load_model(md)
    assert [s.P for s in mdl.species] == [18] * 3 and mdl.species[0].wtype == M.WTYPE_RFUSE
    md = W._normalized(md, st, 300, 34)
    _check(O, md, st)


@pytest.mark.parametrize("symbols,nmax", [(["Cu", "Zr", "Al"], 8), (["Cu"], 5)])
def test_radial_chebyshev_basis(O, symbols, nmax):
    """The radial-only `chebyshev` basis (Chebyshev.java): fp[p] = sum_j fc Rnp[p];
    three species under the default mixing, one without mixing."""
    st = _liquid(symbols, 1500, 25)
    md = M.random_model(symbols, nmax=nmax, lmax=0, rcut=6.0, hidden=(32, 32), seed=35)
    for m in md["models"]:
        b = m["basis"]
        m["basis"] = {"type": "chebyshev", "nmax": b["nmax"], "rcut": b["rcut"], "wtype": b["wtype"]}
    assert all(s.radial for s in M.load_model(md).species)
    md = W._normalized(md, st, 300, 35)
    _check(O, md, st)


@pytest.mark.parametrize("symbols,nmax,lmax,wtype,hidden", [
    (["Cu"], 4, 3, "none", (32, 32)),                  # P = 5, S = 20
    (["Cu", "Zr", "Al"], 5, 8, "exfull", (32, 32)),    # P = 24, S = 216 (the vector MLP)
    (["Cu", "Zr"], 20, 2, "full", (64, 64)),           # K = 21 (nmax at its limit), P = 42
    (["Cu"], 3, 12, "none", (32, 32)),                 # lmax at its limit, M = 169
    (["Cu", "Zr", "Al"], 7, 0, "default", (32, 32)),   # lmax = 0: fp = 4 pi A_00^2
])
def test_untiled_shapes(O, symbols, nmax, lmax, wtype, hidden):
    st = _liquid(symbols, 1000, 26)
    md = M.random_model(symbols, nmax=nmax, lmax=lmax, rcut=6.0, hidden=hidden, seed=36, wtype=wtype)
    md = W._normalized(md, st, 200, 36)
    _check(O, md, st)


def test_untiled_radial_shape(O):
    """Radial-only basis with wtype "exfull" over two species: K = 13, P = 39."""
    st = _liquid(["Cu", "Zr"], 1000, 27)
    md = M.random_model(["Cu", "Zr"], nmax=12, lmax=0, rcut=6.0, hidden=(32, 32), seed=37, wtype="exfull")
    for m in md["models"]:
        b = m["basis"]
        m["basis"] = {"type": "chebyshev", "nmax": b["nmax"], "rcut": b["rcut"], "wtype": b["wtype"]}
    assert [s.P for s in M.load_model(md).species] == [39, 39]
    md = W._normalized(md, st, 200, 37)
    _check(O, md, st)


@pytest.mark.parametrize("wtype,fuse_size", [("fuse", 3), ("exfuse", 2)])
def test_fuse_wtypes(O, wtype, fuse_size):
    """fuse / exfuse mixing (P = 3K here): the network is drawn for the full
    mixing of three species (also P = 3K), then the basis switches to the fuse
    weights."""
    st = _liquid(["Cu", "Zr", "Al"], 1000, 28)
    md = M.random_model(["Cu", "Zr", "Al"], nmax=6, lmax=5, rcut=6.0, hidden=(32, 32), seed=38, wtype="full")
    rng = np.random.default_rng(38)
    for m in md["models"]:
        m["basis"]["wtype"] = wtype
        m["basis"]["fuse_weight"] = rng.normal(size=3 * fuse_size).tolist()
    assert [s.P for s in M.load_model(md).species] == [21] * 3
    md = W._normalized(md, st, 200, 38)
    _check(O, md, st)


def test_network_without_matrix_core_tiles(O):
    """Three hidden layers of widths 48-24-16 (no matrix-core tiles): the FP64
    vector MLP of the plain mode, behind the tiled descriptor of the benchmark
    shape."""
    st = _liquid(["Cu", "Zr", "Al"], 1000, 29)
    md = M.random_model(["Cu", "Zr", "Al"], nmax=8, lmax=6, rcut=6.0, hidden=(48, 24, 16), seed=39)
    md = W._normalized(md, st, 200, 39)
    _check(O, md, st)


def test_bf16_mode_network_without_bf16_tiles(O):
    """The BF16-MLP mode with a network the BF16 tiles do not cover (three
    hidden layers): that species' MLP runs in FP32, inside the BF16 budget
    (1 meV/atom, 20 meV/A RMS)."""
    st = _liquid(["Cu", "Zr", "Al"], 1000, 30)
    md = M.random_model(["Cu", "Zr", "Al"], nmax=8, lmax=6, rcut=6.0, hidden=(48, 24, 16), seed=40)
    md = W._normalized(md, st, 200, 40)
    with nnap.NNAP(md, precision=nnap.BF16_MLP) as pot:
        res = pot.calEnergyForceVirials(nnap.AtomData(st.box, st.xyz, st.types))
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    e = abs(res["energy"] - ref["energies"].sum()) / st.natoms
    f_rms = float(np.sqrt(((res["forces"] - ref["forces"]) ** 2).mean()))
    print(f"e {e:.2e} f_rms {f_rms:.2e}")
    assert e <= 1e-3 and f_rms <= 2e-2


_MERGE_PARTS = [
    {"type": "spherical_chebyshev", "nmax": 8, "lmax": 6, "rcut": 6.0, "wtype": "default"},   # tiled kernels
    {"type": "chebyshev", "nmax": 5, "rcut": 4.5, "wtype": "default"},                        # radial kernels
    {"type": "spherical_chebyshev", "nmax": 3, "lmax": 2, "rcut": 3.8, "wtype": "none"},       # untiled kernels
]


def test_merge_basis(O):
    """merge (MergedBasis.java): a tiled spherical part, a radial part and an
    untiled spherical part with three cutoffs; each part writes its slice of
    the descriptor row and the later parts' backwards add into the pairs' g.
    Plain path at the full bar."""
    st = _liquid(["Cu", "Zr", "Al"], 1500, 27)
    md = M.random_merge_model(["Cu", "Zr", "Al"], _MERGE_PARTS, (64, 64), 37)
    assert M.load_model(md).species[0].S == 126 + 12 + 12
    md = W._normalized(md, st, 300, 37)
    _check(O, md, st)


def test_merge_basis_fp32_fast(O):
    """The same merge model in FP32_FAST mode, within that mode's bar
    (RMS 2e-5 eV/A, max 2e-4 eV/A)."""
    st = _liquid(["Cu", "Zr", "Al"], 1500, 27)
    md = W._normalized(M.random_merge_model(["Cu", "Zr", "Al"], _MERGE_PARTS, (64, 64), 37), st, 300, 37)
    with nnap.NNAP(md, precision=nnap.FP32_FAST) as pot:
        res = pot.calEnergyForceVirials(nnap.AtomData(st.box, st.xyz, st.types))
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    df = res["forces"] - ref["forces"]
    print(f"f_rms {np.sqrt((df ** 2).mean()):.2e} f_max {np.abs(df).max():.2e}")
    assert abs(res["energy"] - ref["energies"].sum()) / st.natoms <= 1e-6
    assert np.sqrt((df ** 2).mean()) <= 2e-5 and np.abs(df).max() <= 2e-4


def _normalized_by_oracle(O, md, st):
    """Normalisation from the oracle's descriptors of every atom (the L3 / L4
    columns have no numpy restatement)."""
    mdl = M.load_model(md)
    fp = O.system_fp(O.Oracle(mdl), st)
    return M.set_normalization(md, [fp[st.types == t + 1, :mdl.species[t].S] for t in range(mdl.ntypes)])


@pytest.mark.parametrize("symbols,nmax,lmax,l3max,l4max,wtype,hidden", [
    (["Cu", "Zr", "Al"], 5, 4, 4, 2, "default", (64, 64)),   # P = 12, S = 12 (5 + 9 + 3) = 204
    (["Cu"], 4, 2, 6, 3, "none", (32, 32)),                  # harmonics to l = 6 for L3, S = 5 (3 + 23 + 9)
    (["Cu", "Zr"], 3, 6, 2, 1, "full", (32, 32)),            # P = 8, S = 8 (7 + 2 + 1)
])
def test_l34_invariants(O, symbols, nmax, lmax, l3max, l4max, wtype, hidden):
    """l3max / l4max > 0 (basis_SphericalUtil.hpp:395-467): the untiled kernels
    form each channel's L3 / L4 invariants from their coupling terms in the
    forward and add their dE/dA in the backward.  Plain path at the full bar."""
    st = _liquid(symbols, 800, 41)
    md = M.random_model(symbols, nmax=nmax, lmax=lmax, rcut=6.0, hidden=hidden, seed=43, wtype=wtype,
                        l3max=l3max, l4max=l4max)
    _check(O, _normalized_by_oracle(O, md, st), st)


def test_l34_invariants_fp32_fast_and_merge(O):
    """An L3 / L4 basis as a part of a merge basis beside a radial part, in the
    plain mode (full bar) and in FP32_FAST (RMS 2e-5 eV/A, max 2e-4 eV/A)."""
    st = _liquid(["Cu", "Zr", "Al"], 800, 42)
    parts = [{"type": "spherical_chebyshev", "nmax": 4, "lmax": 3, "l3max": 4, "l4max": 3, "rcut": 6.0,
              "wtype": "default"},
             {"type": "chebyshev", "nmax": 6, "rcut": 4.5, "wtype": "full"}]
    md = M.random_merge_model(["Cu", "Zr", "Al"], parts, (64, 64), 44)
    assert M.load_model(md).species[0].S == 10 * (4 + 9 + 9) + 21
    md = _normalized_by_oracle(O, md, st)
    _check(O, md, st)
    with nnap.NNAP(md, precision=nnap.FP32_FAST) as pot:
        res = pot.calEnergyForceVirials(nnap.AtomData(st.box, st.xyz, st.types))
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    df = res["forces"] - ref["forces"]
    print(f"fast f_rms {np.sqrt((df ** 2).mean()):.2e} f_max {np.abs(df).max():.2e}")
    assert abs(res["energy"] - ref["energies"].sum()) / st.natoms <= 1e-6
    assert np.sqrt((df ** 2).mean()) <= 2e-5 and np.abs(df).max() <= 2e-4
