"""GPU parity of the HIP path against the FP64 oracle (and the committed
fixtures), called through the C-ABI.

Bars (stated per test):
* neighbour lists: bit-exact -- same counts, same sorted (j, shift) keys, and
  FP32 displacements equal to the oracle's FP64 displacement rounded once;
* plain path (JSE_NNAP_FP32): total energy within 1e-6 eV/atom, every force
  component within 1e-5 eV/A, virial within 1e-5 eV per atom (total / N) --
  the tolerances the plain path is specified to meet against the reference's
  FP64 results;
* FP32-geometry modes (FP32_FAST, BF16_MLP): their own stated bars;
* descriptors: relative 2e-5 of the largest component per atom.
"""
import json
import os

import numpy as np
import pytest

from conftest import DevBuf, GOLDEN
from jse_amd import model as M
from jse_amd import nnap
from jse_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _atoms(st):
    return nnap.AtomData(st.box, st.xyz, st.types)


def _golden(name):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    return json.loads(str(g["model"])), W.Structure(g["box"], g["xyz"], g["types"]), g


def _errors(res, ref, n):
    e_atom = abs(res["energy"] - ref["energies"].sum()) / n
    e_i = np.abs(res["energies"] - ref["energies"]).max()
    df = res["forces"] - ref["forces"]
    f_rms = float(np.sqrt((df ** 2).mean()))
    f_max = float(np.abs(df).max())
    v_tot = float(np.abs(res["virial"] - ref["virials"].sum(0)).max()) / n
    v_i = float(np.abs(res["virials"] - ref["virials"]).max())
    out = dict(e_atom=e_atom, e_i_max=e_i, f_rms=f_rms, f_max=f_max, v_atom=v_tot, v_i_max=v_i)
    print({k: f"{v:.3e}" for k, v in out.items()})
    return out


def _assert_fp32_bar(err):
    assert err["e_atom"] <= 1e-6
    assert err["f_max"] <= 1e-5
    assert err["v_atom"] <= 1e-5


# ------------------------------------------------------------------ neighbours
def _nl_cases():
    return ["cu256", "glass", "ternary2000", "single_cell", "two_cells", "outside_box"]


def _nl_structure(case):
    if case == "cu256":
        return W.cu_fcc(seed=1)
    if case == "glass":
        return _golden("golden_glass")[1]
    if case == "ternary2000":
        return W.ternary_liquid(n=2000, seed=3)
    if case == "single_cell":
        return W.Structure(np.array([9.95] * 3), W.rsa(64, [9.95] * 3, 2.0, 11), np.ones(64, np.int32))
    if case == "two_cells":
        return W.Structure(np.array([12.4, 13.0, 12.7]), W.rsa(110, [12.4, 13.0, 12.7], 2.0, 12),
                           np.ones(110, np.int32))
    st = W.ternary_liquid(n=500, seed=4)   # positions far outside [0, L): wrap on the device
    k = np.random.default_rng(0).integers(-3, 4, size=st.xyz.shape)
    return W.Structure(st.box, st.xyz + k * st.box, st.types)


@pytest.mark.parametrize("case", _nl_cases())
def test_neighbor_list_bit_exact(O, case):
    st = _nl_structure(case)
    mdl_json = W.cu_model() if st.types.max() == 1 else W.ternary_model()
    if case == "glass":
        mdl_json = _golden("golden_glass")[0]
    pot = nnap.NNAP(mdl_json)
    sysm = pot.system(_atoms(st))
    sysm.compute(nnap.ENERGY)
    cnt, key, d = sysm.neighbors()
    x = O.wrap(st.xyz, st.box)
    c2, k2, d2 = O.neighbors(x, st.box, pot.rcutMax(), "bruteforce")
    assert np.array_equal(cnt, c2)
    for i in range(st.natoms):
        n = cnt[i]
        assert np.array_equal(key[i, :n], k2[i, :n]), f"atom {i}"
        assert np.array_equal(d[i, :n], d2[i, :n].astype(np.float32)), f"atom {i}"
    stats = sysm.nl_stats()
    assert stats["unmatched"] == 0
    pot.close()


# ----------------------------------------------------------------- descriptors
@pytest.mark.parametrize("name", ["golden_cu256", "golden_glass"])
def test_descriptors_match_oracle(O, name):
    md, st, _ = _golden(name)
    mdl = M.load_model(md)
    pot = nnap.NNAP(md)
    fp = pot.calFp(_atoms(st))
    orc = O.Oracle(mdl)
    x = O.wrap(st.xyz, st.box)
    cnt, key, d = O.neighbors(x, st.box, mdl.rcut_max, "bruteforce")
    for i in range(0, st.natoms, max(1, st.natoms // 200)):
        S = mdl.species[st.types[i] - 1].S
        want = orc.cal_fp(int(st.types[i]), d[i, :cnt[i]], st.types[key[i, :cnt[i]] >> 5])
        assert np.abs(fp[i, :S] - want).max() <= 2e-5 * np.abs(want).max(), f"atom {i}"
    pot.close()


# -------------------------------------------------------------- energy / force
def test_efv_glass_against_golden():
    md, st, g = _golden("golden_glass")
    pot = nnap.NNAP(md)
    res = pot.calEnergyForceVirials(_atoms(st))
    err = _errors(res, {"energies": g["energies"], "forces": g["forces"], "virials": g["virials"]}, st.natoms)
    _assert_fp32_bar(err)
    pot.close()


def test_efv_cu_crystal_against_golden():
    """Cu FCC 256 with a model normalised on that same near-perfect crystal:
    sigma/mu of the descriptors is as small as 0.03, so (fp - mu)/sigma
    multiplies any relative error of fp by ~30.  The plain path forms the
    geometry, harmonics and A in FP64, so the full bar holds here too."""
    md, st, g = _golden("golden_cu256")
    pot = nnap.NNAP(md)
    res = pot.calEnergyForceVirials(_atoms(st))
    err = _errors(res, {"energies": g["energies"], "forces": g["forces"], "virials": g["virials"]}, st.natoms)
    _assert_fp32_bar(err)
    pot.close()


def test_efv_c2_full_size_against_oracle(O):
    """The benchmark system itself: the 32,000-atom Cu-Zr-Al liquid with the
    nmax 8 / lmax 6 / 64-64 ternary model, plain path against the FP64 oracle
    (16 host threads)."""
    st = W.ternary_liquid(n=32000, seed=7)
    md = W.ternary_model()
    pot = nnap.NNAP(md)
    res = pot.calEnergyForceVirials(_atoms(st))
    ref = O.system_efv_par(O.Oracle(M.load_model(md), fast=False), st, 16)
    _assert_fp32_bar(_errors(res, ref, st.natoms))
    pot.close()


def test_fp32_fast_mode_bar(O):
    """FP32_FAST: geometry, harmonics and A in FP32 with an FP32 MLP.  Its
    stated bar on the liquid recipe: 1e-6 eV/atom, forces 2e-5 eV/A RMS and
    2e-4 eV/A max, virial 1e-5 eV/atom."""
    st = W.ternary_liquid(n=2000, seed=3)
    md = W.ternary_model()
    pot = nnap.NNAP(md, precision=nnap.FP32_FAST)
    res = pot.calEnergyForceVirials(_atoms(st))
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    err = _errors(res, ref, st.natoms)
    assert err["e_atom"] <= 1e-6 and err["f_rms"] <= 2e-5 and err["f_max"] <= 2e-4 and err["v_atom"] <= 1e-5
    pot.close()


def test_pair_gradients_match_per_atom_oracle(O):
    """g_j = -dE_i/d(r_j - r_i) per pair, the output of the reference's
    jse_nnap_calEnergyForce, against the oracle's restatement of that call."""
    st = W.ternary_liquid(n=2000, seed=3)
    md = W.ternary_model()
    mdl = M.load_model(md)
    pot = nnap.NNAP(md)
    sysm = pot.system(_atoms(st))
    sysm.compute(nnap.ENERGY | nnap.FORCE)
    cnt, key, d = sysm.neighbors()
    g = sysm.pair_gradients(key.shape[1])
    E = sysm.results(forces=False, virials=False)["energies"]
    orc = O.Oracle(mdl)
    x = O.wrap(st.xyz, st.box)
    c2, k2, d2 = O.neighbors(x, st.box, mdl.rcut_max, "bruteforce")
    worst = 0.0
    worst_e = 0.0
    for i in range(0, st.natoms, 7):
        e, gr = orc.cal_energy_force(int(st.types[i]), d2[i, :c2[i]], st.types[k2[i, :c2[i]] >> 5])
        worst_e = max(worst_e, abs(E[i] - e))
        worst = max(worst, float(np.abs(g[i, :cnt[i]] - gr).max()))
    print("max |dE_i| =", worst_e, "max |dg| =", worst)
    assert worst_e <= 5e-6 and worst <= 1e-5
    pot.close()


def test_bf16_mlp_error_budget(O):
    """BF16 operands / FP32 accumulation in the MLP: energy within 1 meV/atom
    and force RMSE within 20 meV/A of the FP64 oracle (the precision variant's
    stated budget), on the liquid recipe."""
    st = W.ternary_liquid(n=4000, seed=5)
    md = W.ternary_model()
    pot = nnap.NNAP(md, precision=nnap.BF16_MLP)
    res = pot.calEnergyForceVirials(_atoms(st))
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    err = _errors(res, ref, st.natoms)
    e_rmse = float(np.sqrt(((res["energies"] - ref["energies"]) ** 2).mean()))
    print("per-atom energy RMSE", e_rmse)
    assert err["e_atom"] <= 1e-3 and e_rmse <= 1e-3
    assert err["f_rms"] <= 2e-2
    pot.close()


def test_efv_ternary_against_oracle(O):
    st = W.ternary_liquid(n=2000, seed=3)
    md = W.ternary_model()
    pot = nnap.NNAP(md)
    res = pot.calEnergyForceVirials(_atoms(st))
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    _assert_fp32_bar(_errors(res, ref, st.natoms))
    pot.close()


def test_small_box_multiple_images(O):
    st = _nl_structure("single_cell")
    md = W.ternary_model()
    pot = nnap.NNAP(md)
    res = pot.calEnergyForceVirials(_atoms(st))
    ref = O.Oracle(M.load_model(md)).system_efv(st)
    _assert_fp32_bar(_errors(res, ref, st.natoms))
    pot.close()


# ------------------------------------------------ boxes shorter than the cutoff
def _short_box(case):
    """Boxes with an edge below the 6 A cutoff (the reference's ExpandLinkedCell
    case): every edge short, a slab, and 2 atoms in a 2.9 A cube (three copies
    per axis) with positions several boxes outside [0, L)."""
    box, n = {"all_short": ((4.6, 3.9, 5.2), 5), "slab": ((13.0, 12.2, 3.4), 30),
              "tiny_outside": ((2.9, 2.9, 2.9), 2)}[case]
    box = np.array(box)
    types = np.random.default_rng(43).integers(1, 4, size=n).astype(np.int32)
    xyz = W.rsa(n, box, 2.0, 43)
    if case == "tiny_outside":
        xyz = xyz + np.array([[-2, 1, 3], [4, -1, 0]]) * box
    return W.Structure(box, xyz, types)


@pytest.mark.parametrize("case", ["all_short", "slab", "tiny_outside"])
def test_box_shorter_than_cutoff(O, case):
    """The system evaluates such a box as its supercell (every edge >= rc); the
    caller's atoms get the oracle's energies, forces and virials (its expanded
    walk, NeighborListGetter.java:704-776) to the FP32 bar, totals per box.
    Descriptors and device positions take the same mapping; neighbour rows
    (which name supercell copies) are refused with status 2."""
    st = _short_box(case)
    md = W.ternary_model()
    mdl = M.load_model(md)
    orc = O.Oracle(mdl)
    with nnap.NNAP(md) as pot:
        res = pot.calEnergyForceVirials(_atoms(st))
        fp = pot.calFp(_atoms(st))
        s = pot.system(_atoms(st))
        with pytest.raises(nnap.NNAPError) as e:
            s.neighbors()
        assert e.value.code == 2
        x = DevBuf(np.ascontiguousarray(st.xyz, dtype=np.float64))
        s.set_positions_device(x.ptr)
        s.compute()
        dev = s.results()
    ref = orc.system_efv(st)
    _assert_fp32_bar(_errors(res, ref, st.natoms))
    for k in ("energies", "forces", "virials"):
        assert np.array_equal(dev[k], res[k])
    want = O.system_fp(orc, st)
    assert np.abs(fp - want).max() <= 1e-9 * max(1.0, np.abs(want).max())


def test_box_crossing_the_cutoff_gets_a_new_system(O):
    """The same atoms in a box above the cutoff, below it, and above again: the
    cached system is replaced when the box needs another supercell."""
    st = _short_box("all_short")
    big = W.Structure(st.box * 2.0, st.xyz * 2.0, st.types)
    md = W.ternary_model()
    orc = O.Oracle(M.load_model(md))
    with nnap.NNAP(md) as pot:
        for s_ in (big, st, big):
            _assert_fp32_bar(_errors(pot.calEnergyForceVirials(_atoms(s_)), orc.system_efv(s_), s_.natoms))


def test_batch_with_boxes_shorter_than_cutoff(O):
    """A batch whose replicas include boxes below the cutoff (each held as its
    own supercell) next to an ordinary liquid: every replica's atoms and totals
    against the oracle, and device positions equal to host positions."""
    reps = [_short_box("all_short"), W.ternary_liquid(n=400, seed=9), _short_box("slab"),
            _short_box("tiny_outside")]
    md = W.ternary_model()
    orc = O.Oracle(M.load_model(md))
    with nnap.NNAP(md) as pot:
        b = pot.batch([_atoms(r) for r in reps])
        b.compute()
        res = b.results()
        e_rep, v_rep = b.replica_results()
        x = DevBuf(np.ascontiguousarray(np.concatenate([r.xyz for r in reps]), dtype=np.float64))
        b.set_positions_device(x.ptr)
        b.compute()
        dev = b.results()
        b.close()
    off = 0
    for q, r in enumerate(reps):
        sl = slice(off, off + r.natoms)
        off += r.natoms
        one = {"energy": e_rep[q], "energies": res["energies"][sl], "forces": res["forces"][sl],
               "virial": v_rep[q], "virials": res["virials"][sl]}
        _assert_fp32_bar(_errors(one, orc.system_efv(r), r.natoms))
    for k in ("energies", "forces", "virials"):
        assert np.array_equal(dev[k], res[k])
    assert abs(res["energy"] - e_rep.sum()) <= 1e-9 * max(1.0, abs(res["energy"]))


def test_api_variants_agree():
    md, st, g = _golden("golden_cu256")
    pot = nnap.NNAP(md)
    a = _atoms(st)
    full = pot.calEnergyForceVirials(a)
    assert abs(pot.calEnergy(a) - full["energy"]) < 1e-9
    assert np.allclose(pot.calEnergies(a), full["energies"], atol=1e-9, rtol=0)
    assert np.allclose(pot.calForces(a), full["forces"], atol=1e-9, rtol=0)
    stress = pot.calStress(a)
    assert np.allclose(stress, -full["virial"][:6] / a.volume(), rtol=1e-12, atol=1e-15)
    assert np.allclose(pot.calStresses(a), -full["virials"], rtol=0, atol=0)
    pot.close()


def test_deterministic_repeat():
    md, st, _ = _golden("golden_glass")
    pot = nnap.NNAP(md)
    a = _atoms(st)
    r1 = pot.calEnergyForceVirials(a)
    r2 = pot.calEnergyForceVirials(a)
    for k in ("energies", "forces", "virials"):
        assert np.array_equal(r1[k], r2[k])
    assert r1["energy"] == r2["energy"]
    pot.close()


def test_edge_cases():
    md = W.cu_model()
    pot = nnap.NNAP(md)
    # no atoms
    e = pot.calEnergyForceVirials(nnap.AtomData(np.array([20.0] * 3), np.zeros((0, 3)), np.zeros(0, np.int32)))
    assert e["energy"] == 0.0 and e["forces"].shape == (0, 3)
    # one atom in a large box: no neighbours, E = E(fp = 0), no force
    one = pot.calEnergyForceVirials(nnap.AtomData(np.array([20.0] * 3), np.array([[1.0, 2.0, 3.0]]),
                                                  np.array([1], np.int32)))
    assert np.all(one["forces"] == 0.0)
    # atom type beyond ntypes: status 5 (reference: IllegalArgumentException "Exist type ...")
    with pytest.raises(ValueError):
        pot.calEnergy(nnap.AtomData(np.array([20.0] * 3), np.array([[1.0, 2.0, 3.0]]), np.array([2], np.int32)))
    # non-finite position
    with pytest.raises(nnap.NNAPError) as ei:
        pot.calEnergy(nnap.AtomData(np.array([20.0] * 3), np.array([[np.nan, 2.0, 3.0]]), np.array([1], np.int32)))
    assert ei.value.code == 1
    pot.close()
