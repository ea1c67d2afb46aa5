"""GPU parity of the training derivatives (the rGradNnParam and rGradFpParam
parts of jse_nnap_backwardEnergy, nnap_interface.cpp:111-134, summed over a
system with per-atom weights) against the oracle's FP64 restatement (the
network part on the oracle's own descriptors).  Plain mode: the GPU descriptors are FP64, so the
gradients agree to ~1e-9 of their largest entry; FP32_FAST (FP32 descriptors
into an FP64 gradient): 1e-5."""
import numpy as np
import pytest

from jse_amd import model as M
from jse_amd import nnap
from jse_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _check(O, md, st, precision, tol):
    mdl = M.load_model(md)
    w = np.random.default_rng(5).normal(size=st.natoms)
    with nnap.NNAP(md, precision=precision) as pot:
        sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
        got = sysm.backward_energy(w)
        zero = sysm.backward_energy(np.zeros(st.natoms))
    fp = O.system_fp(O.Oracle(mdl), st)
    for t, sp in enumerate(mdl.species):
        rows = st.types == t + 1
        want = O.nn_param_grad(sp, fp[rows], w[rows])
        err = np.abs(got[t] - want).max() / np.abs(want).max()
        print(f"species {t + 1}: max |dg| / max |g| = {err:.2e}")
        assert err <= tol
        assert not zero[t].any()


def test_param_grad_ternary_plain(O):
    st = W.ternary_liquid(n=2000, seed=3)
    _check(O, W.ternary_model(), st, nnap.FP32, 1e-9)


def test_param_grad_ternary_fp32_fast(O):
    st = W.ternary_liquid(n=2000, seed=3)
    _check(O, W.ternary_model(), st, nnap.FP32_FAST, 1e-5)


def test_param_grad_three_layers_merge_basis(O):
    """A merge basis (spherical + radial parts) with a three-hidden-layer
    network of uneven widths."""
    st = W.ternary_liquid(n=1500, seed=9)
    parts = [{"type": "spherical_chebyshev", "nmax": 4, "lmax": 4, "rcut": 5.5, "wtype": "default"},
             {"type": "chebyshev", "nmax": 5, "rcut": 4.5, "wtype": "default"}]
    md = W._normalized(M.random_merge_model(["Cu", "Zr", "Al"], parts, (24, 16, 8), 91), st, 300, 91)
    _check(O, md, st, nnap.FP32, 1e-9)


def _trainable_mixing_model(symbols, parts, hidden, seed, st, O):
    """A merge (or single) basis whose parts use the rfuse mixing with random
    weights; the networks are drawn for the full mixing of the same channel
    counts.  Normalised by the oracle's descriptors."""
    nt = len(symbols)
    rng = np.random.default_rng(seed)
    full = [dict(p, wtype="full") if p.get("wtype") == "rfuse" else dict(p) for p in parts]
    md = M.random_merge_model(symbols, full, hidden, seed)
    real = []
    for p in parts:
        q = dict(p)
        if q.get("wtype") == "rfuse":
            K = q["nmax"] + 1
            q["rfuse_weight"] = rng.normal(size=nt * K * nt * K).tolist()   # P = ntypes K, as the full mixing
        real.append(q)
    for m in md["models"]:
        m["basis"]["basis"] = real
    mdl = M.load_model(md)
    fp = O.system_fp(O.Oracle(mdl), st)
    return M.set_normalization(md, [fp[st.types == t + 1, :mdl.species[t].S] for t in range(nt)])


@pytest.mark.parametrize("precision,tol", [(nnap.FP32, 1e-7), (nnap.FP32_FAST, 2e-6)])
def test_mixing_param_grad(O, precision, tol):
    """The rGradFpParam part of jse_nnap_backwardEnergy (rfuse mixing weights,
    basis_SphericalChebyshev.hpp:353-440, basis_Chebyshev.hpp:135-220) summed
    over a system: a merge of an rfuse spherical part with L3 invariants, a
    default-mixing part (no parameters: zeros) and an rfuse radial part,
    against the oracle: 5e-9 of the largest entry measured in the plain mode
    (dE/dA formed from the FP32 copies of G and A), 2e-7 in FP32_FAST."""
    st = W.ternary_liquid(n=1200, seed=17)
    parts = [{"type": "spherical_chebyshev", "nmax": 3, "lmax": 4, "l3max": 3, "rcut": 6.0, "wtype": "rfuse"},
             {"type": "spherical_chebyshev", "nmax": 2, "lmax": 2, "rcut": 4.5, "wtype": "default"},
             {"type": "chebyshev", "nmax": 5, "rcut": 5.0, "wtype": "rfuse"}]
    md = _trainable_mixing_model(["Cu", "Zr", "Al"], parts, (32, 32), 23, st, O)
    mdl = M.load_model(md)
    w = np.random.default_rng(6).uniform(0.5, 1.5, size=st.natoms)
    with nnap.NNAP(md, precision=precision) as pot:
        sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
        gnn, gfp = sysm.backward_energy(w, mixing=True)
    want = O.fp_param_grad(O.Oracle(mdl), st, w)
    sizes = [3 * b.P * b.K for b in mdl.species[0].bases()]
    for t in range(3):
        assert gfp[t].shape == want[t].shape == (sum(sizes),)
        off = 0
        for k, sz in enumerate(sizes):
            g, r = gfp[t][off:off + sz], want[t][off:off + sz]
            if k == 1:
                assert not g.any() and not r.any()
            else:
                err = np.abs(g - r).max() / np.abs(r).max()
                print(f"species {t + 1} part {k}: max |dg| / max |g| = {err:.2e}")
                assert err <= tol
            off += sz


# --------------------------------------------------------- force / stress loss
def _lin_loss(orc, st, w, bf, bv):
    """sum_i w_i E_i + sum_k bf_k . F_k + bv . virial on the oracle."""
    r = orc.system_efv(st)
    return w @ r["energies"] + float((bf * r["forces"]).sum()) + bv @ r["virial"]


def _fd_check(O, mdl, st, got_nn, got_fp, w, bf, bv, tol, n_nn=8, n_w=6, seed=0):
    """Central differences (h = 1e-5) of the oracle's linear loss by random
    network and mixing entries against the gradient blocks."""
    orc = O.Oracle(mdl)
    rng = np.random.default_rng(seed)
    h = 1e-5
    def fd(arr, idx):
        old = arr[idx]
        arr[idx] = old + h
        lp = _lin_loss(orc, st, w, bf, bv)
        arr[idx] = old - h
        lm = _lin_loss(orc, st, w, bf, bv)
        arr[idx] = old
        return (lp - lm) / (2 * h)
    worst = 0.0
    b0 = 0
    for t, sp in enumerate(mdl.species):
        scale = max(1.0, np.abs(got_nn[t]).max())
        for idx in rng.choice(got_nn[t].size, size=n_nn, replace=False):
            e = abs(fd(orc._nn[t], idx) - got_nn[t][idx]) / scale
            worst = max(worst, e)
        if got_fp is not None:
            off = 0
            for k, b in enumerate(sp.bases()):
                sz = mdl.ntypes * b.P * b.K
                blk = got_fp[t][off:off + sz]
                if b.typed_wtype in (4, 5, 6):
                    scale = max(1.0, np.abs(blk).max())
                    for idx in rng.choice(sz, size=n_w, replace=False):
                        e = abs(fd(orc._W[b0 + k], idx) - blk[idx]) / scale
                        worst = max(worst, e)
                else:
                    assert not blk.any()
                off += sz
        b0 += len(sp.bases())
    print(f"max |fd - g| / max |g| = {worst:.2e}")
    assert worst <= tol


def _upstream(st, seed):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0.5, 1.5, size=st.natoms), rng.normal(size=(st.natoms, 3)),
            np.array([0.3, -0.2, 0.1, 0.05, -0.07, 0.02, 0.0, 0.0, 0.0]))


@pytest.mark.parametrize("precision,tol", [(nnap.FP32, 1e-7), (nnap.FP32_FAST, 2e-6)])
def test_force_loss_grad_networks(O, precision, tol):
    """The force / stress loss' derivative by the networks (backwardEnergyForce's
    rBGradNnParam) on the ternary shapes, against central differences of the
    oracle's linear loss sum w E + bf . F + bv . virial (measured 4.5e-9 of the
    largest entry in the plain mode, 2.4e-7 in FP32_FAST)."""
    st = W.ternary_liquid(n=400, seed=19)
    md = M.random_model(["Cu", "Zr", "Al"], nmax=4, lmax=4, rcut=5.5, hidden=(16, 16), seed=29)
    md = W._normalized(md, st, 200, 29)
    mdl = M.load_model(md)
    w, bf, bv = _upstream(st, 1)
    with nnap.NNAP(md, precision=precision) as pot:
        sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
        gnn = sysm.backward_energy_force(w, bf, bv)
        # without force / virial terms it is the energy loss' derivative
        ge = sysm.backward_energy(w)
        ge2 = sysm.backward_energy_force(w)
    for a, b in zip(ge, ge2):
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(a).max())
    _fd_check(O, mdl, st, gnn, None, w, bf, bv, tol)


def test_force_loss_grad_mixing_merge_l34(O):
    """The same for every parameter block of a merge basis: an rfuse spherical
    part with L3 / L4 invariants, a default part (no parameters) and an rfuse
    radial part (backwardEnergyForce's rBGradFpParam: the descriptor's
    backward-backward, basis_SphericalChebyshev.hpp:527-656)."""
    st = W.ternary_liquid(n=400, seed=21)
    parts = [{"type": "spherical_chebyshev", "nmax": 3, "lmax": 3, "l3max": 3, "l4max": 2, "rcut": 5.5,
              "wtype": "rfuse"},
             {"type": "spherical_chebyshev", "nmax": 2, "lmax": 2, "rcut": 4.5, "wtype": "default"},
             {"type": "chebyshev", "nmax": 4, "rcut": 5.0, "wtype": "rfuse"}]
    md = _trainable_mixing_model(["Cu", "Zr", "Al"], parts, (16, 16), 31, st, O)
    mdl = M.load_model(md)
    w, bf, bv = _upstream(st, 2)
    with nnap.NNAP(md, precision=nnap.FP32) as pot:
        sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
        gnn, gfp = sysm.backward_energy_force(w, bf, bv, mixing=True)
        ge = sysm.backward_energy(w, mixing=True)
        ge2 = sysm.backward_energy_force(w, mixing=True)
    # the energy loss' mixing gradient forms dE/dA from the FP32 copies of G and
    # A, this one from FP64 adjoints: they agree to FP32 rounding
    for a, b in zip(ge[0], ge2[0]):
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(a).max())
    for a, b in zip(ge[1], ge2[1]):
        assert np.abs(a - b).max() <= 1e-6 * max(1.0, np.abs(a).max())
    _fd_check(O, mdl, st, gnn, gfp, w, bf, bv, 2e-7)   # measured 2.5e-8


def test_force_loss_grad_mirror_and_absent_species(O):
    """A mirror species (its neighbours' roles swapped, evaluated with its own
    copies of the mirrored species' network and rfuse mixing) and a species
    with no atoms in the structure (zero blocks): every block against central
    differences of the oracle's linear loss."""
    st0 = W.ternary_liquid(n=400, seed=23)
    st = W.Structure(st0.box, st0.xyz, np.where(st0.types == 3, 1, st0.types).astype(np.int32))  # no Al atoms
    parts = [{"type": "spherical_chebyshev", "nmax": 3, "lmax": 3, "rcut": 5.5, "wtype": "rfuse"}]
    md = _trainable_mixing_model(["Cu", "Zr", "Al"], parts, (12, 12), 37, st, O)
    md["models"][1] = {"symbol": "Zr", "basis": {"type": "mirror", "mirror": 1}}
    mdl = M.load_model(md)
    assert mdl.species[1].mirror == 1
    w, bf, bv = _upstream(st, 3)
    with nnap.NNAP(md, precision=nnap.FP32) as pot:
        sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
        gnn, gfp = sysm.backward_energy_force(w, bf, bv, mixing=True)
    assert not gnn[2].any() and not gfp[2].any()
    _fd_check(O, mdl, st, gnn, None, w, bf, bv, 2e-7, seed=4)
    # the mirror species evaluates with the mirrored species' mixing (the
    # reference adds its mixing gradient into that species' block,
    # nnap_main.hpp:345-346): one perturbed W moves both, so the difference
    # quotient is the sum of the two blocks
    orc = O.Oracle(mdl)
    assert np.shares_memory(orc._W[0], orc._W[1])
    tot = gfp[0] + gfp[1]
    scale = np.abs(tot).max()
    h = 1e-5
    for idx in np.random.default_rng(5).choice(tot.size, size=8, replace=False):
        old = orc._W[0][idx]
        orc._W[0][idx] = old + h
        lp = _lin_loss(orc, st, w, bf, bv)
        orc._W[0][idx] = old - h
        lm = _lin_loss(orc, st, w, bf, bv)
        orc._W[0][idx] = old
        assert abs((lp - lm) / (2 * h) - tot[idx]) <= 2e-7 * scale


def test_force_loss_grad_box_shorter_than_cutoff(O):
    """Training on a structure whose box is shorter than the cutoff (a DFT-sized
    cell: 6 atoms in 4.6 x 3.9 x 5.2 A, cutoff 5.5 A, evaluated as its
    supercell): the energy and force / stress loss' gradients against central
    differences of the oracle's linear loss (its expanded walk)."""
    box = np.array([4.6, 3.9, 5.2])
    st = W.Structure(box, W.rsa(6, box, 2.0, 47), np.array([1, 2, 3, 1, 2, 1], dtype=np.int32))
    parts = [{"type": "spherical_chebyshev", "nmax": 3, "lmax": 3, "rcut": 5.5, "wtype": "rfuse"}]
    big = W.ternary_liquid(n=400, seed=23)
    md = _trainable_mixing_model(["Cu", "Zr", "Al"], parts, (12, 12), 41, big, O)
    mdl = M.load_model(md)
    w, bf, bv = _upstream(st, 5)
    with nnap.NNAP(md, precision=nnap.FP32) as pot:
        sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
        gnn, gfp = sysm.backward_energy_force(w, bf, bv, mixing=True)
        ge = sysm.backward_energy(w, mixing=True)
    _fd_check(O, mdl, st, gnn, gfp, w, bf, bv, 2e-7, seed=6)
    _fd_check(O, mdl, st, ge[0], ge[1], w, np.zeros_like(bf), np.zeros(9), 2e-7, seed=7)


@pytest.mark.parametrize("case", ["default_lmax6", "rfuse_lmax6"])
def test_force_loss_grad_tiled_shapes(O, case):
    """The register-tiled shapes take the matrix-core tangent kernel: the
    ternary default mixing (lmax 6, nmax 8, P 18) for the networks, and a
    two-species rfuse mixing of the same shape (P 18, trainable) for every
    block, against central differences of the oracle's linear loss."""
    if case == "default_lmax6":
        st = W.ternary_liquid(n=300, seed=29)
        md = M.random_model(["Cu", "Zr", "Al"], nmax=8, lmax=6, rcut=5.5, hidden=(16, 16), seed=31)
        md = W._normalized(md, st, 150, 31)
        mixing = False
    else:
        st0 = W.ternary_liquid(n=300, seed=31)
        st = W.Structure(st0.box, st0.xyz, np.where(st0.types == 3, 1, st0.types).astype(np.int32))
        parts = [{"type": "spherical_chebyshev", "nmax": 8, "lmax": 6, "rcut": 5.5, "wtype": "rfuse"}]
        md = _trainable_mixing_model(["Cu", "Zr"], parts, (12, 12), 33, st, O)
        mixing = True
    mdl = M.load_model(md)
    assert all(b.lmax == 6 and b.K == 9 and b.P == 18 for s in mdl.species for b in s.bases())
    w, bf, bv = _upstream(st, 8)
    with nnap.NNAP(md, precision=nnap.FP32) as pot:
        sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
        got = sysm.backward_energy_force(w, bf, bv, mixing=mixing)
    gnn, gfp = got if mixing else (got, None)
    _fd_check(O, mdl, st, gnn, gfp, w, bf, bv, 2e-7, seed=9)


def test_batch_of_structures_sums_the_per_structure_gradients(O):
    """Many training structures in one call: a batch (jse_nnap_batch_create) of
    three liquids and a 6-atom cell below the cutoff, with per-atom w, dL/dF
    and one dL/dvirial per replica, gives the sum of the single-system
    gradients, for the energy and the force / stress loss, networks and rfuse
    mixing (to summation order)."""
    reps = [W.ternary_liquid(n=n, seed=s) for n, s in ((300, 41), (420, 42), (360, 43))]
    small = np.array([4.6, 3.9, 5.2])   # a box below the cutoff: held as its supercell in the batch too
    reps.append(W.Structure(small, W.rsa(6, small, 2.0, 44), np.array([1, 2, 3, 1, 2, 3], dtype=np.int32)))
    parts = [{"type": "spherical_chebyshev", "nmax": 3, "lmax": 4, "rcut": 5.5, "wtype": "rfuse"}]
    md = _trainable_mixing_model(["Cu", "Zr", "Al"], parts, (12, 12), 45, reps[0], O)
    rng = np.random.default_rng(11)
    ups = [(rng.uniform(0.5, 1.5, r.natoms), rng.normal(size=(r.natoms, 3)), rng.normal(size=9)) for r in reps]
    with nnap.NNAP(md, precision=nnap.FP32) as pot:
        singles_f, singles_e = [], []
        for r, (w, bf, bv) in zip(reps, ups):
            s1 = nnap.System(pot, r.natoms, r.types, r.box, nnap.FP32)
            s1.set_positions(r.xyz)
            singles_f.append(s1.backward_energy_force(w, bf, bv, mixing=True))
            singles_e.append(s1.backward_energy(w, mixing=True))
            s1.close()
        b = pot.batch([nnap.AtomData(r.box, r.xyz, r.types) for r in reps])
        W_ = np.concatenate([u[0] for u in ups])
        BF = np.concatenate([u[1] for u in ups])
        BV = np.stack([u[2] for u in ups])
        bat_f = b.backward_energy_force(W_, BF, BV, mixing=True)
        bat_e = b.backward_energy(W_, mixing=True)
        b.close()
    for bat, singles in ((bat_f, singles_f), (bat_e, singles_e)):
        for blk in (0, 1):
            for t in range(3):
                want = sum(s_[blk][t] for s_ in singles)
                err = np.abs(bat[blk][t] - want).max() / max(1.0, np.abs(want).max())
                assert err <= 1e-11, (blk, t, err)
