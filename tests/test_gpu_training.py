"""GPU parity of the training derivative of the networks (the rGradNnParam
part of jse_nnap_backwardEnergy, nnap_interface.cpp:111-134, summed over a
system with per-atom weights) against the oracle's FP64 restatement on the
oracle's own descriptors.  Plain mode: the GPU descriptors are FP64, so the
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
