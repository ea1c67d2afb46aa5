"""Synthetic inputs for the NNAP evaluation path, with the sizes the
benchmarks and tests are quoted on:

* ``cu_fcc``          Cu FCC 4x4x4 conventional cells (256 atoms), a = 3.61 A,
                      Gaussian displacements sigma = 0.1 A from a fixed seed.
* ``ternary_liquid``  Cu:Zr:Al 50:40:10 at 0.058 A^-3, uniform random
                      sequential addition rejecting pairs closer than 2.0 A
                      (32,000 or 256,000 atoms).
* ``replicas``        independent periodic replicas, sizes log-uniform in a
                      range, 1-3 species drawn per replica, densities
                      uniform in [0.05, 0.065] A^-3, 2.0 A minimum separation.

Models: ``cu_model`` (1 species, nmax=5, lmax=6, rcut=6, MLP 32-32) and
``ternary_model`` (3 species, nmax=8, lmax=6, wtype default, MLP 64-64), both
seeded random-init with normalisation fixed from descriptor statistics.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import model as _model

_HERE = os.path.dirname(os.path.abspath(__file__))
_WL = None


def _wl():
    global _WL
    if _WL is None:
        path = os.path.join(_HERE, "libjse_workload.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        lib.jse_workload_rsa.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                         ctypes.c_uint64, ctypes.c_long, ctypes.POINTER(ctypes.c_double)]
        lib.jse_workload_rsa.restype = ctypes.c_int
        _WL = lib
    return _WL


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Structure:
    """An atomic configuration: orthogonal box [0,L) per axis, Cartesian
    positions (N x 3, FP64) and 1-based model types."""
    box: np.ndarray
    xyz: np.ndarray
    types: np.ndarray

    @property
    def natoms(self) -> int:
        return int(self.xyz.shape[0])


def rsa(n: int, box, dmin: float, seed: int) -> np.ndarray:
    box = np.ascontiguousarray(box, dtype=np.float64)
    xyz = np.zeros((n, 3), dtype=np.float64)
    placed = _wl().jse_workload_rsa(n, _dptr(box), float(dmin), int(seed) & 0xFFFFFFFFFFFFFFFF,
                                    int(2000 * n + 100000), _dptr(xyz))
    if placed != n:
        raise RuntimeError(f"random sequential addition placed {placed} of {n} atoms")
    return xyz


def cu_fcc(seed: int = 1, ncell: int = 4, a: float = 3.61, sigma: float = 0.1) -> Structure:
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    pts = []
    for k in range(ncell):
        for j in range(ncell):
            for i in range(ncell):
                pts.append((basis + np.array([i, j, k])) * a)
    xyz = np.concatenate(pts, axis=0)
    rng = np.random.Generator(np.random.PCG64(seed))
    if sigma > 0:
        xyz = xyz + rng.normal(0.0, sigma, size=xyz.shape)
    L = ncell * a
    return Structure(np.array([L, L, L]), xyz, np.ones(len(xyz), dtype=np.int32))


def ternary_liquid(n: int = 32000, rho: float = 0.058, dmin: float = 2.0, seed: int = 7,
                   fractions=(0.5, 0.4, 0.1)) -> Structure:
    L = (n / rho) ** (1.0 / 3.0)
    box = np.array([L, L, L])
    xyz = rsa(n, box, dmin, seed)
    counts = [int(round(f * n)) for f in fractions]
    counts[0] += n - sum(counts)
    types = np.concatenate([np.full(c, t + 1, dtype=np.int32) for t, c in enumerate(counts)])
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    rng.shuffle(types)
    return Structure(box, xyz, types)


def replicas(count: int, nmin: int, nmax: int, seed: int, rho_range=(0.05, 0.065), dmin: float = 2.0,
             ntypes: int = 3, workers: int = 1) -> List[Structure]:
    """`count` independent replicas: N log-uniform in [nmin, nmax], density
    uniform in `rho_range`, 1..ntypes species drawn per replica.  The random
    draws are sequential; the per-replica placements are independent (own
    seed), so `workers` > 1 only runs them concurrently."""
    rng = np.random.Generator(np.random.PCG64(seed))
    plan = []
    for r in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        rho = float(rng.uniform(*rho_range))
        k = int(rng.integers(1, ntypes + 1))
        chosen = np.sort(rng.choice(np.arange(1, ntypes + 1), size=k, replace=False))
        L = (n / rho) ** (1.0 / 3.0)
        box = np.array([L, L, L])
        types = chosen[rng.integers(0, k, size=n)].astype(np.int32)
        plan.append((n, box, seed * 100003 + r, types))

    def place(p):
        n, box, sd, types = p
        return Structure(box, rsa(n, box, dmin, sd), types)

    if workers > 1 and count > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(max_workers=workers) as ex:
            return list(ex.map(place, plan))
    return [place(p) for p in plan]


# ---------------------------------------------------------------------------
# models
# ---------------------------------------------------------------------------

def _real_sph_np(lmax, dx, dy, dz):
    """Real orthonormal Y_lm at index l^2+l+m (m>0 ~ cos, m<0 ~ sin), from
    scipy's complex harmonics.  Used only to pick normalisation constants."""
    from scipy.special import sph_harm_y
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    th = np.arccos(np.clip(dz / r, -1.0, 1.0))
    ph = np.arctan2(dy, dx)
    M = (lmax + 1) ** 2
    Y = np.zeros((len(dx), M))
    for l in range(lmax + 1):
        Y[:, l * l + l] = np.real(sph_harm_y(l, 0, th, ph))
        for m in range(1, l + 1):
            c = sph_harm_y(l, m, th, ph)
            Y[:, l * l + l + m] = math.sqrt(2.0) * np.real(c)
            Y[:, l * l + l - m] = math.sqrt(2.0) * np.imag(c)
    return Y


def descriptors_np(mdl: _model.Model, st: Structure, atoms: Optional[np.ndarray] = None) -> List[np.ndarray]:
    """Plain numpy FP64 descriptors of selected atoms (minimum image over 27
    shifts); returns one array per species (rows = atoms of that species)."""
    box = st.box
    x = st.xyz - np.floor(st.xyz / box) * box
    if atoms is None:
        atoms = np.arange(st.natoms)
    out = [[] for _ in range(mdl.ntypes)]
    shifts = np.array([[sx, sy, sz] for sz in (-1, 0, 1) for sy in (-1, 0, 1) for sx in (-1, 0, 1)], float)
    for i in atoms:
        sp = mdl.species[st.types[i] - 1]
        d0 = ((x[None, :, :] + shifts[:, None, :] * box) - x[i]).reshape(-1, 3)
        jidx = np.tile(np.arange(st.natoms), 27)
        r20 = (d0 * d0).sum(1)
        parts = []
        for b in sp.bases():  # a merge basis: the parts' fp concatenated
            keep = (r20 < b.rcut ** 2) & (r20 > 0)
            d = d0[keep]
            tj = st.types[jidx[keep]]
            r = np.sqrt((d * d).sum(1))
            xx = 1.0 - 2.0 * r / b.rcut
            Rn = np.zeros((len(r), b.K))
            Rn[:, 0] = 1.0
            if b.K > 1:
                Rn[:, 1] = xx
            for q in range(2, b.K):
                Rn[:, q] = 2 * xx * Rn[:, q - 1] - Rn[:, q - 2]
            fc = (1.0 - (r / b.rcut) ** 2) ** 4
            c = np.einsum("jpk,jk->jp", b.W[tj - 1], Rn) * fc[:, None]
            if b.radial:
                parts.append(c.sum(0))
                continue
            if b.l3max or b.l4max:
                raise ValueError("descriptors_np evaluates the L2 power spectrum only (l3max = l4max = 0)")
            Y = _real_sph_np(b.lmax, d[:, 0], d[:, 1], d[:, 2])
            A = c.T @ Y
            fp = np.zeros((b.P, b.lmax + 1))
            for l in range(b.lmax + 1):
                fp[:, l] = 4 * math.pi / (2 * l + 1) * (A[:, l * l:(l + 1) * (l + 1)] ** 2).sum(1)
            parts.append(fp.reshape(-1))
        out[st.types[i] - 1].append(np.concatenate(parts))
    return [np.array(v) if v else None for v in out]


def _normalized(model_dict: dict, st: Structure, nsample: int, seed: int) -> dict:
    mdl = _model.load_model(model_dict)
    rng = np.random.Generator(np.random.PCG64(seed))
    atoms = np.sort(rng.choice(st.natoms, size=min(nsample, st.natoms), replace=False))
    fps = descriptors_np(mdl, st, atoms)
    return _model.set_normalization(model_dict, fps)


def cu_model(seed: int = 11) -> dict:
    m = _model.random_model(["Cu"], nmax=5, lmax=6, rcut=6.0, hidden=(32, 32), seed=seed)
    return _normalized(m, cu_fcc(seed=1), 256, seed)


def ternary_model(seed: int = 13) -> dict:
    m = _model.random_model(["Cu", "Zr", "Al"], nmax=8, lmax=6, rcut=6.0, hidden=(64, 64), seed=seed)
    return _normalized(m, ternary_liquid(n=2000, seed=seed + 1), 600, seed)
