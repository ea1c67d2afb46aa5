"""ctypes front end of the FP64 oracle (nnap_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
MAX_LAYERS = 8

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)


class OrcModel(ctypes.Structure):
    _fields_ = [
        ("ntypes", ctypes.c_int),
        ("nmax", _c_int_p),
        ("lmax", _c_int_p),
        ("size_np", _c_int_p),
        ("rcut", _c_dbl_p),
        ("nlayers", _c_int_p),
        ("hidden", _c_int_p),
        ("fp_param", ctypes.POINTER(_c_dbl_p)),
        ("nn_param", ctypes.POINTER(_c_dbl_p)),
        ("norm_param", ctypes.POINTER(_c_dbl_p)),
        ("mirror", _c_int_p),
        ("radial", _c_int_p),
        ("part0", _c_int_p),
        ("nparts", _c_int_p),
        ("l3max", _c_int_p),
        ("l4max", _c_int_p),
        ("train_w", _c_int_p),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_LIBS = {}


def lib(fast: bool = False):
    name = "liborc_fast.so" if fast else "liborc.so"
    if name not in _LIBS:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built (make -C oracle)")
        L = ctypes.CDLL(path)
        L.orc_neighbors_celllist.restype = ctypes.c_int
        L.orc_neighbors_bruteforce.restype = ctypes.c_int
        L.orc_system_efv.restype = ctypes.c_int
        _LIBS[name] = L
    return _LIBS[name]


def _ip(a):
    return a.ctypes.data_as(_c_int_p)


def _dp(a):
    return a.ctypes.data_as(_c_dbl_p)


class Oracle:
    """Holds a model (jse_amd.model.Model) in the oracle's C layout."""

    def __init__(self, mdl, fast: bool = False):
        self.mdl = mdl
        self.L = lib(fast)
        nt = mdl.ntypes
        sp = mdl.species
        # bases in species order (a merge species owns several, in fp order)
        bl = [b for s in sp for b in s.bases()]
        self._part0 = np.cumsum([0] + [len(s.bases()) for s in sp])[:-1].astype(np.int32)
        self._nparts = np.array([len(s.bases()) for s in sp], dtype=np.int32)
        self._nmax = np.array([b.nmax for b in bl], dtype=np.int32)
        self._lmax = np.array([b.lmax for b in bl], dtype=np.int32)
        self._P = np.array([b.P for b in bl], dtype=np.int32)
        self._rc = np.array([b.rcut for b in bl], dtype=np.float64)
        self._nl = np.array([len(s.hidden) for s in sp], dtype=np.int32)
        hid = np.zeros((nt, MAX_LAYERS), dtype=np.int32)
        for t, s in enumerate(sp):
            hid[t, :len(s.hidden)] = s.hidden
        self._hid = hid
        self._W = [np.ascontiguousarray(b.W.reshape(-1), dtype=np.float64) for b in bl]
        self._nn = [np.ascontiguousarray(s.nn_param, dtype=np.float64) for s in sp]
        self._norm = [np.ascontiguousarray(s.norm_param, dtype=np.float64) for s in sp]
        self._Wp = (_c_dbl_p * len(bl))(*[_dp(a) for a in self._W])
        self._nnp = (_c_dbl_p * nt)(*[_dp(a) for a in self._nn])
        self._normp = (_c_dbl_p * nt)(*[_dp(a) for a in self._norm])
        self._mirror = np.array([getattr(s, "mirror", 0) for s in sp], dtype=np.int32)
        self._radial = np.array([1 if b.radial else 0 for b in bl], dtype=np.int32)
        self._l3 = np.array([getattr(b, "l3max", 0) for b in bl], dtype=np.int32)
        self._l4 = np.array([getattr(b, "l4max", 0) for b in bl], dtype=np.int32)
        # W is a parameter block for the fuse, rfuse and exfuse wtypes (WTypeBasis.parameterSize)
        self._train = np.array([1 if b.typed_wtype in (4, 5, 6) else 0 for b in bl], dtype=np.int32)
        self.c = OrcModel(nt, _ip(self._nmax), _ip(self._lmax), _ip(self._P), _dp(self._rc), _ip(self._nl),
                          _ip(self._hid), self._Wp, self._nnp, self._normp, _ip(self._mirror), _ip(self._radial),
                          _ip(self._part0), _ip(self._nparts), _ip(self._l3), _ip(self._l4),
                          _ip(self._train))

    # --- per-atom entry points (jse_nnap_calFp / calEnergy / calEnergyForce) ---
    def cal_fp(self, ctype, d, tj):
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
        n = len(d)
        dx, dy, dz = (np.ascontiguousarray(d[:, k]) for k in range(3))
        tj = np.ascontiguousarray(tj, dtype=np.int32)
        S = self.mdl.species[ctype - 1].S
        fp = np.zeros(S)
        rc = self.L.orc_cal_fp(ctypes.byref(self.c), ctype, _dp(dx), _dp(dy), _dp(dz), _ip(tj), n, _dp(fp))
        if rc:
            raise ValueError(f"oracle cal_fp code {rc}")
        return fp

    def cal_energy_force(self, ctype, d, tj):
        d = np.ascontiguousarray(d, dtype=np.float64).reshape(-1, 3)
        n = len(d)
        dx, dy, dz = (np.ascontiguousarray(d[:, k]) for k in range(3))
        tj = np.ascontiguousarray(tj, dtype=np.int32)
        e = np.zeros(1)
        g = np.zeros((3, max(n, 1)))
        rc = self.L.orc_cal_energy_force(ctypes.byref(self.c), ctype, _dp(dx), _dp(dy), _dp(dz), _ip(tj), n,
                                         _dp(e), _dp(g[0]), _dp(g[1]), _dp(g[2]))
        if rc:
            raise ValueError(f"oracle cal_energy_force code {rc}")
        return float(e[0]), g[:, :n].T.copy()

    def mlp(self, ctype, fp):
        fp = np.ascontiguousarray(fp, dtype=np.float64)
        e = np.zeros(1)
        g = np.zeros_like(fp)
        rc = self.L.orc_mlp(ctypes.byref(self.c), ctype, _dp(fp), _dp(e), _dp(g))
        if rc:
            raise ValueError(f"oracle mlp code {rc}")
        return float(e[0]), g

    # --- whole system ---
    def system_efv(self, st, cap: int = 256):
        n = st.natoms
        xyz = np.ascontiguousarray(st.xyz, dtype=np.float64)
        ty = np.ascontiguousarray(st.types, dtype=np.int32)
        box = np.ascontiguousarray(st.box, dtype=np.float64)
        E = np.zeros(n)
        F = np.zeros((n, 3))
        V = np.zeros((n, 9))
        cnt = np.zeros(n, dtype=np.int32)
        while True:
            rc = self.L.orc_system_efv(ctypes.byref(self.c), n, _dp(xyz), _ip(ty), _dp(box), cap, _dp(E), _dp(F),
                                       _dp(V), _ip(cnt))
            if rc == 2:
                cap *= 2
                continue
            if rc:
                raise ValueError(f"oracle system_efv code {rc}")
            return {"energies": E, "forces": F, "virials": V, "energy": E.sum(), "virial": V.sum(0),
                    "nei_count": cnt}


def system_fp(orc: "Oracle", st, cap: int = 256):
    """Descriptors of every atom (orc_system_fp: jse_nnap_calFp per atom over the
    cell-list neighbours), rows [natoms][Smax]."""
    n = st.natoms
    smax = max(s.S for s in orc.mdl.species)
    xyz = np.ascontiguousarray(st.xyz, dtype=np.float64)
    ty = np.ascontiguousarray(st.types, dtype=np.int32)
    box = np.ascontiguousarray(st.box, dtype=np.float64)
    out = np.zeros((n, smax))
    while True:
        rc = orc.L.orc_system_fp(ctypes.byref(orc.c), n, _dp(xyz), _ip(ty), _dp(box), cap, smax, _dp(out))
        if rc == 2:
            cap *= 2
            continue
        if rc:
            raise ValueError(f"oracle system_fp code {rc}")
        return out


def fp_param_grad(orc: "Oracle", st, w, cap: int = 256):
    """Sum over atoms of w[i] dE_i/dW for every basis whose mixing W is a
    parameter block (orc_system_fp_param_grad); one array per species, its
    bases' [ntypes][P][K] blocks concatenated (zeros for a basis without
    parameters)."""
    mdl = orc.mdl
    n = st.natoms
    xyz = np.ascontiguousarray(st.xyz, dtype=np.float64)
    ty = np.ascontiguousarray(st.types, dtype=np.int32)
    box = np.ascontiguousarray(st.box, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = [np.zeros(sum(mdl.ntypes * b.P * b.K for b in sp.bases())) for sp in mdl.species]
    ptrs = (_c_dbl_p * len(out))(*[_dp(a) for a in out])
    while True:
        rc = orc.L.orc_system_fp_param_grad(ctypes.byref(orc.c), n, _dp(xyz), _ip(ty), _dp(box), cap, _dp(w), ptrs)
        if rc == 2:
            cap *= 2
            continue
        if rc:
            raise ValueError(f"oracle fp_param_grad code {rc}")
        return out


def nn_param_grad(sp, fp, w):
    """The network part of jse_nnap_backwardEnergy summed over rows
    (nnap_interface.cpp:111-134; normedNnBackward with GRAD_PARAM,
    nnap_main.hpp:297-323; the layer backward of nn_FeedForward.hpp:131-261):
    d(sum_i w_i E_i)/d nn_param for one species, fp rows [n][S], in the packed
    layout [hidden W per layer [out][in] | output W | hidden b | output b].
    Plain numpy FP64."""
    S = sp.S
    fp = np.asarray(fp, dtype=np.float64)[:, :S]
    w = np.asarray(w, dtype=np.float64)
    nn = sp.nn_param
    norm = sp.norm_param
    mu, sg = norm[2:2 + S], norm[2 + S:2 + 2 * S]
    hidden = list(sp.hidden)
    Ws, off, cin = [], 0, S
    for h in hidden:
        Ws.append(nn[off:off + h * cin].reshape(h, cin))
        off += h * cin
        cin = h
    wo = nn[off:off + cin]
    off += cin
    bs = []
    for h in hidden:
        bs.append(nn[off:off + h])
        off += h
    acts, ders = [(fp - mu) / sg], []
    for W, b in zip(Ws, bs):
        z = acts[-1] @ W.T + b
        s = 1.0 / (1.0 + np.exp(-z))
        acts.append(z * s)
        ders.append(s * (1.0 + z * (1.0 - s)))
    dy = w * norm[1]                      # dL/dy = w_i sigma_E
    g_wo = acts[-1].T @ dy
    g_bo = dy.sum()
    dz = (dy[:, None] * wo[None, :]) * ders[-1]
    gW, gb = [None] * len(hidden), [None] * len(hidden)
    for l in range(len(hidden) - 1, -1, -1):
        gW[l] = dz.T @ acts[l]
        gb[l] = dz.sum(0)
        if l > 0:
            dz = (dz @ Ws[l]) * ders[l - 1]
    return np.concatenate([g.ravel() for g in gW] + [g_wo] + gb + [[g_bo]])


def wrap(xyz, box):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    out = np.zeros_like(xyz)
    lib().orc_wrap(len(xyz), _dp(xyz), _dp(np.ascontiguousarray(box, dtype=np.float64)), _dp(out))
    return out


def neighbors(xyz_wrapped, box, rmax, mode="celllist", cap=256):
    """Returns (count, key, d) with key = j*32 + shift_code; celllist keeps the
    reference's visiting order, bruteforce is sorted by key."""
    x = np.ascontiguousarray(xyz_wrapped, dtype=np.float64)
    n = len(x)
    box = np.ascontiguousarray(box, dtype=np.float64)
    fn = lib().orc_neighbors_celllist if mode == "celllist" else lib().orc_neighbors_bruteforce
    while True:
        cnt = np.zeros(n, dtype=np.int32)
        key = np.zeros((n, cap), dtype=np.int32)
        d = np.zeros((n, cap, 3))
        rc = fn(n, _dp(x), _dp(box), ctypes.c_double(rmax), cap, _ip(cnt), _ip(key), _dp(d))
        if rc:
            raise ValueError(f"oracle neighbours code {rc}")
        if n and cnt.max() > cap:
            cap = int(cnt.max()) + 16
            continue
        return cnt, key, d


def real_sph(lmax, d):
    M = (lmax + 1) ** 2
    y = np.zeros(M)
    yt = np.zeros(M)
    yp = np.zeros(M)
    L = lib()
    L.orc_real_sph(lmax, ctypes.c_double(d[0]), ctypes.c_double(d[1]), ctypes.c_double(d[2]), _dp(y), _dp(yt),
                   _dp(yp))
    return y, yt, yp


L3_COLS = (0, 0, 2, 4, 9, 14, 23)
L4_COLS = (0, 1, 3, 9)


def sh_invariants(l3max: int, l4max: int, A):
    """The L3 then L4 invariants of one radial channel's A row (index l^2+l+m)."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    out = np.zeros(L3_COLS[l3max] + L4_COLS[l4max])
    if lib().orc_sh_invariants(l3max, l4max, _dp(A), _dp(out)):
        raise ValueError("l3max / l4max out of range")
    return out


def sh_invariants_grad(l3max: int, l4max: int, A, g):
    """sum_k g[k] d invariant_k / dA for one A row."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    g = np.ascontiguousarray(g, dtype=np.float64)
    gA = np.zeros_like(A)
    if lib().orc_sh_invariants_grad(l3max, l4max, _dp(A), _dp(g), _dp(gA)):
        raise ValueError("l3max / l4max out of range")
    return gA


def system_efv_par(orc: "Oracle", st, nthreads: int, cap: int = 256):
    """Thread-parallel evaluation (liborc_fast.so) used as the CPU baseline."""
    n = st.natoms
    xyz = np.ascontiguousarray(st.xyz, dtype=np.float64)
    ty = np.ascontiguousarray(st.types, dtype=np.int32)
    box = np.ascontiguousarray(st.box, dtype=np.float64)
    E = np.zeros(n)
    F = np.zeros((n, 3))
    V = np.zeros((n, 9))
    rc = orc.L.orc_system_efv_par(ctypes.byref(orc.c), n, _dp(xyz), _ip(ty), _dp(box), cap, int(nthreads), _dp(E),
                                  _dp(F), _dp(V))
    if rc:
        raise ValueError(f"oracle system_efv_par code {rc}")
    return {"energies": E, "forces": F, "virials": V, "energy": E.sum(), "virial": V.sum(0)}
