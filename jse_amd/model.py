"""NNAP model files: load the reference's JSON/YAML format into the packed
per-species parameter blocks, and write seeded random-init models.

The packing follows the reference exactly (paths relative to the reference
repository root):

* ``fp_param[t]``   internal radial mixing ``W[ntypes][P][K]`` built from the
  basis ``wtype`` as in ``src/nnap/java/jsex/nnap/basis/WTypeBasis.java:202-304``
  (sizes ``:331-354``; a one-species model collapses full/exfull/none/default
  to ``none``, ``:89-101``).
* ``nn_param[t]``   ``[hidden W (per layer, [out][in]) | output W | hidden b |
  output b]`` (``nn/FeedForward.java:131-221,310-316``).
* ``norm_param[t]`` ``[mu_E + ref_E, sigma_E, mu[S], sigma[S]]`` with
  ``norm_mu_eng``/``norm_sigma_eng`` taken from the first species that has
  them (``NNAP.java:272-306``).

Bases are ``spherical_chebyshev`` with ``l3max = l4max = 0``, the radial-only
``chebyshev`` (``basis/Chebyshev.java``; descriptor sizes as lmax = 0), or ``mirror`` /
``share`` entries that take another species' basis (``basis/MirrorBasis.java``,
``basis/SharedBasis.java``; a mirror also swaps the roles of the two species
among the neighbours of its centres).  Every error is raised as ``ValueError`` with the reference's wording
where it has one (the reference raises ``IllegalArgumentException``).
"""
from __future__ import annotations

import json
import math
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence

import numpy as np

VERSION = 6  # NNAP.java:77

WTYPES = {"default": 0, "none": -1, "full": 2, "exfull": 3, "fuse": 4, "rfuse": 5, "exfuse": 6}
WTYPE_DEFAULT, WTYPE_NONE, WTYPE_FULL, WTYPE_EXFULL, WTYPE_FUSE, WTYPE_RFUSE, WTYPE_EXFUSE = 0, -1, 2, 3, 4, 5, 6

DEFAULT_NMAX, DEFAULT_LMAX, DEFAULT_RCUT = 5, 6, 6.0  # SphericalChebyshev.java:18-22
DEFAULT_HIDDEN = (32, 32)  # FeedForward.java:28
WS_EPS = 1e-5  # weight standardisation epsilon of the rfuse weight (WTypeBasis.java:20, used at :216)


# columns the L3 / L4 invariants add per radial channel, by l3max / l4max
# (SphericalChebyshev.java:15-16)
L3_COLS = (0, 0, 2, 4, 9, 14, 23)
L4_COLS = (0, 1, 3, 9)


@dataclass
class Basis:
    """One spherical-Chebyshev or radial-only Chebyshev basis (a part of a
    merge basis)."""
    nmax: int
    lmax: int
    rcut: float
    wtype: int
    typed_wtype: int
    P: int
    W: np.ndarray       # [ntypes][P][K]
    radial: bool = False
    l3max: int = 0      # L3 / L4 invariants (SphericalChebyshev.java:33-42)
    l4max: int = 0

    @property
    def K(self) -> int:
        return self.nmax + 1

    @property
    def lmm(self) -> int:
        """The harmonics' lmax: max(lmax, l3max, l4max)."""
        return max(self.lmax, self.l3max, self.l4max)

    @property
    def M(self) -> int:
        return (self.lmm + 1) ** 2

    @property
    def SL(self) -> int:
        """Descriptor columns per radial channel: L2, then L3, then L4."""
        return self.lmax + 1 + L3_COLS[self.l3max] + L4_COLS[self.l4max]

    @property
    def S(self) -> int:
        return self.P * self.SL


@dataclass
class Species:
    symbol: str
    nmax: int
    lmax: int
    rcut: float
    wtype: int          # wtype as written in the file
    typed_wtype: int    # after the one-species collapse
    P: int
    W: np.ndarray       # [ntypes][P][K]
    hidden: List[int]
    nn_param: np.ndarray
    norm_param: np.ndarray
    mirror: int = 0     # mirror basis: neighbour species `mirror` and this one swap roles (0: none)
    radial: bool = False  # radial-only chebyshev basis: fp[p] = sum_j fc Rnp[p] (sizes as lmax = 0)
    parts: Optional[List[Basis]] = None  # merge basis (MergedBasis.java): fp = the parts' fp concatenated;
                                         # the single-basis fields above then hold part 0, rcut the largest
    l3max: int = 0
    l4max: int = 0

    @property
    def K(self) -> int:
        return self.nmax + 1

    @property
    def M(self) -> int:
        return (max(self.lmax, self.l3max, self.l4max) + 1) ** 2

    @property
    def S(self) -> int:
        return sum(b.S for b in self.bases())

    def bases(self) -> List[Basis]:
        """The species' bases in fp order (one unless it is a merge basis)."""
        if self.parts:
            return list(self.parts)
        return [Basis(self.nmax, self.lmax, self.rcut, self.wtype, self.typed_wtype, self.P, self.W, self.radial,
                      self.l3max, self.l4max)]


@dataclass
class Model:
    species: List[Species]
    units: Optional[str] = None
    source: Dict[str, Any] = field(default_factory=dict)

    @property
    def ntypes(self) -> int:
        return len(self.species)

    @property
    def symbols(self) -> List[str]:
        return [s.symbol for s in self.species]

    @property
    def rcut_max(self) -> float:
        return max(s.rcut for s in self.species)


def _get(m: dict, *keys, default=None):
    for k in keys:
        if k in m and m[k] is not None:
            return m[k]
    return default


def _size_n(wtype: int, ntypes: int, nmax: int, fuse_size: int) -> int:
    """WTypeBasis.getSizeN_ (WTypeBasis.java:331-354)."""
    if wtype == WTYPE_EXFULL:
        return (ntypes + 1) * (nmax + 1) if ntypes > 1 else nmax + 1
    if wtype in (WTYPE_FULL, WTYPE_RFUSE):
        return ntypes * (nmax + 1)
    if wtype == WTYPE_NONE:
        return nmax + 1
    if wtype == WTYPE_DEFAULT:
        return 2 * nmax + 2 if ntypes > 1 else nmax + 1
    if wtype == WTYPE_FUSE:
        return fuse_size * (nmax + 1)
    if wtype == WTYPE_EXFUSE:
        return (fuse_size + 1) * (nmax + 1)
    raise ValueError(f"Input wtype MUST be in {{-1, 0, 2, 3, 4, 5, 6}}, input: {wtype}")


def _mixing(wtype: int, ntypes: int, nmax: int, fuse_w, rfuse_w, rfuse_scale: float, ws: bool):
    """Internal W[ntypes][P][K] of WTypeBasis.updateParameters_ (WTypeBasis.java:202-304)."""
    K = nmax + 1
    typed = wtype
    if ntypes == 1 and wtype in (WTYPE_EXFULL, WTYPE_FULL, WTYPE_NONE, WTYPE_DEFAULT):
        typed = WTYPE_NONE
    fuse_size = 0 if wtype not in (WTYPE_FUSE, WTYPE_EXFUSE) else len(fuse_w) // ntypes
    size_n = _size_n(wtype, ntypes, nmax, fuse_size)
    if rfuse_w is not None:
        P = len(rfuse_w) // size_n
    else:
        P = size_n
    W = np.zeros((ntypes, P, K))
    if rfuse_w is not None:
        rw = np.asarray(rfuse_w, dtype=float)
        # the rfuse weight is the internal weight itself (fillD of mRFuseWeight)
        rw = rw.reshape(ntypes, P, K)
        if ws:
            out = np.empty_like(rw)
            for p in range(P):
                norm = WS_EPS + np.abs(rw[:, p, :]).sum()
                norm /= size_n
                out[:, p, :] = rw[:, p, :] / norm
            rw = out
        W[:] = rfuse_scale * rw
        return typed, P, W
    for t in range(1, ntypes + 1):
        for n in range(K):
            if typed == WTYPE_NONE:
                W[t - 1, n, n] = 1.0
            elif typed == WTYPE_DEFAULT:
                wt = float(t) if (t & 1) == 1 else float(-t)
                W[t - 1, n, n] = 1.0
                W[t - 1, K + n, n] = wt
            elif typed == WTYPE_FULL:
                W[t - 1, (t - 1) * K + n, n] = 1.0
            elif typed == WTYPE_EXFULL:
                W[t - 1, n, n] = 1.0
                W[t - 1, t * K + n, n] = 1.0
            elif typed == WTYPE_FUSE:
                for k in range(fuse_size):
                    W[t - 1, k * K + n, n] = fuse_w[(t - 1) * fuse_size + k]
            elif typed == WTYPE_EXFUSE:
                W[t - 1, n, n] = 1.0
                for k in range(fuse_size):
                    W[t - 1, (k + 1) * K + n, n] = fuse_w[(t - 1) * fuse_size + k]
            else:
                raise ValueError(f"Unsupported wtype {typed}")
    return typed, P, W


def _fuse_weight(b, wtype: int, ntypes: int):
    """WTypeBasis.getFuseWeight_ (WTypeBasis.java:365-389): read for the fuse
    wtypes only (other wtypes ignore the key); a nested list without
    `fuse_size` is an older file whose inner lists are the columns of the
    [ntypes][fuse] matrix; `fuse_size` alone gives zeros."""
    if wtype not in (WTYPE_FUSE, WTYPE_EXFUSE):
        return None
    fs = b.get("fuse_size")
    fw = b.get("fuse_weight")
    if fw is not None:
        if len(fw) == 0:
            raise ValueError("Size of fuse weight mismatch")
        nested = isinstance(fw[0], (list, tuple))
        if fs is None and nested:
            out = np.asarray(fw, dtype=float).T.ravel()
        elif nested:
            mat = np.asarray(fw, dtype=float)
            if mat.ndim != 2 or mat.shape[1] != int(fs):
                raise ValueError("Column number of fuse weight mismatch")
            out = mat.ravel()
        else:
            out = np.asarray(fw, dtype=float)
            if fs is not None and len(out) != int(fs) * ntypes:
                raise ValueError("Size of fuse weight mismatch")
    else:
        if fs is None:
            raise ValueError("Key `fuse_weight` or `fuse_size` required for fuse wtype")
        out = np.zeros(int(fs) * ntypes)
    if len(out) == 0 or len(out) % ntypes:
        raise ValueError("Size of fuse weight mismatch")
    return out


def _rfuse_weight(b, wtype: int, size_n: int):
    """WTypeBasis.getRFuseWeight_ (WTypeBasis.java:391-410): read for rfuse
    only; `rfuse_size` alone gives zeros."""
    if wtype != WTYPE_RFUSE:
        return None
    if b.get("post_fuse_size") is not None or b.get("post_fuse_weight") is not None:
        raise ValueError("Key `post_fuse_size` or `post_fuse_weight` is invalid for wtype=='rfuse'")
    rs = b.get("rfuse_size")
    rw = b.get("rfuse_weight")
    if rw is not None:
        out = np.asarray(rw, dtype=float)
        if rs is not None and len(out) != int(rs) * size_n:
            raise ValueError("Size of rfuse weight mismatch")
    else:
        if rs is None:
            raise ValueError("Key `rfuse_weight` or `rfuse_size` required for wtype=='rfuse'")
        out = np.zeros(int(rs) * size_n)
    if len(out) == 0 or len(out) % size_n:
        raise ValueError("Size of rfuse weight mismatch")
    return out


def _post_fuse(b, wtype: int, ntypes: int, nmax: int, fuse_w):
    """The rfuse weight [ntypes][P][K] (flat) equal to a basis' post-fuse weight
    (WTypeBasis.getPostFuseWeight_ and postFuse2RFuse_, WTypeBasis.java:
    391-540): post[P][sizeN] applied after the base mixing, so
    rfuse[t] = post . W_base[t].  None without post-fuse keys."""
    has = any(b.get(k) is not None for k in ("post_fuse", "post_fuse_size", "post_fuse_weight"))
    if wtype == WTYPE_RFUSE:
        if b.get("post_fuse_size") is not None or b.get("post_fuse_weight") is not None:
            raise ValueError("Key `post_fuse_size` or `post_fuse_weight` is invalid for wtype=='rfuse'")
        return None
    flag = b.get("post_fuse")
    if (flag is not None and not flag) or not has:
        return None
    if b.get("post_fuse_weight") is None:
        raise ValueError("Key `post_fuse_weight` always required for post_fuse, training with post_fuse is invalid now")
    post = np.asarray(b["post_fuse_weight"], dtype=float)
    if post.ndim != 1:
        raise ValueError("post_fuse_weight must be a flat list")
    fuse_size = 0 if wtype not in (WTYPE_FUSE, WTYPE_EXFUSE) else len(fuse_w) // ntypes
    sn = _size_n(wtype, ntypes, nmax, fuse_size)
    if b.get("post_fuse_size") is not None and len(post) != int(b["post_fuse_size"]) * sn:
        raise ValueError("Size of post fuse weight mismatch")
    _, _, Wb = _mixing(wtype, ntypes, nmax, fuse_w, None, 1.0, False)   # [ntypes][sizeN][K]
    P = len(post) // sn
    return np.einsum("pq,tqn->tpn", post.reshape(P, sn), Wb).ravel()


def _feed_forward(nn: dict, S: int):
    """FeedForward.load (FeedForward.java:131-221): (hidden dims, packed
    [hidden W | output W | hidden b | output b])."""
    in_dim = nn.get("input_dim")
    if in_dim is not None and int(in_dim) != S:
        raise ValueError(f"Input dimension mismatch: {int(in_dim)} vs {S}")
    hidden = [int(h) for h in (_get(nn, "hidden_dims", "nnarch", default=list(DEFAULT_HIDDEN)))]
    if len(hidden) == 0:
        raise ValueError("At least one hidden layer is required")
    hw_list, hb_list = [], []
    hws = nn.get("hidden_weights")
    hbs = nn.get("hidden_biases")
    if hws is not None and len(hws) != len(hidden):
        raise ValueError(f"The number of hidden weights mismatch: {len(hws)} vs {len(hidden)}")
    if hbs is not None and len(hbs) != len(hidden):
        raise ValueError(f"The number of hidden biases mismatch: {len(hbs)} vs {len(hidden)}")
    cols = S
    for li, h in enumerate(hidden):
        if hws is not None:
            w = np.asarray(hws[li], dtype=float)
            if w.shape != (h, cols):
                raise ValueError(f"shape of hidden weight '{li}' mismatch: {w.shape} vs {(h, cols)}")
        else:
            w = np.zeros((h, cols))
        hw_list.append(w.reshape(-1))
        if hbs is not None:
            bb = np.asarray(hbs[li], dtype=float)
            if bb.shape != (h,):
                raise ValueError(f"Size of hidden bias '{li}' mismatch: {bb.shape[0]} vs {h}")
        else:
            bb = np.zeros(h)
        hb_list.append(bb)
        cols = h
    ow = np.asarray(nn.get("output_weight", np.zeros(hidden[-1])), dtype=float)
    if ow.shape != (hidden[-1],):
        raise ValueError(f"Size of output weight mismatch: {ow.shape[0]} vs {hidden[-1]}")
    ob = float(nn.get("output_bias", 0.0) or 0.0)
    return hidden, np.concatenate(hw_list + [ow] + hb_list + [np.array([ob])])


def _shared_feed_forward(nn: dict, S: int, base_of, i: int, bases):
    """SharedFeedForward.load (SharedFeedForward.java:235-353) as the packed
    block its updateParameters_ (:163-198) fills: the base species' network
    with each hidden layer's first (hidden - shared) rows, and the output layer
    unless shared, taken from this entry."""
    if nn.get("share") is None:
        raise ValueError("Key `share` required for shared_feed_forward")
    b = int(nn["share"])
    hidden, base = base_of(i, b)
    if nn.get("shared_flags") is not None:
        raise ValueError("shared in layer nn is invalid now")
    in_dim = nn.get("input_dim")
    if in_dim is not None and int(in_dim) != S:
        raise ValueError(f"Input dimension mismatch: {int(in_dim)} vs {S}")
    if sum(bp.S for bp in bases[b - 1]) != S:
        raise ValueError("Input dimensions mismatch for shared nn")
    shd = _get(nn, "shared_hidden_dims", "shared_nnarch")
    if shd is None:
        raise ValueError("Key `shared_hidden_dims` required for shared_feed_forward")
    shd = [int(v) for v in shd]
    L = len(hidden)
    if len(shd) == L:
        shd.append(0)   # the output layer defaults to not shared
    if len(shd) != L + 1:
        raise ValueError(f"The number of hidden layers mismatch: {len(shd) - 1} vs {L}")
    out = base.copy()
    hw_size = sum(h * c for h, c in zip(hidden, [S] + hidden[:-1]))
    hb_off = hw_size + hidden[-1]
    ob_off = hb_off + sum(hidden)
    hws = nn.get("hidden_weights")
    hbs = nn.get("hidden_biases")
    if hws is not None and len(hws) != L:
        raise ValueError(f"The number of hidden weights mismatch: {len(hws)} vs {L}")
    if hbs is not None and len(hbs) != L:
        raise ValueError(f"The number of hidden biases mismatch: {len(hbs)} vs {L}")
    woff = boff = 0
    cols = S
    for li, h in enumerate(hidden):
        sh = shd[li]
        if sh < 0 or sh > h:
            raise ValueError(f"invalid shared dimension: {sh} vs {h}")
        own = h - sh
        if hws is not None:
            rows = hws[li]
            if len(rows) != own:
                raise ValueError(f"nrows of hidden weight '{li}' mismatch: {len(rows)} vs {own}")
            w = np.asarray(rows, dtype=float).reshape(own, -1) if own else np.zeros((0, cols))
            if own and w.shape[1] != cols:
                raise ValueError(f"ncols of hidden weight '{li}' mismatch: {w.shape[1]} vs {cols}")
            out[woff:woff + own * cols] = w.ravel()
        else:
            out[woff:woff + own * cols] = 0.0
        if hbs is not None:
            bb = np.asarray(hbs[li], dtype=float)
            if bb.shape != (own,):
                raise ValueError(f"Size of hidden bias '{li}' mismatch: {bb.shape[0]} vs {own}")
            out[hb_off + boff:hb_off + boff + own] = bb
        else:
            out[hb_off + boff:hb_off + boff + own] = 0.0
        woff += h * cols
        boff += h
        cols = h
    if shd[L] not in (0, 1):
        raise ValueError(f"invalid shared dimension: {shd[L]} vs 1")
    if shd[L] == 0:
        ow = np.asarray(nn.get("output_weight", np.zeros(cols)), dtype=float)
        if ow.shape != (cols,):
            raise ValueError(f"Size of output weight mismatch: {ow.shape[0]} vs {cols}")
        out[hw_size:hw_size + cols] = ow
        out[ob_off] = float(nn.get("output_bias", 0.0) or 0.0)
    return list(hidden), out


def _basis(b: dict, btype: str, ntypes: int) -> Basis:
    """SphericalChebyshev.load / Chebyshev.load (SphericalChebyshev.java:60-100,
    Chebyshev.java) with the WTypeBasis mixing (WTypeBasis.java:202-304)."""
    radial = btype in ("chebyshev", "cheby")  # Chebyshev.java: no angular part
    if not radial and btype not in ("spherical_chebyshev", "sph_cheby"):
        raise ValueError(f"Unsupported basis type: {btype}")
    nmax = int(_get(b, "nmax", default=DEFAULT_NMAX))
    lmax = 0 if radial else int(_get(b, "lmax", default=DEFAULT_LMAX))
    l3 = 0 if radial else int(_get(b, "l3max", default=0))
    l4 = 0 if radial else int(_get(b, "l4max", default=0))
    if lmax < 0 or lmax > 12:   # SphericalChebyshev.java:32-34 (lmax checked as max(lmax, l3max, l4max))
        raise ValueError(f"Input lmax MUST be in [0, 12], input: {lmax}")
    if l3 < 0 or l3 > 6:
        raise ValueError(f"Input l3max MUST be in [0, 6], input: {l3}")
    if l4 < 0 or l4 > 3:
        raise ValueError(f"Input l4max MUST be in [0, 3], input: {l4}")
    if nmax < 0 or nmax > 20:   # WTypeBasis.java:56
        raise ValueError(f"Input nmax MUST be in [0, 20], input: {nmax}")
    rcut = float(_get(b, "rcut", default=DEFAULT_RCUT))
    wt = _get(b, "wtype", default="default")
    wtype = int(wt) if isinstance(wt, (int, float)) else WTYPES.get(str(wt))
    if wtype is None:
        raise ValueError(f"Input wtype MUST be in {{-1, 0, 2, 3, 4, 5, 6}}, input: {wt}")
    if wtype not in WTYPES.values():
        raise ValueError(f"Input wtype MUST be in {{-1, 0, 2, 3, 4, 5, 6}}, input: {wtype}")
    fuse_w = _fuse_weight(b, wtype, ntypes)
    rfuse_w = _rfuse_weight(b, wtype, _size_n(wtype, ntypes, nmax, 0))
    rfuse_scale = float(b.get("rfuse_scale", 1.0) or 1.0)
    ws = bool(_get(b, "weight_standardization", "ws", default=False))
    post = _post_fuse(b, wtype, ntypes, nmax, fuse_w)
    if post is not None:
        # an older file's post-fuse weight becomes the rfuse weight it equals
        # (SphericalChebyshev.java:81-94, WTypeBasis.postFuse2RFuse_)
        rfuse_w, wtype, fuse_w = post, WTYPE_RFUSE, None
        rfuse_scale = float(b.get("post_fuse_scale", 1.0) or 1.0)
    typed, P, W = _mixing(wtype, ntypes, nmax, fuse_w, rfuse_w, rfuse_scale, ws)
    return Basis(nmax, lmax, rcut, wtype, typed, P, W, radial, l3, l4)


def load_model(src) -> Model:
    """Parse a model (path to .json/.yaml/.yml, JSON text, or a dict) as
    ``new NNAP(path)`` does (NNAP.java:129-307, 402-414)."""
    if isinstance(src, dict):
        info = src
    elif isinstance(src, str) and src.lstrip().startswith("{"):
        info = json.loads(src)
    else:
        path = str(src)
        with open(path, "r") as f:
            if path.endswith((".yaml", ".yml")):
                import yaml
                info = yaml.safe_load(f)
            else:
                info = json.load(f)
    ver = info.get("version")
    if ver is not None and int(ver) > VERSION:
        raise ValueError(f"Unsupported version: {int(ver)}")
    models = info.get("models")
    if models is None:
        raise ValueError("No models in ModelInfo")
    ntypes = len(models)
    norm_sigma_eng = None
    norm_mu_eng = None
    for m in models:
        if norm_sigma_eng is None:
            norm_sigma_eng = m.get("norm_sigma_eng")
        if norm_mu_eng is None:
            norm_mu_eng = m.get("norm_mu_eng")
    sig_e = 1.0 if norm_sigma_eng is None else float(norm_sigma_eng)
    mu_e = 0.0 if norm_mu_eng is None else float(norm_mu_eng)
    # pass 1 (Basis.java:22-50): the species' own bases; mirror / share entries
    # only record the species whose basis they take
    bases = [None] * ntypes     # per species: its list of Basis (one unless merge)
    btypes = [""] * ntypes
    ref = [0] * ntypes
    is_mirror = [False] * ntypes
    for i, m in enumerate(models):
        if m.get("symbol") is None:
            raise ValueError("No symbol in model")
        if m.get("torch") is not None:
            raise ValueError("torch model is invalid now.")
        b = m.get("basis") or {"type": "spherical_chebyshev"}
        btype = str(b.get("type", "spherical_chebyshev"))
        btypes[i] = btype
        if btype in ("mirror", "mirror_basis", "share", "shared_basis"):
            mir = btype.startswith("mirror")
            key = "mirror" if mir else "share"
            if b.get(key) is None:  # MirrorBasis.java:61-66, SharedBasis.java:61-66
                raise ValueError(f"Key `{key}` required for {'mirror_basis' if mir else 'shared_basis'}")
            ref[i] = int(b[key])
            if not 1 <= ref[i] <= ntypes:
                raise ValueError("mirror/share type out of range")
            is_mirror[i] = mir
            continue
        if btype == "merge":  # MergedBasis.load (MergedBasis.java:204-229)
            sub = b.get("basis")
            if sub is None:
                raise ValueError("Key `basis` required for merge load")
            parts = []
            for bm in sub:
                st = str(bm.get("type", "spherical_chebyshev"))
                if st not in ("spherical_chebyshev", "sph_cheby", "chebyshev", "cheby"):
                    raise ValueError(f"Unsupported basis type: {st}")
                parts.append(_basis(bm, st, ntypes))
            if not parts:
                raise ValueError("Merge basis can not be null or empty")
            bases[i] = parts
            continue
        bases[i] = [_basis(b, btype, ntypes)]
    # pass 2 (Basis.java:51-66): mirror / share take the named species' basis
    for i in range(ntypes):
        if ref[i]:
            if ref[ref[i] - 1]:
                raise ValueError("a mirror/share basis of a mirror/share basis is not supported")
            bases[i] = bases[ref[i] - 1]
    # pass 3 (NeuralNetwork.java:24-65): the networks.  feed_forward entries
    # first; then a mirror species takes the full network of the species it
    # mirrors (SharedFeedForward.full; its `nn` must be absent) and a
    # shared_feed_forward entry takes another species' network with its own
    # leading rows per layer (SharedFeedForward.java:163-198, 235-353)
    nets = [None] * ntypes      # (hidden, nn_param)
    shared_nn = [False] * ntypes
    for i, m in enumerate(models):
        S = sum(bp.S for bp in bases[i])
        if is_mirror[i]:
            if m.get("nn") is not None:
                raise ValueError("nn data in mirror ModelInfo MUST be empty")
            continue
        nn = m.get("nn") or {}
        ntype = str(nn.get("type", "feed_forward"))
        if ntype == "shared_feed_forward":
            shared_nn[i] = True
            continue
        if ntype != "feed_forward":
            raise ValueError(f"Unsupported nn type: {nn.get('type')}")
        nets[i] = _feed_forward(nn, S)

    def base_of(i, b):
        if not (1 <= b <= ntypes) or is_mirror[b - 1] or shared_nn[b - 1]:
            raise ValueError(f"species {i + 1} takes the network of species {b}, "
                             "which has no feed_forward network of its own")
        return nets[b - 1]

    for i, m in enumerate(models):
        if is_mirror[i]:
            nets[i] = base_of(i, ref[i])
        elif shared_nn[i]:
            nets[i] = _shared_feed_forward(m["nn"], sum(bp.S for bp in bases[i]), base_of, i, bases)
    # normalisation (NNAP.java:272-306): a mirror species must not carry its
    # own and uses that of the species it mirrors
    out: List[Species] = []
    for i, m in enumerate(models):
        sym = m.get("symbol")
        b0 = bases[i][0]
        S = sum(bp.S for bp in bases[i])
        src = m
        if is_mirror[i]:
            if m.get("ref_eng") is not None:
                raise ValueError("ref_eng in mirror_basis MUST be empty")
            if _get(m, "norm_vec", "norm_sigma", "norm_mu") is not None:
                raise ValueError("norm_vec/norm_sigma/norm_mu in mirror_basis MUST be empty")
            src = models[ref[i] - 1]
        ref_eng = float(src.get("ref_eng", 0.0) or 0.0)
        nsig = _get(src, "norm_sigma", "norm_vec")
        nmu = src.get("norm_mu")
        mu = np.zeros(S) if nmu is None else np.asarray(nmu, dtype=float)
        sg = np.ones(S) if nsig is None else np.asarray(nsig, dtype=float)
        if mu.shape != (S,) or sg.shape != (S,):
            raise ValueError("norm_mu/norm_sigma size mismatch")
        norm = np.concatenate([[mu_e + ref_eng, sig_e], mu, sg])
        hidden, nn_param = nets[i]
        rcut = max(bp.rcut for bp in bases[i])
        out.append(Species(str(sym), b0.nmax, b0.lmax, rcut, b0.wtype, b0.typed_wtype, b0.P, b0.W, list(hidden),
                           nn_param.copy(), norm, ref[i] if is_mirror[i] else 0, b0.radial,
                           list(bases[i]) if len(bases[i]) > 1 or btypes[i] == "merge" else None,
                           b0.l3max, b0.l4max))
    return Model(out, info.get("units"), info)


def _random_network(rng, S: int, hidden: Sequence[int]) -> dict:
    """A feed_forward entry drawn as FeedForward.initParameters draws it
    (FeedForward.java:105-128)."""
    cols = S
    hws, hbs = [], []
    for h in hidden:
        bw = math.sqrt(6.0 / cols)
        hws.append(rng.uniform(-bw, bw, size=(h, cols)).tolist())
        cols = h
    cols = S
    for h in hidden:
        bb = math.sqrt(1.0 / cols)
        hbs.append(rng.uniform(-bb, bb, size=h).tolist())
        cols = h
    bo = math.sqrt(3.0 / cols)
    ow = rng.uniform(-bo, bo, size=cols).tolist()
    ob = float(rng.uniform(-math.sqrt(1.0 / cols), math.sqrt(1.0 / cols)))
    return {"type": "feed_forward", "input_dim": S, "hidden_dims": list(hidden), "hidden_weights": hws,
            "hidden_biases": hbs, "output_weight": ow, "output_bias": ob}


def random_merge_model(symbols: Sequence[str], parts: Sequence[dict], hidden: Sequence[int], seed: int) -> dict:
    """Seeded random-init model whose species all carry a merge basis of the
    given part entries (each a spherical_chebyshev or chebyshev basis entry
    with its nmax, lmax, rcut, wtype; MergedBasis.java)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ntypes = len(symbols)
    S = 0
    for b in parts:
        radial = str(b.get("type", "spherical_chebyshev")) in ("chebyshev", "cheby")
        wt = WTYPES[str(b.get("wtype", "default"))]
        nmax = int(b.get("nmax", DEFAULT_NMAX))
        P = _size_n(wt, ntypes, nmax, 0)
        if ntypes == 1 and wt in (WTYPE_DEFAULT, WTYPE_FULL, WTYPE_EXFULL):
            P = _size_n(WTYPE_NONE, 1, nmax, 0)
        S += P * (1 if radial else int(b.get("lmax", DEFAULT_LMAX)) + 1 + L3_COLS[int(b.get("l3max", 0))]
                  + L4_COLS[int(b.get("l4max", 0))])
    models = []
    for sym in symbols:
        models.append({"symbol": sym, "basis": {"type": "merge", "basis": [dict(b) for b in parts]},
                       "nn": _random_network(rng, S, hidden), "ref_eng": 0.0,
                       "norm_mu": [0.0] * S, "norm_sigma": [1.0] * S})
    models[0]["norm_mu_eng"] = 0.0
    models[0]["norm_sigma_eng"] = 1.0
    return {"version": VERSION, "units": "metal", "models": models}


def random_model(symbols: Sequence[str], nmax: int, lmax: int, rcut: float, hidden: Sequence[int],
                 seed: int, wtype: str = "default", l3max: int = 0, l4max: int = 0) -> dict:
    """Seeded random-init model in the reference file format, drawn from the
    distributions of FeedForward.initParameters (FeedForward.java:105-128):
    hidden W ~ U(+-sqrt(6/fan_in)), biases ~ U(+-sqrt(1/fan_in)),
    output W ~ U(+-sqrt(3/fan_in)).  Normalisation is identity until
    ``set_normalization`` is called."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ntypes = len(symbols)
    P = _size_n(WTYPES[wtype], ntypes, nmax, 0)
    if ntypes == 1 and WTYPES[wtype] in (WTYPE_DEFAULT, WTYPE_FULL, WTYPE_EXFULL):
        P = _size_n(WTYPE_NONE, 1, nmax, 0)
    S = P * (lmax + 1 + L3_COLS[l3max] + L4_COLS[l4max])
    models = []
    for sym in symbols:
        models.append({
            "symbol": sym,
            "basis": {"type": "spherical_chebyshev", "nmax": nmax, "lmax": lmax, "l3max": l3max, "l4max": l4max,
                      "rcut": rcut, "wtype": wtype},
            "nn": _random_network(rng, S, hidden),
            "ref_eng": 0.0,
            "norm_mu": [0.0] * S,
            "norm_sigma": [1.0] * S,
        })
    models[0]["norm_mu_eng"] = 0.0
    models[0]["norm_sigma_eng"] = 1.0
    return {"version": VERSION, "units": "metal", "models": models}


def set_normalization(model: dict, fp_by_species: Sequence[Optional[np.ndarray]], digits: int = 6) -> dict:
    """Fix norm_mu / norm_sigma per species from descriptor samples (rows =
    atoms).  Values are rounded to ``digits`` significant figures so that a
    model regenerated on another host is identical."""
    def rnd(v):
        return float(f"{v:.{digits}g}")
    for m, fp in zip(model["models"], fp_by_species):
        if fp is None or len(fp) == 0:
            continue
        mu = fp.mean(axis=0)
        sg = fp.std(axis=0)
        sg = np.where(sg > 1e-12 * np.maximum(1.0, np.abs(mu)), sg, 1.0)
        m["norm_mu"] = [rnd(v) for v in mu]
        m["norm_sigma"] = [rnd(v) for v in sg]
    return model


def model_json(model: dict) -> str:
    return json.dumps(model)
