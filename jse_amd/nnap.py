"""Host-side mirror of jse's NNAP potential over libjsenn_hip.

``NNAP`` keeps the names, argument meaning and error behaviour of the
reference's Java API (paths relative to the jse repository root):

* ``NNAP(model_path_or_dict)``                 NNAP.java:402-414
* ``calEnergy / calEnergies / calForces``      IPotential.java:145-185
* ``calStress / calStresses``                  IPotential.java:197-311 (stress = -virial / volume)
* ``calEnergyForceVirials``                    IPotential.java:536-562 -> IPairPotential.java:811-1022
* ``symbols / ntypes / rcutMax / close``       NNAP.java:470-490

An ``AtomData`` carries an orthogonal box, Cartesian FP64 positions and
1-based types (optionally with species symbols; types are then mapped onto
the model's species by symbol, as IHasSymbol.typeMap does, IHasSymbol.java:
51-99).  Every evaluation runs the whole path on the GPU: wrap, cell list,
descriptor, MLP, backward, FP64 force/virial accumulation.  Native errors are
raised as ``NNAPError`` carrying the library's status code (the reference
raises IllegalStateException("Exit code: n"), NNAP.java:756-787).

The native library is required: importing works without it, but any
evaluation raises ``RuntimeError`` when ``libjsenn_hip.so`` cannot be loaded.
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence, Union

import numpy as np

from . import model as _model

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libjsenn_hip.so")

# precision modes (include/jse_nnap_hip.h): FP32 is the plain path that meets the
# reference's FP64 results to 1e-6 eV/atom, 1e-5 eV/A (FP64 geometry, descriptor
# and MLP; only the backward's dE/dA contraction rounds to FP32); BF16_MLP and
# FP32_FAST evaluate the geometry and the descriptor in FP32, with a BF16 or
# FP32 MLP
FP32 = 0
BF16_MLP = 1
FP32_FAST = 2
ENERGY, FORCE, VIRIAL = 1, 2, 4

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_flt_p = ctypes.POINTER(ctypes.c_float)
_c_ll_p = ctypes.POINTER(ctypes.c_longlong)

EXPORTS = [
    "jse_nnap_last_error", "jse_nnap_model_load", "jse_nnap_model_parse", "jse_nnap_model_parse_blocks",
    "jse_nnap_sh_invariants",
    "jse_nnap_model_create", "jse_nnap_model_info",
    "jse_nnap_model_fp_size", "jse_nnap_model_destroy", "jse_nnap_system_create", "jse_nnap_system_destroy",
    "jse_nnap_system_set_box", "jse_nnap_system_set_positions", "jse_nnap_system_set_positions_device",
    "jse_nnap_system_compute", "jse_nnap_system_get_results", "jse_nnap_system_compute_fp",
    "jse_nnap_system_backward_energy", "jse_nnap_system_backward_energy_force",
    "jse_nnap_system_get_neighbors", "jse_nnap_system_benchmark", "jse_nnap_system_nl_stats",
    "jse_nnap_system_get_pair_gradients", "jse_nnap_batch_create", "jse_nnap_system_get_replica_results",
    "jse_nnap_system_set_graphs", "jse_nnap_system_autotune", "jse_nnap_system_set_tuning", "jse_nnap_server_create", "jse_nnap_server_add_replica", "jse_nnap_server_submit",
    "jse_nnap_server_wait", "jse_nnap_server_collect", "jse_nnap_server_drain", "jse_nnap_server_stats",
    "jse_nnap_server_destroy", "jse_nnap_lammps_bind", "jse_nnap_statNeiNumLammps", "jse_nnap_lammps2cuda",
    "jse_nnap_computeLammpsCuda", "jse_nnap_cuda2lammps", "jse_nnap_computeGPUMD_noalloc",
]

SERVER_STATS = ["completed", "atom_evals", "batches", "mean_batch_atoms", "rebuilds", "admissions", "evictions",
                "growths", "resident_replicas", "resident_atoms", "arena_slots", "arena_bytes", "candidate_cap",
                "row_cap", "pending", "device_bytes"]


class NNAPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"Exit code: {code}: {msg}")
        self.code = code


_LIB = None


def lib():
    """Load libjsenn_hip.so (raises if it is missing: there is no fallback)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m jse_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.jse_nnap_last_error.restype = ctypes.c_char_p
    L.jse_nnap_model_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    L.jse_nnap_model_parse.argtypes = [ctypes.c_char_p, _c_int_p, _c_int_p, _c_int_p]
    L.jse_nnap_model_parse_blocks.argtypes = [ctypes.c_char_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_int_p,
                                              _c_int_p, _c_int_p]
    L.jse_nnap_sh_invariants.argtypes = [ctypes.c_int, ctypes.c_int, _c_dbl_p, _c_dbl_p]
    L.jse_nnap_model_create.argtypes = [ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_dbl_p, _c_int_p, _c_int_p,
                                        ctypes.POINTER(_c_dbl_p), ctypes.POINTER(_c_dbl_p), ctypes.POINTER(_c_dbl_p),
                                        ctypes.POINTER(vp)]
    L.jse_nnap_model_info.argtypes = [vp, _c_int_p, _c_dbl_p]
    L.jse_nnap_model_fp_size.argtypes = [vp, ctypes.c_int, _c_int_p]
    L.jse_nnap_model_destroy.argtypes = [vp]
    L.jse_nnap_model_destroy.restype = None
    L.jse_nnap_system_create.argtypes = [vp, ctypes.c_int, _c_int_p, _c_dbl_p, ctypes.c_int, ctypes.POINTER(vp)]
    L.jse_nnap_system_destroy.argtypes = [vp]
    L.jse_nnap_system_destroy.restype = None
    L.jse_nnap_system_set_box.argtypes = [vp, _c_dbl_p]
    L.jse_nnap_system_set_positions.argtypes = [vp, _c_dbl_p]
    L.jse_nnap_system_set_positions_device.argtypes = [vp, vp]
    L.jse_nnap_system_compute.argtypes = [vp, ctypes.c_int]
    L.jse_nnap_system_get_results.argtypes = [vp, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]
    L.jse_nnap_system_compute_fp.argtypes = [vp, _c_dbl_p, ctypes.c_int]
    L.jse_nnap_system_backward_energy.argtypes = [vp, _c_dbl_p, ctypes.POINTER(_c_dbl_p), ctypes.POINTER(_c_dbl_p)]
    L.jse_nnap_system_backward_energy_force.argtypes = [vp, _c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.POINTER(_c_dbl_p),
                                                        ctypes.POINTER(_c_dbl_p)]
    L.jse_nnap_system_get_neighbors.argtypes = [vp, ctypes.c_int, _c_int_p, _c_int_p, _c_flt_p]
    L.jse_nnap_system_benchmark.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p, _c_dbl_p,
                                            ctypes.c_int]
    L.jse_nnap_system_nl_stats.argtypes = [vp, _c_int_p, _c_int_p, _c_ll_p, _c_ll_p]
    L.jse_nnap_system_get_pair_gradients.argtypes = [vp, ctypes.c_int, _c_flt_p]
    L.jse_nnap_batch_create.argtypes = [vp, ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p, ctypes.c_int,
                                        ctypes.POINTER(vp)]
    L.jse_nnap_system_get_replica_results.argtypes = [vp, _c_dbl_p, _c_dbl_p]
    L.jse_nnap_system_set_graphs.argtypes = [vp, ctypes.c_int]
    L.jse_nnap_system_autotune.argtypes = [vp, ctypes.c_int, ctypes.c_int, _c_int_p, _c_dbl_p]
    L.jse_nnap_system_set_tuning.argtypes = [vp, _c_int_p, ctypes.c_int]
    L.jse_nnap_server_create.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_double, ctypes.POINTER(vp)]
    L.jse_nnap_server_add_replica.argtypes = [vp, ctypes.c_int, _c_int_p, _c_dbl_p, _c_int_p]
    L.jse_nnap_server_submit.argtypes = [vp, ctypes.c_int, _c_dbl_p, ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p,
                                         _c_ll_p]
    L.jse_nnap_server_wait.argtypes = [vp, ctypes.c_longlong, _c_dbl_p]
    L.jse_nnap_server_collect.argtypes = [vp, _c_ll_p, _c_dbl_p, _c_int_p, ctypes.c_int, _c_int_p]
    L.jse_nnap_server_drain.argtypes = [vp]
    L.jse_nnap_server_stats.argtypes = [vp, _c_dbl_p, ctypes.c_int]
    L.jse_nnap_server_destroy.argtypes = [vp]
    L.jse_nnap_server_destroy.restype = None
    _LIB = L
    return L


def _check(rc: int):
    if rc != 0:
        raise NNAPError(rc, lib().jse_nnap_last_error().decode(errors="replace"))


def _dp(a):
    return a.ctypes.data_as(_c_dbl_p) if a is not None else None


def _ip(a):
    return a.ctypes.data_as(_c_int_p) if a is not None else None


@dataclass
class AtomData:
    """Orthogonal periodic box [0,L) per axis, FP64 positions [N][3] and 1-based types."""
    box: np.ndarray
    xyz: np.ndarray
    types: np.ndarray
    symbols: Optional[Sequence[str]] = None   # symbol of type t at symbols[t-1]

    @property
    def natoms(self) -> int:
        return int(np.asarray(self.xyz).shape[0])

    def volume(self) -> float:
        b = np.asarray(self.box, dtype=float)
        return float(b[0] * b[1] * b[2])


class System:
    """A device-resident configuration bound to one model: positions,
    neighbour list and outputs stay on the GPU between evaluations.  With
    ``natoms_per_replica`` it holds a batch of independent periodic replicas
    (``box`` then has one row per replica) evaluated in the same launches."""

    def __init__(self, nnap: "NNAP", natoms: int, types, box, precision: int = FP32, natoms_per_replica=None):
        self.nnap = nnap
        self.natoms = int(natoms)
        self.types = np.ascontiguousarray(types, dtype=np.int32)
        if self.types.shape != (self.natoms,):
            raise ValueError("types must have one entry per atom")
        self.box = np.ascontiguousarray(box, dtype=np.float64)
        h = ctypes.c_void_p()
        if natoms_per_replica is None:
            self.nrep = 1
            _check(lib().jse_nnap_system_create(nnap._h, self.natoms, _ip(self.types), _dp(self.box),
                                                int(precision), ctypes.byref(h)))
        else:
            per = np.ascontiguousarray(natoms_per_replica, dtype=np.int32)
            if int(per.sum()) != self.natoms or self.box.shape != (len(per), 3):
                raise ValueError("replica sizes / boxes do not match")
            self.nrep = len(per)
            self.rep_natoms = per
            _check(lib().jse_nnap_batch_create(nnap._h, len(per), _ip(per), _ip(self.types), _dp(self.box),
                                               int(precision), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().jse_nnap_system_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_box(self, box):
        self.box = np.ascontiguousarray(box, dtype=np.float64)
        _check(lib().jse_nnap_system_set_box(self._h, _dp(self.box)))

    def set_positions_device(self, ptr: int):
        """positions from a device pointer (e.g. a torch tensor's data_ptr())"""
        _check(lib().jse_nnap_system_set_positions_device(self._h, ctypes.c_void_p(ptr)))

    def set_positions(self, xyz):
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        if xyz.shape != (self.natoms, 3):
            raise ValueError("positions must be [natoms][3]")
        _check(lib().jse_nnap_system_set_positions(self._h, _dp(xyz)))

    def compute(self, what: int = ENERGY | FORCE | VIRIAL):
        _check(lib().jse_nnap_system_compute(self._h, int(what)))

    def results(self, energies=True, forces=True, virials=True) -> Dict[str, np.ndarray]:
        n = self.natoms
        e = np.zeros(1)
        E = np.zeros(n) if energies else None
        F = np.zeros((n, 3)) if forces else None
        v = np.zeros(9) if virials else None
        V = np.zeros((n, 9)) if virials else None
        _check(lib().jse_nnap_system_get_results(self._h, _dp(e), _dp(E), _dp(F), _dp(v), _dp(V)))
        out = {"energy": float(e[0])}
        if energies:
            out["energies"] = E
        if forces:
            out["forces"] = F
        if virials:
            out["virial"] = v
            out["virials"] = V
        return out

    def replica_results(self, virials=True):
        e = np.zeros(self.nrep)
        v = np.zeros((self.nrep, 9)) if virials else None
        _check(lib().jse_nnap_system_get_replica_results(self._h, _dp(e), _dp(v)))
        return (e, v) if virials else e

    def set_graphs(self, enable: bool):
        _check(lib().jse_nnap_system_set_graphs(self._h, 1 if enable else 0))

    def descriptors(self) -> np.ndarray:
        smax = max(s.S for s in self.nnap.model.species)
        fp = np.zeros((self.natoms, smax))
        _check(lib().jse_nnap_system_compute_fp(self._h, _dp(fp), smax))
        return fp

    def backward_energy(self, w, mixing: bool = False):
        """Parameter gradients of sum_i w[i] E_i (jse_nnap_backwardEnergy summed over
        the atoms, aGradEng = w[i]): one array per species in the packed nn_param
        layout (rGradNnParam); with ``mixing`` also, per species, the
        concatenated [ntypes][P][K] blocks of its bases' mixing weights
        (rGradFpParam: non-zero for fuse / rfuse / exfuse), returned as a second
        list."""
        w = np.ascontiguousarray(w, dtype=np.float64)
        if w.shape != (self.natoms,):
            raise ValueError("w must have one entry per atom")
        sp = self.nnap.model.species
        nt = len(sp)
        outs = [np.zeros(s.nn_param.size) for s in sp]
        ptrs = (_c_dbl_p * nt)(*[_dp(o) for o in outs])
        fps = [np.zeros(sum(nt * b.P * b.K for b in s.bases())) for s in sp] if mixing else None
        fptrs = (_c_dbl_p * nt)(*[_dp(o) for o in fps]) if mixing else None
        _check(lib().jse_nnap_system_backward_energy(self._h, _dp(w), ptrs, fptrs))
        return (outs, fps) if mixing else outs

    def backward_energy_force(self, w=None, bgrad_forces=None, bgrad_virial=None, mixing: bool = False):
        """Parameter gradients of sum_i w[i] E_i + sum_k bgrad_forces[k] . F_k +
        sum bgrad_virial . virial (jse_nnap_backwardEnergyForce over every atom, as
        TrainerNNAP drives it): the network blocks, and with ``mixing`` also the
        mixing-weight blocks, as backward_energy returns them.  For a batch,
        bgrad_virial is [nrep][9] (each replica's by its own virial), so one
        call gives the summed gradient of many training structures."""
        n = self.natoms
        def arr(v, shape):
            if v is None:
                return None
            v = np.ascontiguousarray(v, dtype=np.float64)
            if v.shape != shape:
                raise ValueError(f"expected shape {shape}")
            return v
        w = arr(w, (n,))
        bf = arr(bgrad_forces, (n, 3))
        bv = arr(bgrad_virial, (9,) if self.nrep == 1 else (self.nrep, 9))  # a batch: one per replica
        sp = self.nnap.model.species
        nt = len(sp)
        outs = [np.zeros(s.nn_param.size) for s in sp]
        ptrs = (_c_dbl_p * nt)(*[_dp(o) for o in outs])
        fps = [np.zeros(sum(nt * b.P * b.K for b in s.bases())) for s in sp] if mixing else None
        fptrs = (_c_dbl_p * nt)(*[_dp(o) for o in fps]) if mixing else None
        nul = ctypes.cast(None, _c_dbl_p)
        _check(lib().jse_nnap_system_backward_energy_force(
            self._h, _dp(w) if w is not None else nul, _dp(bf) if bf is not None else nul,
            _dp(bv) if bv is not None else nul, ptrs, fptrs))
        return (outs, fps) if mixing else outs

    def neighbors(self, cap: Optional[int] = None):
        n = self.natoms
        c = ctypes.c_int()
        mx = ctypes.c_int()
        pairs = ctypes.c_longlong()
        un = ctypes.c_longlong()
        _check(lib().jse_nnap_system_nl_stats(self._h, ctypes.byref(c), ctypes.byref(mx), ctypes.byref(pairs),
                                              ctypes.byref(un)))
        cap = max(1, mx.value) if cap is None else cap
        cnt = np.zeros(n, dtype=np.int32)
        key = np.zeros((n, cap), dtype=np.int32)
        d = np.zeros((n, cap, 3), dtype=np.float32)
        _check(lib().jse_nnap_system_get_neighbors(self._h, cap, _ip(cnt), _ip(key),
                                                   d.ctypes.data_as(_c_flt_p)))
        return cnt, key, d

    def pair_gradients(self, cap: int):
        g = np.zeros((self.natoms, cap, 3), dtype=np.float32)
        _check(lib().jse_nnap_system_get_pair_gradients(self._h, int(cap), g.ctypes.data_as(_c_flt_p)))
        return g

    def nl_stats(self):
        c, mx, pairs, un = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().jse_nnap_system_nl_stats(self._h, ctypes.byref(c), ctypes.byref(mx), ctypes.byref(pairs),
                                              ctypes.byref(un)))
        return {"cap": c.value, "max_count": mx.value, "pairs": pairs.value, "unmatched": un.value}

    TUNE_KNOBS = ["desc_fwd_wg_per_cu", "desc_bwd_wg_per_cu", "mlp_wg_per_cu", "nl_wg_per_cu", "gather_waves",
                  "desc_fwd_variant", "desc_bwd_variant"]

    def autotune(self, what: int = ENERGY | FORCE | VIRIAL, iters: int = 5):
        """times candidate launch shapes on this system's data; returns (knobs, ms per evaluation)"""
        t = np.zeros(len(self.TUNE_KNOBS), dtype=np.int32)
        ms = ctypes.c_double()
        _check(lib().jse_nnap_system_autotune(self._h, int(what), int(iters), _ip(t), ctypes.byref(ms)))
        return dict(zip(self.TUNE_KNOBS, t.tolist())), ms.value

    def set_tuning(self, knobs: Dict[str, int]):
        t = np.array([int(knobs.get(k, 0)) for k in self.TUNE_KNOBS], dtype=np.int32)
        _check(lib().jse_nnap_system_set_tuning(self._h, _ip(t), len(t)))

    def benchmark(self, what: int = ENERGY | FORCE | VIRIAL, warmup: int = 3, iters: int = 10):
        ms = ctypes.c_double()
        ph = np.zeros(6)
        _check(lib().jse_nnap_system_benchmark(self._h, int(what), int(warmup), int(iters), ctypes.byref(ms),
                                               _dp(ph), 6))
        return ms.value, ph


class Server:
    """Device-resident replicas evaluated by continuous batching (see
    jse_nnap_server_* in include/jse_nnap_hip.h).  ``submit`` returns a ticket
    at once; ``wait(ticket)`` / ``collect()`` hand back results.  The output
    arrays of a request are kept here until its ticket is released."""

    def __init__(self, nnap: "NNAP", precision: int = BF16_MLP, skin: float = 1.0, hbm_budget_bytes: float = 2e9,
                 max_batch_atoms: int = 65536, max_replicas: int = 4096, max_density: float = 0.08):
        self.nnap = nnap
        h = ctypes.c_void_p()
        _check(lib().jse_nnap_server_create(nnap._h, int(precision), float(skin), float(hbm_budget_bytes),
                                            int(max_batch_atoms), int(max_replicas), float(max_density),
                                            ctypes.byref(h)))
        self._h = h
        self._natoms: List[int] = []
        self._out: Dict[int, tuple] = {}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().jse_nnap_server_destroy(self._h)
            self._h = None
            self._out.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_replica(self, atoms: AtomData) -> int:
        t = np.ascontiguousarray(self.nnap._types_of(atoms), dtype=np.int32)
        box = np.ascontiguousarray(atoms.box, dtype=np.float64)
        rid = ctypes.c_int()
        _check(lib().jse_nnap_server_add_replica(self._h, atoms.natoms, _ip(t), _dp(box), ctypes.byref(rid)))
        self._natoms.append(atoms.natoms)
        return rid.value

    def submit(self, rid: int, xyz, what: int = ENERGY | FORCE | VIRIAL) -> int:
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        n = self._natoms[rid]
        if xyz.shape != (n, 3):
            raise ValueError("positions must be [natoms][3]")
        e = np.zeros(1)
        f = np.zeros((n, 3)) if what & FORCE else None
        v = np.zeros(9) if what & VIRIAL else None
        tk = ctypes.c_longlong()
        _check(lib().jse_nnap_server_submit(self._h, int(rid), _dp(xyz), int(what), _dp(e), _dp(f), _dp(v),
                                            ctypes.byref(tk)))
        self._out[tk.value] = (rid, e, f, v)
        return tk.value

    def _result(self, ticket, latency):
        rid, e, f, v = self._out.pop(ticket)
        out = {"replica": rid, "energy": float(e[0]), "latency_s": latency}
        if f is not None:
            out["forces"] = f
        if v is not None:
            out["virial"] = v
        return out

    def wait(self, ticket: int) -> dict:
        lat = ctypes.c_double()
        rc = lib().jse_nnap_server_wait(self._h, int(ticket), ctypes.byref(lat))
        if rc != 0:
            self._out.pop(ticket, None)
            _check(rc)
        return self._result(ticket, lat.value)

    def collect(self, cap: int = 4096) -> Dict[int, dict]:
        """released completed requests: ticket -> result (with 'status')"""
        tk = np.zeros(cap, dtype=np.int64)
        lat = np.zeros(cap)
        st = np.zeros(cap, dtype=np.int32)
        cnt = ctypes.c_int()
        _check(lib().jse_nnap_server_collect(self._h, tk.ctypes.data_as(_c_ll_p), _dp(lat), _ip(st), cap,
                                             ctypes.byref(cnt)))
        out = {}
        for k in range(cnt.value):
            r = self._result(int(tk[k]), float(lat[k]))
            r["status"] = int(st[k])
            out[int(tk[k])] = r
        return out

    def drain(self):
        _check(lib().jse_nnap_server_drain(self._h))

    def stats(self) -> Dict[str, float]:
        v = np.zeros(len(SERVER_STATS))
        _check(lib().jse_nnap_server_stats(self._h, _dp(v), len(v)))
        return dict(zip(SERVER_STATS, v.tolist()))


class NNAP:
    """jse's NNAP potential on one MI355X.  ``nthreads`` is accepted for API
    parity with the reference and ignored (the work runs on the GPU)."""

    def __init__(self, model_src: Union[str, dict], nthreads: int = 1, precision: int = FP32):
        if isinstance(model_src, str) and model_src.endswith((".yaml", ".yml")):
            import yaml
            with open(model_src) as f:
                model_src = yaml.safe_load(f)
        self.model = _model.load_model(model_src)
        self.precision = int(precision)
        text = json.dumps(model_src) if isinstance(model_src, dict) else model_src
        if not isinstance(text, str):
            raise TypeError("model must be a path, JSON text or a dict")
        h = ctypes.c_void_p()
        _check(lib().jse_nnap_model_load(text.encode(), ctypes.byref(h)))
        self._h = h
        self._dead = False
        self._systems: Dict[tuple, System] = {}

    # --- reference accessors ---
    def symbols(self) -> List[str]:
        return self.model.symbols

    def ntypes(self) -> int:
        return self.model.ntypes

    def rcutMax(self) -> float:
        return self.model.rcut_max

    def nthreads(self) -> int:
        return 1

    def isClosed(self) -> bool:
        return self._dead

    def close(self):
        if self._dead:
            return
        for s in self._systems.values():
            s.close()
        self._systems.clear()
        lib().jse_nnap_model_destroy(self._h)
        self._dead = True

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- evaluation ---
    def _types_of(self, atoms: AtomData) -> np.ndarray:
        t = np.asarray(atoms.types, dtype=np.int32)
        if atoms.symbols is not None:
            syms = self.model.symbols
            mp = np.zeros(len(atoms.symbols) + 1, dtype=np.int32)
            for k, s in enumerate(atoms.symbols):
                if s not in syms:
                    raise ValueError(f"Symbol {s} not in the model symbols {syms}")
                mp[k + 1] = syms.index(s) + 1
            t = mp[t]
        if t.size and (t.min() < 1 or t.max() > self.ntypes()):
            raise ValueError(f"Exist type ({int(t.max())}) greater than the input ntypes ({self.ntypes()})")
        return t

    def system(self, atoms: AtomData) -> System:
        if self._dead:
            raise RuntimeError("This NNAP is dead")
        t = self._types_of(atoms)
        key = (atoms.natoms, t.tobytes())
        s = self._systems.get(key)
        if s is not None:
            try:
                s.set_box(atoms.box)
            except NNAPError as e:
                if e.code != 2:  # 2: a box edge crossed the cutoff (another supercell): a new system
                    raise
                s.close()
                s = None
        if s is None:
            s = System(self, atoms.natoms, t, atoms.box, self.precision)
            self._systems = {key: s}   # keep the last configuration resident
        s.set_positions(atoms.xyz)
        return s

    def calEnergyForceVirials(self, atoms: AtomData, energies=True, forces=True, virials=True):
        what = ENERGY | (FORCE if forces else 0) | (VIRIAL if virials else 0)
        s = self.system(atoms)
        s.compute(what)
        return s.results(energies=energies, forces=forces or virials, virials=virials)

    def calEnergy(self, atoms: AtomData) -> float:
        s = self.system(atoms)
        s.compute(ENERGY)
        return s.results(energies=False, forces=False, virials=False)["energy"]

    def calEnergies(self, atoms: AtomData) -> np.ndarray:
        s = self.system(atoms)
        s.compute(ENERGY)
        return s.results(energies=True, forces=False, virials=False)["energies"]

    def calForces(self, atoms: AtomData) -> np.ndarray:
        return self.calEnergyForceVirials(atoms, energies=False, forces=True, virials=False)["forces"]

    def calStresses(self, atoms: AtomData) -> np.ndarray:
        """per-atom stress columns xx yy zz xy xz yz yx zx zy = -virial (IPotential.java:197-253)"""
        return -self.calEnergyForceVirials(atoms, energies=False, forces=False, virials=True)["virials"]

    def calStress(self, atoms: AtomData) -> List[float]:
        """[xx, yy, zz, xy, xz, yz] = -virial / volume (IPotential.java:264-311)"""
        v = self.calEnergyForceVirials(atoms, energies=False, forces=False, virials=True)["virial"]
        vol = atoms.volume()
        return [-v[0] / vol, -v[1] / vol, -v[2] / vol, -v[3] / vol, -v[4] / vol, -v[5] / vol]

    def batch(self, structures: Sequence[AtomData]) -> System:
        """A device-resident batch of independent replicas (positions set)."""
        per = [a.natoms for a in structures]
        types = np.concatenate([self._types_of(a) for a in structures]) if structures else np.zeros(0, np.int32)
        boxes = np.array([np.asarray(a.box, dtype=float) for a in structures])
        b = System(self, int(sum(per)), types, boxes, self.precision, natoms_per_replica=per)
        b.set_positions(np.concatenate([np.asarray(a.xyz, dtype=float) for a in structures]))
        return b

    def server(self, **kw) -> Server:
        """a continuous-batching server of device-resident replicas"""
        return Server(self, **kw)

    def calFp(self, atoms: AtomData) -> np.ndarray:
        """descriptors of every atom (jse_nnap_calFp, nnap_interface.cpp:11-24)"""
        return self.system(atoms).descriptors()
