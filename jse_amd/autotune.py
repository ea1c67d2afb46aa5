"""Launch-shape autotuner with an on-disk cache.

``tune(system)`` looks up the system's shape signature (precision, per-species
descriptor and MLP shapes, atoms and replicas rounded to powers of two, mean
neighbours per atom rounded to 8) in ``_tuned.json`` beside this file.  On a
hit it applies the cached launch shapes; otherwise it runs the native
autotuner (jse_nnap_system_autotune: coordinate descent over the descriptor,
MLP, neighbour-build and gather launch shapes, each candidate timed with HIP
events on the system's own data) and stores the winner.
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Tuple

from . import nnap

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_tuned.json")
# bumped whenever a knob's meaning or range changes, so older entries are not applied
SCHEMA = 4


def _p2(v: float) -> int:
    return 1 << max(0, int(round(math.log2(max(1.0, v)))))


def signature(system: "nnap.System", what: int) -> str:
    pot = system.nnap
    st = system.nl_stats()
    mean_nb = st["pairs"] / max(1, system.natoms)
    sp = ";".join(f"P{s.P}K{s.K}L{s.lmax}H{'x'.join(map(str, s.hidden))}" for s in pot.model.species)
    return (f"v{SCHEMA}|prec{pot.precision}|what{what}|{sp}|atoms{_p2(system.natoms)}|reps{_p2(system.nrep)}"
            f"|nb{8 * round(mean_nb / 8)}")


def _load() -> Dict[str, dict]:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def tune(system: "nnap.System", what: int = nnap.ENERGY | nnap.FORCE | nnap.VIRIAL, iters: int = 5,
         use_cache: bool = True) -> Tuple[Dict[str, int], float, bool]:
    """Returns (knobs, tuned ms per evaluation or None on a cache hit, cache hit)."""
    system.compute(what)
    system.results(energies=False, forces=False, virials=False)
    key = signature(system, what)
    cache = _load() if use_cache else {}
    if key in cache:
        system.set_tuning(cache[key]["knobs"])
        return cache[key]["knobs"], None, True
    knobs, ms = system.autotune(what, iters)
    cache = _load()
    cache[key] = {"knobs": knobs, "ms": ms}
    try:
        with open(CACHE + ".tmp", "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(CACHE + ".tmp", CACHE)
    except OSError:
        pass
    return knobs, ms, False
