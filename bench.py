#!/usr/bin/env python3
"""Benchmark of the NNAP evaluation path on one MI355X.

Workload (the headline configuration): ternary Cu-Zr-Al liquid, 32,000 atoms
in the ratio 50:40:10 at 0.058 A^-3 (random sequential addition, 2.0 A
minimum separation, ~53 neighbours within 6 A), 3-species spherical-Chebyshev
descriptor (nmax = 8, lmax = 6, wtype default -> 126 inputs) and a 64-64 SiLU
MLP per species with seeded random-init weights.  One step = one full
energy + force + virial evaluation of the configuration: cell-list neighbour
build, descriptor forward, MLP, backward, FP64 force/virial accumulation.
Positions are resident in HBM before the timed region starts.

Prints one JSON line.  ``--gpus N`` with N > 1 exits: this build runs on one
GPU (replicas and spatial domains would shard across GPUs with no change to
the per-GPU kernels).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from jse_amd import model as M  # noqa: E402
from jse_amd import nnap  # noqa: E402
from jse_amd import workloads as W  # noqa: E402

FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 peak
HBM_PEAK_GBS = 8000.0


def algorithmic_flops(mdl, types, pairs_per_species):
    """FLOPs per evaluation by kernel (multiply and add counted separately):
    descriptor forward: per pair 2PK (radial mixing) + 2PM (A), per atom 2PM (power spectrum);
    MLP: per atom 4 (S H1 + H1 H2 + H2) (forward + input gradient);
    descriptor backward: per pair 4PM (dE/dY, dE/dc) + 2PK + 4M (angles), per atom 2PM (dE/dA)."""
    out = {"desc_fwd": 0.0, "mlp": 0.0, "desc_bwd": 0.0}
    for t, sp in enumerate(mdl.species):
        na = int((types == t + 1).sum())
        npair = pairs_per_species[t]
        P, K, Mm, S = sp.P, sp.K, sp.M, sp.S
        out["desc_fwd"] += npair * (2 * P * K + 2 * P * Mm) + na * 2 * P * Mm
        dims = [S] + list(sp.hidden)
        mac = sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1)) + dims[-1]
        out["mlp"] += na * 4 * mac
        out["desc_bwd"] += npair * (4 * P * Mm + 2 * P * K + 4 * Mm) + na * 2 * P * Mm
    return out


def cpu_baseline(md, st, gpu_res, budget_s=12.0):
    """Times the FP64 C restatement (OpenMP, reference-like per-thread partial
    sums) on this host's cores, and states the GPU result's error against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    mdl = M.load_model(md)
    orc = O.Oracle(mdl, fast=True)
    cores = min(16, os.cpu_count() or 1)
    t0 = time.perf_counter()
    reps = 0
    ref = None
    while True:
        ref = O.system_efv_par(orc, st, cores)
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 50:
            break
    dt = time.perf_counter() - t0
    n = st.natoms
    err = {
        "energy_per_atom_eV": abs(gpu_res["energy"] - ref["energy"]) / n,
        "force_rms_eV_per_A": float(np.sqrt(((gpu_res["forces"] - ref["forces"]) ** 2).mean())),
        "force_max_eV_per_A": float(np.abs(gpu_res["forces"] - ref["forces"]).max()),
        "virial_per_atom_eV": float(np.abs(gpu_res["virial"] - ref["virial"]).max()) / n,
    }
    base = {"value": n * reps / dt, "unit": "atom-evals/s", "cores": cores, "kind": "port",
            "sample": f"{reps} full E+F+virial evaluations of the {n}-atom workload ({dt:.1f} s)"}
    return base, err


def batched_c4(nrep=512, seed=17, steps=10):
    """Batched variant: `nrep` independent replicas (log-uniform 64-4096 atoms,
    1-3 species, 0.05-0.065 A^-3) evaluated together with the BF16 MLP and a
    replayed step graph, against the plain path evaluated one replica per call
    (FP32, launches issued per call, synchronised per call)."""
    md = W.ternary_model()
    reps = W.replicas(nrep, 64, 4096, seed)
    natoms = sum(r.natoms for r in reps)
    atoms = [nnap.AtomData(r.box, r.xyz, r.types) for r in reps]
    out = {"replicas": nrep, "atoms": natoms}
    with nnap.NNAP(md, precision=nnap.BF16_MLP) as pot:
        b = pot.batch(atoms)
        ms, ph = b.benchmark(nnap.ENERGY | nnap.FORCE | nnap.VIRIAL, warmup=3, iters=steps)
        out["batched_bf16_graph"] = {"ms_per_step": ms, "atom_evals_per_s": natoms / (ms * 1e-3)}
        b.close()
    with nnap.NNAP(md, precision=nnap.FP32) as pot:
        systems = []
        for a in atoms:
            s_ = nnap.System(pot, a.natoms, pot._types_of(a), a.box, nnap.FP32)
            s_.set_positions(a.xyz)
            s_.set_graphs(False)
            systems.append(s_)
        for s_ in systems:          # warm-up (neighbour capacities settle)
            s_.compute()
            s_.results(energies=False, forces=False, virials=False)
        t0 = time.perf_counter()
        for s_ in systems:
            s_.compute()
            s_.results(energies=False, forces=False, virials=False)
        dt = time.perf_counter() - t0
        out["plain_fp32_one_replica_per_call"] = {"ms_per_pass": dt * 1e3, "atom_evals_per_s": natoms / dt}
        for s_ in systems:
            s_.close()
    out["speedup"] = out["batched_bf16_graph"]["atom_evals_per_s"] / out["plain_fp32_one_replica_per_call"][
        "atom_evals_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["fp32", "bf16"], default="fp32")
    ap.add_argument("--natoms", type=int, default=32000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(replica and domain sharding across GPUs is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    md = W.ternary_model()
    st = W.ternary_liquid(n=args.natoms, seed=7)
    prec = nnap.FP32 if args.precision == "fp32" else nnap.BF16_MLP
    pot = nnap.NNAP(md, precision=prec)
    sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
    what = nnap.ENERGY | nnap.FORCE | nnap.VIRIAL
    sysm.compute(what)
    res = sysm.results()
    stats = sysm.nl_stats()

    t0 = time.perf_counter()
    ms_step, phase = sysm.benchmark(what, warmup=args.warmup, iters=args.steps)
    wall = time.perf_counter() - t0

    # pairs per species of the centre atom (the kernels' work units)
    cnt, _, _ = sysm.neighbors(cap=stats["max_count"])
    pairs = [int(cnt[st.types == t + 1].sum()) for t in range(pot.ntypes())]
    flops = algorithmic_flops(pot.model, st.types, pairs)
    names = ["neighbor_build", "reverse_index", "desc_fwd", "mlp", "desc_bwd", "force_gather"]
    phase_ms = {k: float(v) for k, v in zip(names, phase)}
    dom = max(("desc_fwd", "mlp", "desc_bwd"), key=lambda k: phase_ms[k])
    achieved = flops[dom] / (phase_ms[dom] * 1e-3) / 1e12
    roof = {"bound": "mfma", "pipe": "fp32 vector ALU (MI355X quotes the same 157.3 TFLOP/s for FP32 vector and matrix)",
            "kernel": dom, "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
            "flop_per_step": {k: float(v) for k, v in flops.items()}}
    value = st.natoms / (ms_step * 1e-3)
    line = {
        "metric": "NNAP atom-evals/s (E+F+virial)",
        "value": value,
        "unit": "atom-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if prec == nnap.FP32 else "fp32+bf16-mlp",
        "data": "synthetic (seeded RSA liquid, seeded random-init model)",
        "config": {"workload": f"ternary Cu-Zr-Al liquid, {st.natoms} atoms, rho 0.058, rcut 6, nmax 8, lmax 6, "
                               "MLP 64-64 SiLU, plain FP32 path" if prec == nnap.FP32 else
                               f"ternary Cu-Zr-Al liquid, {st.natoms} atoms, BF16 MLP",
                   "natoms": st.natoms, "pairs": int(stats["pairs"]), "parallelism": "single GPU"},
        "phase_ms": phase_ms,
        "wall_s_timed": wall,
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        base, err = cpu_baseline(md, st, res)
        line["cpu_baseline"] = base
        line["error_vs_fp64"] = err
    if not args.no_batched:
        line["batched_512_replicas"] = batched_c4()
    print(json.dumps(line), flush=True)
    pot.close()


if __name__ == "__main__":
    main()
