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

Prints one JSON line.  ``--gpus N`` with N > 1 exits: the benchmark runs on
one GPU (replicas and spatial domains would shard across GPUs with no change
to the per-GPU kernels).
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

FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 peak (matrix cores; vector with packed-FP32 issue)
FP64_PEAK_TFLOPS = 78.6    # MI355X published FP64 peak (matrix cores and vector)
BF16_PEAK_TFLOPS = 2500.0  # MI355X published dense BF16 matrix peak


def roofline_pipe(kernel: str, precision: str):
    """(pipe, published peak in TFLOP/s, bound) of the pipe a kernel's counted
    arithmetic runs on in a precision mode ("fp32" = the plain mode, "fast",
    "bf16")."""
    if kernel == "desc_fwd":
        if precision == "fp32":
            return "fp64 matrix cores (v_mfma_f64_16x16x4_f64)", FP64_PEAK_TFLOPS, "mfma"
        return "fp32 matrix cores (v_mfma_f32_16x16x4_f32)", FP32_PEAK_TFLOPS, "mfma"
    if kernel == "mlp":
        if precision == "fp32":
            return "fp64 matrix cores (v_mfma_f64_16x16x4_f64)", FP64_PEAK_TFLOPS, "mfma"
        if precision == "fast":
            return "fp32 matrix cores (v_mfma_f32_16x16x4_f32)", FP32_PEAK_TFLOPS, "mfma"
        return "bf16 matrix cores (v_mfma_f32_16x16x32_bf16)", BF16_PEAK_TFLOPS, "mfma"
    # the backward's counted work is FP32 FMA on the vector ALU (its FP64 geometry is not counted)
    return "fp32 vector ALU (FMA; the published FP32 peak assumes packed issue)", FP32_PEAK_TFLOPS, "valu"
HBM_PEAK_GBS = 8000.0
# Launch shapes of the 32,000-atom workload per precision mode, as the
# autotuner (jse_amd/autotune.py) chose them on an MI355X; the line and the
# committed profiles use the same shapes (--autotune tunes afresh instead).
TUNED_C2 = {
    "fp32": {"desc_fwd_wg_per_cu": 32, "desc_bwd_wg_per_cu": 16, "mlp_wg_per_cu": 0, "nl_wg_per_cu": 0,
             "gather_waves": 8, "desc_fwd_variant": 0, "desc_bwd_variant": 0},
    "fast": {"desc_fwd_wg_per_cu": 32, "desc_bwd_wg_per_cu": 16, "mlp_wg_per_cu": 8, "nl_wg_per_cu": 0,
             "gather_waves": 8, "desc_fwd_variant": 0, "desc_bwd_variant": 2},
    "bf16": {"desc_fwd_wg_per_cu": 32, "desc_bwd_wg_per_cu": 16, "mlp_wg_per_cu": 4, "nl_wg_per_cu": 0,
             "gather_waves": 0, "desc_fwd_variant": 0, "desc_bwd_variant": 0},
}
# the batched 512-replica step's shapes (BF16 mode), for profiling that leg
TUNED_C4 = {
    "bf16": {"desc_fwd_wg_per_cu": 32, "desc_bwd_wg_per_cu": 32, "mlp_wg_per_cu": 8, "nl_wg_per_cu": 16,
             "gather_waves": 8, "desc_fwd_variant": 0, "desc_bwd_variant": 0},
}


def algorithmic_flops(mdl, types, pairs_per_species):
    """FLOPs per evaluation by kernel (multiply and add counted separately):
    descriptor forward: per pair 2PK (radial mixing) + 2PM (A), per atom 2PM (power spectrum);
    MLP: per atom 4 (S H1 + H1 H2 + H2) (forward + input gradient);
    descriptor backward (dE/dA folded through each species' mixing first): per pair 4KM (dE/dY,
    dE/dRn) + 2K + 4M (radial chain, angles), per atom 2PM (dE/dA) + 2 T K M (default mixing) or
    2 T P K M (dense) for the folding over T species."""
    out = {"desc_fwd": 0.0, "mlp": 0.0, "desc_bwd": 0.0}
    for t, sp in enumerate(mdl.species):
        na = int((types == t + 1).sum())
        npair = pairs_per_species[t]
        P, K, Mm, S = sp.P, sp.K, sp.M, sp.S
        out["desc_fwd"] += npair * (2 * P * K + 2 * P * Mm) + na * 2 * P * Mm
        dims = [S] + list(sp.hidden)
        mac = sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1)) + dims[-1]
        out["mlp"] += na * 4 * mac
        T = mdl.ntypes
        fold = 0 if P == K else (2 * T * K * Mm if P == 2 * K else 2 * T * P * K * Mm)
        out["desc_bwd"] += npair * (4 * K * Mm + 2 * K + 4 * Mm) + na * (2 * P * Mm + fold)
    return out


def cpu_baseline(md, st, gpu_res, budget_s=12.0):
    """Times the FP64 C restatement (OpenMP, reference-like per-thread partial
    sums) on this host's cores, and states the GPU result's error against it."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    mdl = M.load_model(md)
    orc = O.Oracle(mdl, fast=True)
    cores = host_cores()
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
            "host_cpus_visible": os.cpu_count(),
            "sample": f"{reps} full E+F+virial evaluations of the {n}-atom workload ({dt:.1f} s), "
                      f"one OpenMP thread per core of this process's CPU share"}
    return base, err


def host_cores():
    """Threads for the CPU baseline: the cores this process may use (its CPU
    affinity), at most the 16-core share of one GPU on a shared host."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cu_c1(steps=50, budget_s=10.0):
    """The Cu FCC 256-atom crystal (a = 3.61 A, Gaussian displacements 0.1 A,
    seed 1) with the one-species nmax 5 / lmax 6 / 32-32 model: the plain
    path's time per E+F+virial evaluation on the GPU and the FP64 CPU
    restatement's on the host cores, with the GPU result's error."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    md = W.cu_model()
    st = W.cu_fcc(seed=1)
    what = nnap.ENERGY | nnap.FORCE | nnap.VIRIAL
    with nnap.NNAP(md) as pot:
        s_ = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
        s_.compute(what)
        r = s_.results()
        ms, _ = s_.benchmark(what, warmup=5, iters=steps)
    orc = O.Oracle(M.load_model(md), fast=True)
    cores = host_cores()
    t0 = time.perf_counter()
    reps = 0
    while True:
        ref = O.system_efv_par(orc, st, cores)
        reps += 1
        if time.perf_counter() - t0 > budget_s or reps >= 20000:
            break
    dt = time.perf_counter() - t0
    n = st.natoms
    return {"natoms": n, "gpu_ms_per_step": ms, "gpu_atom_evals_per_s": n / (ms * 1e-3),
            "cpu_baseline": {"value": n * reps / dt, "unit": "atom-evals/s", "cores": cores, "kind": "port",
                             "sample": f"{reps} evaluations ({dt:.1f} s)"},
            "error_vs_fp64": {"energy_per_atom_eV": abs(r["energy"] - ref["energy"]) / n,
                              "force_max_eV_per_A": float(np.abs(r["forces"] - ref["forces"]).max()),
                              "virial_per_atom_eV": float(np.abs(r["virial"] - ref["virial"]).max()) / n},
            "note": "a 256-atom system fills a small fraction of the GPU; the time is launch-bound"}


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
    from jse_amd import autotune
    what = nnap.ENERGY | nnap.FORCE | nnap.VIRIAL
    with nnap.NNAP(md, precision=nnap.BF16_MLP) as pot:
        b = pot.batch(atoms)
        ms0, _ = b.benchmark(what, warmup=3, iters=steps)
        knobs, _, hit = autotune.tune(b, what, use_cache=False)
        ms, ph = b.benchmark(what, warmup=3, iters=steps)
        out["batched_bf16_graph"] = {"ms_per_step": ms, "atom_evals_per_s": natoms / (ms * 1e-3),
                                     "default_launch_ms_per_step": ms0, "tuned_launch_shapes": knobs}
        b.close()
    for name, prec in (("batched_plain_graph", nnap.FP32), ("batched_fp32_fast_graph", nnap.FP32_FAST)):
        with nnap.NNAP(md, precision=prec) as pot:
            b = pot.batch(atoms)
            autotune.tune(b, what, use_cache=False)
            ms, _ = b.benchmark(what, warmup=3, iters=steps)
            out[name] = {"ms_per_step": ms, "atom_evals_per_s": natoms / (ms * 1e-3)}
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
    out["bf16_vs_plain_batched"] = out["batched_bf16_graph"]["atom_evals_per_s"] / out["batched_plain_graph"][
        "atom_evals_per_s"]
    return out


def batched_leg(prec_name, knobs, steps):
    """Only the batched graph step of the 512-replica workload in one precision
    mode at the given launch shapes (for profiling runs)."""
    md = W.ternary_model()
    reps = W.replicas(512, 64, 4096, 17)
    natoms = sum(r.natoms for r in reps)
    atoms = [nnap.AtomData(r.box, r.xyz, r.types) for r in reps]
    prec = {"fp32": nnap.FP32, "fast": nnap.FP32_FAST, "bf16": nnap.BF16_MLP}[prec_name]
    what = nnap.ENERGY | nnap.FORCE | nnap.VIRIAL
    with nnap.NNAP(md, precision=prec) as pot:
        b = pot.batch(atoms)
        if knobs:
            b.set_tuning(knobs)
        ms, _ = b.benchmark(what, warmup=3, iters=steps)
        b.close()
    print(json.dumps({"leg": "batched_512_replicas", "precision": prec_name, "atoms": natoms, "ms_per_step": ms,
                      "atom_evals_per_s": natoms / (ms * 1e-3), "launch_shapes": knobs}), flush=True)


def precision_c3(n=256000, seed=7, steps=10):
    """Precision variant at 256,000 atoms (same liquid recipe, L = 164.035 A):
    the BF16-MLP path's throughput and its energy / force / virial errors
    against the FP64 restatement on the same inputs, beside the plain FP32
    path's errors."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    md = W.ternary_model()
    st = W.ternary_liquid(n=n, seed=seed)
    ref = O.system_efv_par(O.Oracle(M.load_model(md), fast=True), st, host_cores())
    out = {"natoms": n, "box_A": float(st.box[0])}
    names = ["neighbor_build", "reverse_index", "desc_fwd", "mlp", "desc_bwd", "force_gather"]
    mdl = M.load_model(md)
    for name, prec in (("bf16_mlp", nnap.BF16_MLP), ("fp32", nnap.FP32), ("fp32_fast", nnap.FP32_FAST)):
        with nnap.NNAP(md, precision=prec) as pot:
            s_ = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
            s_.compute(nnap.ENERGY | nnap.FORCE | nnap.VIRIAL)
            r = s_.results()
            ms, ph = s_.benchmark(nnap.ENERGY | nnap.FORCE | nnap.VIRIAL, warmup=2, iters=steps)
            stats = s_.nl_stats()
            cnt, _, _ = s_.neighbors(cap=stats["max_count"])
            pairs = [int(cnt[st.types == t + 1].sum()) for t in range(pot.ntypes())]
        df = r["forces"] - ref["forces"]
        phase_ms = {k: float(v) for k, v in zip(names, ph)}
        mlp_flops = algorithmic_flops(mdl, st.types, pairs)["mlp"]
        pipe, peak, _ = roofline_pipe("mlp", {"bf16_mlp": "bf16", "fp32": "fp32", "fp32_fast": "fast"}[name])
        out[name] = {"ms_per_step": ms, "atom_evals_per_s": n / (ms * 1e-3), "phase_ms": phase_ms,
                     "mlp_matrix_core_utilisation": {
                         "achieved_TFLOPs": mlp_flops / (phase_ms["mlp"] * 1e-3) / 1e12, "peak_TFLOPs": peak,
                         "frac": mlp_flops / (phase_ms["mlp"] * 1e-3) / 1e12 / peak, "pipe": pipe},
                     "energy_per_atom_eV": abs(r["energy"] - ref["energy"]) / n,
                     "energy_rmse_per_atom_eV": float(np.sqrt(((r["energies"] - ref["energies"]) ** 2).mean())),
                     "force_rms_eV_per_A": float(np.sqrt((df ** 2).mean())),
                     "force_max_eV_per_A": float(np.abs(df).max()),
                     "virial_per_atom_eV": float(np.abs(r["virial"] - ref["virial"]).max()) / n}
    out["bf16_speedup_vs_fp32"] = out["bf16_mlp"]["atom_evals_per_s"] / out["fp32"]["atom_evals_per_s"]
    out["note"] = ("fp32 = the plain path (FP64 geometry, descriptor and MLP; meets the reference bar); "
                   "fp32_fast = FP32 geometry and descriptor with an FP32 MLP; bf16_mlp = FP32 geometry and "
                   "descriptor with a BF16 MLP; the BF16 MLP's forward products carry hi + lo BF16 pairs")
    return out


def pcie_inclusive(pot, st, iters=20):
    """Host positions in, host energies + forces + virial out, per call."""
    a = nnap.AtomData(st.box, st.xyz, st.types)
    s_ = pot.system(a)
    for _ in range(3):
        s_.set_positions(st.xyz)
        s_.compute(nnap.ENERGY | nnap.FORCE | nnap.VIRIAL)
        s_.results(energies=True, forces=True, virials=False)
    t0 = time.perf_counter()
    for _ in range(iters):
        s_.set_positions(st.xyz)
        s_.compute(nnap.ENERGY | nnap.FORCE | nnap.VIRIAL)
        s_.results(energies=True, forces=True, virials=False)
    dt = (time.perf_counter() - t0) / iters
    return {"ms_per_call": dt * 1e3, "atom_evals_per_s": st.natoms / dt}


def gpumd_hook(pot, st, iters=20):
    """The GPUMD hook (jse_nnap_computeGPUMD, packed arguments) on the same
    configuration: GPUMD's device-resident full neighbour list (column-major,
    FP32 minimum-image displacements within rc + 0.5 A, built once), forces,
    per-atom virials and energies added into FP64 device arrays.  Each call is
    synchronous, as the reference's; the time per call includes the hook's
    species grouping (a copy of the centres' types) and capacity check."""
    import ctypes
    import struct
    from scipy.spatial import cKDTree
    lib = nnap.lib()
    for name in ("jse_nnap_gpumd_bind", "jse_nnap_computeGPUMD"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
    L = np.asarray(st.box, float)
    x = st.xyz - np.floor(st.xyz / L) * L
    n = st.natoms
    rg = pot.rcutMax() + 0.5
    pairs = cKDTree(x, boxsize=L).query_pairs(rg, output_type="ndarray")
    i_all = np.concatenate([pairs[:, 0], pairs[:, 1]])
    j_all = np.concatenate([pairs[:, 1], pairs[:, 0]])
    order = np.lexsort((j_all, i_all))
    i_all, j_all = i_all[order], j_all[order]
    num = np.bincount(i_all, minlength=n).astype(np.int32)
    slot = np.arange(len(i_all)) - np.repeat(np.cumsum(num) - num, num)
    mn = int(num.max())
    nbr = np.zeros((mn, n), dtype=np.int32)
    nbr[slot, i_all] = j_all
    dv = x[j_all] - x[i_all]
    dv -= np.round(dv / L) * L
    d = np.zeros((3, mn, n), dtype=np.float32)
    for c in range(3):
        d[c, slot, i_all] = dv[:, c]
    # device buffers from the HIP runtime the library itself uses (as GPUMD's would be)
    hip = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    bufs = []

    def dev(a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(p), max(1, a.nbytes)) == 0
        assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        bufs.append(p)
        return p.value
    ptrs = [dev(num), dev(nbr)] + [dev(d[c]) for c in range(3)] + [dev((st.types - 1).astype(np.int32))] + [0] * 7 + \
        [dev(np.zeros(n)) for _ in range(3)] + [dev(np.zeros(9 * n)), dev(np.zeros(n))] + [0] * 9
    assert lib.jse_nnap_gpumd_bind(ctypes.create_string_buffer(struct.pack("<Qi", pot._h.value, nnap.FP32))) == 0
    blk = ctypes.create_string_buffer(struct.pack("<iii" + "Q" * len(ptrs), n, 0, n, *ptrs))
    for _ in range(3):
        assert lib.jse_nnap_computeGPUMD(blk) == 0, lib.jse_nnap_last_error()
    t0 = time.perf_counter()
    for _ in range(iters):
        assert lib.jse_nnap_computeGPUMD(blk) == 0, lib.jse_nnap_last_error()
    dt = (time.perf_counter() - t0) / iters
    for p in bufs:
        hip.hipFree(p)
    return {"ms_per_call": dt * 1e3, "atom_evals_per_s": n / dt, "pairs_listed": int(num.sum()),
            "note": "device-resident GPUMD neighbour list; rows grouped by species on each call"}


def copy_calibration(nbytes, reps=5):
    """A plain device-to-device copy of `nbytes` (torch's copy kernel), the
    reference point that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    y = torch.empty_like(x)
    for _ in range(reps):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": x.numel() * 4, "reps": reps}), flush=True)


def untiled_shapes(natoms, steps, warmup):
    """Phase times (ms) of E+F+virial on the ternary liquid for model shapes
    without tiled descriptor kernels, beside the tiled nmax 8, lmax 6,
    default-mixing shape; networks 64-64, seeded random weights, default
    launch shapes (no autotuning)."""
    st = W.ternary_liquid(n=natoms, seed=7)
    sample = W.ternary_liquid(n=2000, seed=7)
    names = ["neighbor_build", "reverse_index", "desc_fwd", "mlp", "desc_bwd", "force_gather"]
    what = nnap.ENERGY | nnap.FORCE | nnap.VIRIAL
    for nmax, lmax, wtype in ((8, 6, "default"), (8, 6, "full"), (8, 6, "exfull"), (5, 8, "default")):
        md = M.random_model(["Cu", "Zr", "Al"], nmax=nmax, lmax=lmax, rcut=6.0, hidden=(64, 64), seed=13,
                            wtype=wtype)
        md = W._normalized(md, sample, 300, 13)
        with nnap.NNAP(md) as pot:
            sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
            sysm.compute(what)
            ms, ph = sysm.benchmark(what, warmup=warmup, iters=steps)
            sp = pot.model.species[0]
            print(json.dumps({"nmax": nmax, "lmax": lmax, "wtype": wtype, "P": sp.P, "ms_per_step": ms,
                              "atom_evals_per_s": natoms / (ms * 1e-3),
                              "phase_ms": {k: float(v) for k, v in zip(names, ph)}}), flush=True)


# kernel of each timed phase, by name prefix (the plain mode's; the traffic
# summary covers that mode)
PHASE_KERNEL = {"desc_fwd": "jnn::k_desc_fwd_acc", "mlp": "jnn::k_mlp2_mfma_multi", "desc_bwd": "jnn::k_desc_bwd",
                "neighbor_build": "jnn::k_nl_build_cells", "force_gather": "jnn::k_gather"}


def committed_traffic(phase):
    """HBM bytes per launch of the phase's kernel in the plain mode, from the
    committed counter summary profiles/ternary32k_plain_traffic.md (lines
    'traffic_bytes_per_launch[<kernel>]: N', FETCH_SIZE and WRITE_SIZE scaled
    by a device copy of known size)."""
    import re
    path = os.path.join(ROOT, "profiles", "ternary32k_plain_traffic.md")
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None
    for m in re.finditer(r"traffic_bytes_per_launch\[([^\]]+)\]:\s*([0-9.eE+]+)", text):
        if m.group(1).startswith(PHASE_KERNEL.get(phase, "?")):
            return float(m.group(2))
    return None


def nl_bytes(natoms, npairs, fp64_rows):
    """Algorithmic HBM bytes of the neighbour build: per atom the position read
    and the wrapped position written (48), the cell id and rank (8), the
    cell-sorted copy read and written (56), the count (4); per pair the key
    and the FP32 displacement with its tag (20), plus the FP64 low parts (16)
    in the plain mode."""
    return natoms * 116 + npairs * (36 if fp64_rows else 20)


def gather_bytes(natoms, npairs, fp64_rows):
    """Algorithmic HBM bytes of the force/virial gather: per pair the own g,
    key and displacement (36, plus 16 for the FP64 low parts in the plain
    mode) and the reverse pair's g (16); per atom F and V written (96)."""
    return npairs * (52 + (16 if fp64_rows else 0)) + natoms * 96


def kernel_rooflines(flops, phase_ms, pmode, natoms, npairs, fp64_rows):
    """Achieved work rate of every phase against the peak of the pipe that
    bounds it: GB/s against HBM for the neighbour build and the gather,
    TFLOP/s against the matrix-core or vector peak for the others."""
    out = {}
    for k, b in (("neighbor_build", nl_bytes(natoms, npairs, fp64_rows)),
                 ("force_gather", gather_bytes(natoms, npairs, fp64_rows))):
        gbs = b / (phase_ms[k] * 1e-3) / 1e9
        out[k] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                  "algorithmic_bytes": b}
    for k in ("desc_fwd", "mlp", "desc_bwd"):
        pipe, peak, bound = roofline_pipe(k, pmode)
        tf = flops[k] / (phase_ms[k] * 1e-3) / 1e12
        out[k] = {"bound": bound, "pipe": pipe, "achieved": tf, "peak": peak, "unit": "TFLOP/s", "frac": tf / peak}
    if pmode == "fp32":  # measured HBM bytes per launch (committed counter summary of this mode)
        for k in out:
            out[k]["traffic"] = committed_traffic(k)
    return out


def other_modes_c2(md, st):
    """The FP32_FAST and BF16-MLP modes on the same 32,000-atom system
    (autotuned launch shapes), with their errors against the FP64 CPU
    restatement."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from jse_amd import autotune
    ref = O.system_efv_par(O.Oracle(M.load_model(md), fast=True), st, host_cores())
    what = nnap.ENERGY | nnap.FORCE | nnap.VIRIAL
    names = ["neighbor_build", "reverse_index", "desc_fwd", "mlp", "desc_bwd", "force_gather"]
    out = {}
    for name, prec in (("fp32_fast", nnap.FP32_FAST), ("bf16_mlp", nnap.BF16_MLP)):
        with nnap.NNAP(md, precision=prec) as pot:
            s_ = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
            s_.compute(what)
            r = s_.results()
            knobs, _, _ = autotune.tune(s_, what, use_cache=False)
            ms, ph = s_.benchmark(what, warmup=3, iters=20)
        n = st.natoms
        out[name] = {"ms_per_step": ms, "atom_evals_per_s": n / (ms * 1e-3), "launch_shapes": knobs,
                     "phase_ms": {k: float(v) for k, v in zip(names, ph)},
                     "energy_per_atom_eV": abs(r["energy"] - ref["energy"]) / n,
                     "force_rms_eV_per_A": float(np.sqrt(((r["forces"] - ref["forces"]) ** 2).mean())),
                     "force_max_eV_per_A": float(np.abs(r["forces"] - ref["forces"]).max()),
                     "virial_per_atom_eV": float(np.abs(r["virial"] - ref["virial"]).max()) / n}
    return out


def _pct(a, q):
    return float(np.percentile(np.asarray(a), q)) if len(a) else None


def served_c5(nrep=2000, nmin=100, nmax=20000, seed=23, sat_s=8.0, rate_s=6.0, fractions=(0.5, 0.8),
              budget_gb=80.0, max_batch=262144):
    """Served variant: `nrep` replicas (log-uniform nmin-nmax atoms, 1-3
    species, 0.05-0.065 A^-3) resident on the device; every request moves a
    replica's atoms by N(0, 0.01 A) per component and asks for E+F+virial
    (BF16 MLP, Verlet skin 1.0 A).  Saturation throughput from a closed loop
    with two outstanding requests per replica, then open-loop Poisson
    arrivals (replica drawn uniformly) at fractions of the saturation
    request rate, with p50/p99 latency (submission -> results on the host)
    and the generator's lag behind the Poisson schedule."""
    md = W.ternary_model()
    reps = W.replicas(nrep, nmin, nmax, seed, workers=min(16, os.cpu_count() or 1))
    natoms = np.array([r.natoms for r in reps])
    rng = np.random.default_rng(seed)
    noise = rng.normal(0.0, 0.01, size=1 << 24)
    pos = [r.xyz.copy() for r in reps]

    def step(k):
        n3 = 3 * natoms[k]
        o = int(rng.integers(0, noise.size - n3))
        pos[k] += noise[o:o + n3].reshape(-1, 3)
        return pos[k]

    out = {"replicas": nrep, "resident_atoms": int(natoms.sum()), "sizes": [nmin, nmax], "skin_A": 1.0,
           "step_sigma_A": 0.01, "precision": "bf16-mlp"}
    with nnap.NNAP(md) as pot, pot.server(precision=nnap.BF16_MLP, skin=1.0, hbm_budget_bytes=budget_gb * 1e9,
                                           max_batch_atoms=max_batch, max_replicas=nrep,
                                           max_density=0.065) as srv:
        rids = [srv.add_replica(nnap.AtomData(r.box, r.xyz, r.types)) for r in reps]
        t0 = time.perf_counter()
        for k in range(nrep):
            srv.submit(rids[k], pos[k])
        srv.drain()
        srv.collect(nrep + 16)
        out["admission_s"] = time.perf_counter() - t0
        print(f"served: {nrep} replicas admitted in {out['admission_s']:.2f} s", file=sys.stderr, flush=True)

        # closed loop, two requests in flight per replica
        owner = {}
        for k in range(nrep):
            for _ in range(2):
                owner[srv.submit(rids[k], step(k))] = k
        done_atoms = 0
        lat = []
        t0 = time.perf_counter()
        t_end = t0 + sat_s
        while time.perf_counter() < t_end:
            got = srv.collect(8192)
            if not got:
                time.sleep(0.0002)
                continue
            for tk, r in got.items():
                k = owner.pop(tk)
                done_atoms += natoms[k]
                lat.append(r["latency_s"])
                owner[srv.submit(rids[k], step(k))] = k
        dt = time.perf_counter() - t0
        srv.drain()
        srv.collect(1 << 16)
        sat_rate = done_atoms / dt
        req_rate = len(lat) / dt
        st = srv.stats()
        out["saturation"] = {"atom_evals_per_s": sat_rate, "requests_per_s": req_rate, "window_s": dt,
                             "p50_ms": _pct(lat, 50) * 1e3, "p99_ms": _pct(lat, 99) * 1e3,
                             "mean_batch_atoms": st["mean_batch_atoms"]}
        print(f"served: saturation {sat_rate:.3e} atom-evals/s, {req_rate:.0f} req/s", file=sys.stderr, flush=True)

        # open loop: Poisson arrivals at fractions of the saturation request rate
        out["poisson"] = []
        for f in fractions:
            lam = f * req_rate
            lat, lag, atoms_done = [], [], 0
            owner = {}
            t0 = time.perf_counter()
            t_end = t0 + rate_s
            t_next = t0 + rng.exponential(1.0 / lam)
            while True:
                now = time.perf_counter()
                if now >= t_end:
                    break
                while t_next <= now:
                    k = int(rng.integers(0, nrep))
                    owner[srv.submit(rids[k], step(k))] = k
                    lag.append(time.perf_counter() - t_next)
                    t_next += rng.exponential(1.0 / lam)
                for tk, r in srv.collect(8192).items():
                    atoms_done += natoms[owner.pop(tk)]
                    lat.append(r["latency_s"])
                wait = min(t_next, t_end) - time.perf_counter()
                if wait > 0:
                    time.sleep(min(wait, 0.0005))
            srv.drain()
            for tk, r in srv.collect(1 << 16).items():
                atoms_done += natoms[owner.pop(tk)]
                lat.append(r["latency_s"])
            span = time.perf_counter() - t0
            out["poisson"].append({"offered_fraction_of_saturation": f, "offered_requests_per_s": lam,
                                   "requests": len(lat), "atom_evals_per_s": atoms_done / span,
                                   "p50_ms": _pct(lat, 50) * 1e3, "p99_ms": _pct(lat, 99) * 1e3,
                                   "generator_lag_p99_ms": _pct(lag, 99) * 1e3})
            print(f"served: poisson {f:.2f} -> p50 {out['poisson'][-1]['p50_ms']:.2f} ms "
                  f"p99 {out['poisson'][-1]['p99_ms']:.2f} ms", file=sys.stderr, flush=True)
        st = srv.stats()
        out["server"] = {k: st[k] for k in ("batches", "mean_batch_atoms", "rebuilds", "admissions", "evictions",
                                            "growths", "resident_atoms", "candidate_cap", "row_cap")}
        out["server"]["device_GB"] = st["device_bytes"] / 1e9
    return out


def training_derivatives(pot, st, iters=5, minibatch=True):
    """The training entry points on the same configuration (plain mode), per
    call from host arrays to host gradient blocks: the energy loss'
    derivative (jse_nnap_system_backward_energy) and the force / stress
    loss' (jse_nnap_system_backward_energy_force), each running its own
    neighbour build, descriptors and network passes.  Atom-evals/s counts one
    atom per atom and call."""
    a = nnap.AtomData(st.box, st.xyz, st.types)
    s_ = pot.system(a)
    rng = np.random.default_rng(0)
    w = rng.uniform(0.5, 1.5, size=st.natoms)
    bf = rng.normal(size=(st.natoms, 3))
    bv = np.array([0.3, -0.2, 0.1, 0.05, -0.07, 0.02, 0.0, 0.0, 0.0])
    out = {}
    for name, call in (("energy_loss", lambda: s_.backward_energy(w)),
                       ("force_loss", lambda: s_.backward_energy_force(w, bf, bv))):
        call()
        t0 = time.perf_counter()
        for _ in range(iters):
            call()
        dt = (time.perf_counter() - t0) / iters
        out[name] = {"ms_per_call": dt * 1e3, "atom_evals_per_s": st.natoms / dt}
    out["note"] = "parameter gradients summed over the 32,000 atoms; host upstream in, host gradient blocks out"
    if not minibatch:
        return out
    # a training mini-batch: 256 small cells (8-64 atoms, boxes from 4.9 A up, the
    # short ones held as supercells) in one batch, one force / stress loss call
    reps = W.replicas(256, 8, 64, 23, rho_range=(0.055, 0.065))
    nat = sum(r.natoms for r in reps)
    b = pot.batch([nnap.AtomData(r.box, r.xyz, r.types) for r in reps])
    wb = rng.uniform(0.5, 1.5, size=nat)
    bfb = rng.normal(size=(nat, 3))
    bvb = rng.normal(size=(len(reps), 9))
    b.backward_energy_force(wb, bfb, bvb)
    t0 = time.perf_counter()
    for _ in range(iters):
        b.backward_energy_force(wb, bfb, bvb)
    dt = (time.perf_counter() - t0) / iters
    b.close()
    short = sum(1 for r in reps if min(r.box) < pot.rcutMax())
    out["force_loss_minibatch_256_cells"] = {
        "ms_per_call": dt * 1e3, "structures_per_s": len(reps) / dt, "atom_evals_per_s": nat / dt,
        "atoms": nat, "cells_below_cutoff": short}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=["fp32", "fast", "bf16"], default="fp32")
    ap.add_argument("--natoms", type=int, default=32000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-256k", action="store_true")
    ap.add_argument("--no-modes", action="store_true", help="skip the FP32_FAST / BF16 lines at 32,000 atoms")
    ap.add_argument("--no-c1", action="store_true", help="skip the Cu-256 line")
    ap.add_argument("--no-hooks", action="store_true", help="skip the PCIe-inclusive and GPUMD-hook lines")
    ap.add_argument("--no-training", action="store_true", help="skip the training-derivative line")
    ap.add_argument("--served-replicas", type=int, default=2000)
    ap.add_argument("--knobs", type=str, default="",
                    help="JSON launch shapes for the 32,000-atom system (default: the committed tuned shapes "
                         "of the precision mode)")
    ap.add_argument("--autotune", action="store_true", help="autotune the 32,000-atom system's launch shapes")
    ap.add_argument("--shapes", action="store_true",
                    help="only time model shapes of the untiled descriptor kernels on the 32,000-atom liquid "
                         "(one JSON line per shape, the tiled default shape first)")
    ap.add_argument("--leg", choices=["batched", "training"], default="",
                    help="batched: only the batched graph step of the 512-replica workload (with --precision and "
                         "--knobs); training: only the training entry points on the 32,000-atom liquid "
                         "(plain mode, --steps calls each, no mini-batch); for profiling")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a plain device copy of this many bytes (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on exactly one GPU; --gpus {args.gpus} is not supported "
              "(replicas would shard across GPUs by HBM budget, one large system by spatial bricks)",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_bytes:
        copy_calibration(args.copy_bytes)
        return
    if args.shapes:
        untiled_shapes(args.natoms, args.steps, args.warmup)
        return
    if args.leg == "batched":
        knobs = json.loads(args.knobs) if args.knobs else TUNED_C4.get(args.precision)
        batched_leg(args.precision, knobs, args.steps)
        return
    if args.leg == "training":
        st = W.ternary_liquid(n=args.natoms, seed=7)
        with nnap.NNAP(W.ternary_model(), precision=nnap.FP32) as pot:
            out = training_derivatives(pot, st, iters=args.steps, minibatch=False)
        print(json.dumps({"leg": "training_derivatives", "atoms": st.natoms, **out}), flush=True)
        return

    md = W.ternary_model()
    st = W.ternary_liquid(n=args.natoms, seed=7)
    prec = {"fp32": nnap.FP32, "fast": nnap.FP32_FAST, "bf16": nnap.BF16_MLP}[args.precision]
    pot = nnap.NNAP(md, precision=prec)
    sysm = pot.system(nnap.AtomData(st.box, st.xyz, st.types))
    what = nnap.ENERGY | nnap.FORCE | nnap.VIRIAL
    sysm.compute(what)
    res = sysm.results()
    stats = sysm.nl_stats()

    ms_default = None
    if args.autotune:
        ms_default, _ = sysm.benchmark(what, warmup=2, iters=5)
        from jse_amd import autotune
        knobs, _, _ = autotune.tune(sysm, what, use_cache=False)
    else:
        knobs = json.loads(args.knobs) if args.knobs else TUNED_C2[args.precision]
        sysm.set_tuning(knobs)
    t0 = time.perf_counter()
    ms_step, phase = sysm.benchmark(what, warmup=args.warmup, iters=args.steps)
    wall = time.perf_counter() - t0

    # pairs per species of the centre atom (the kernels' work units)
    cnt, _, _ = sysm.neighbors(cap=stats["max_count"])
    pairs = [int(cnt[st.types == t + 1].sum()) for t in range(pot.ntypes())]
    npairs = int(sum(pairs))
    flops = algorithmic_flops(pot.model, st.types, pairs)
    names = ["neighbor_build", "reverse_index", "desc_fwd", "mlp", "desc_bwd", "force_gather"]
    phase_ms = {k: float(v) for k, v in zip(names, phase)}
    pmode = {nnap.FP32: "fp32", nnap.BF16_MLP: "bf16", nnap.FP32_FAST: "fast"}[prec]
    kernels = kernel_rooflines(flops, phase_ms, pmode, st.natoms, npairs, prec == nnap.FP32)
    dom = max(("desc_fwd", "mlp", "desc_bwd"), key=lambda k: phase_ms[k])
    roof = dict(kernels[dom])
    roof["kernel"] = dom
    roof["traffic"] = committed_traffic(dom) if prec == nnap.FP32 else None
    roof["flop_per_step"] = {k: float(v) for k, v in flops.items()}
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
        "dtype": {"fp32": "fp32 storage; fp64 geometry, descriptor, MLP and sums",
                  "fast": "fp32", "bf16": "fp32 + bf16 MLP"}[pmode],
        "data": "synthetic (seeded RSA liquid, seeded random-init model)",
        "config": {"workload": f"ternary Cu-Zr-Al liquid, {st.natoms} atoms, rho 0.058, rcut 6, nmax 8, lmax 6, "
                               "MLP 64-64 SiLU, " + {"fp32": "plain path (JSE_NNAP_FP32)",
                                                     "fast": "FP32_FAST mode", "bf16": "BF16-MLP mode"}[pmode],
                   "natoms": st.natoms, "pairs": int(stats["pairs"]), "parallelism": "single GPU",
                   "launch_shapes": knobs, "launch_shapes_source": "autotuned" if args.autotune else
                   ("--knobs" if args.knobs else "committed tuned shapes"),
                   "default_shapes_ms_per_step": ms_default},
        "phase_ms": phase_ms,
        "wall_s_timed": wall,
        "roofline": roof,
        "kernel_rooflines": kernels,
    }
    if prec == nnap.FP32 and not args.no_modes:
        line["other_modes_c2"] = other_modes_c2(md, st)
    if not args.no_hooks:
        line["pcie_inclusive"] = pcie_inclusive(pot, st)
        line["gpumd_hook"] = gpumd_hook(pot, st)
    if prec == nnap.FP32 and not args.no_training:
        line["training_derivatives"] = training_derivatives(pot, st)
    if not args.no_cpu_baseline:
        base, err = cpu_baseline(md, st, res)
        line["cpu_baseline"] = base
        line["error_vs_fp64"] = err
    if not args.no_c1:
        line["cu256_c1"] = cu_c1()
    if not args.no_256k:
        line["precision_256k"] = precision_c3()
    if not args.no_batched:
        line["batched_512_replicas"] = batched_c4()
    if not args.no_served:
        sv = served_c5(nrep=args.served_replicas)
        sv["saturation"]["speedup_vs_plain_fp32_c2"] = sv["saturation"]["atom_evals_per_s"] / value
        line["served_2000_replicas"] = sv
        pts = sv["poisson"]
        line["p50_ms"] = pts[-1]["p50_ms"]
        line["p99_ms"] = pts[-1]["p99_ms"]
        line["latency_note"] = (f"served, Poisson arrivals at {pts[-1]['offered_fraction_of_saturation']:.0%} "
                                "of the saturation request rate, submission to results on the host")
    print(json.dumps(line), flush=True)
    pot.close()


if __name__ == "__main__":
    main()
