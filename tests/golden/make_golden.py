"""Writes the committed fixtures of tests/golden (run in the build container):

* structures taken from the reference's own example data files
  (example/data/xyz/CuFCC108.xyz first frame; example/data/lmp/data/data-glass,
  a 4000-atom Cu/Zr glass whose box starts at xlo = -2.65, shifted to the origin);
* seeded random-init models (JSON text in the reference's model format);
* FP64 oracle outputs (per-atom energies, forces, per-atom virials) on the
  Cu FCC 256-atom benchmark structure and on the glass.

Inputs and expected outputs only: no reference source is copied.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
REF = "/root/reference"


def read_xyz_frame0(path):
    with open(path) as f:
        n = int(f.readline())
        hdr = f.readline()
        lat = hdr.split('Lattice="')[1].split('"')[0].split()
        box = np.array([float(lat[0]), float(lat[4]), float(lat[8])])
        xyz = np.array([[float(v) for v in f.readline().split()[1:4]] for _ in range(n)])
    return box, xyz


def read_lmp_atomic(path):
    with open(path) as f:
        lines = f.read().splitlines()
    lo = {}
    for ln in lines:
        for ax in ("x", "y", "z"):
            if ln.strip().endswith(f"{ax}lo {ax}hi"):
                a, b = ln.split()[:2]
                lo[ax] = (float(a), float(b))
    start = [k for k, ln in enumerate(lines) if ln.startswith("Atoms")][0] + 2
    rows = []
    for ln in lines[start:]:
        if not ln.strip():
            break
        p = ln.split()
        rows.append((int(p[0]), int(p[1]), float(p[2]), float(p[3]), float(p[4])))
    rows.sort()
    origin = np.array([lo["x"][0], lo["y"][0], lo["z"][0]])
    box = np.array([lo["x"][1] - lo["x"][0], lo["y"][1] - lo["y"][0], lo["z"][1] - lo["z"][0]])
    xyz = np.array([[r[2], r[3], r[4]] for r in rows]) - origin
    types = np.array([r[1] for r in rows], dtype=np.int32)
    return box, xyz, types


def main():
    import oracle as O
    from jse_amd import model as M
    from jse_amd import workloads as W

    box, xyz = read_xyz_frame0(os.path.join(REF, "example/data/xyz/CuFCC108.xyz"))
    np.savez_compressed(os.path.join(HERE, "cufcc108.npz"), box=box, xyz=xyz)
    gbox, gxyz, gtypes = read_lmp_atomic(os.path.join(REF, "example/data/lmp/data/data-glass"))
    np.savez_compressed(os.path.join(HERE, "data_glass.npz"), box=gbox, xyz=gxyz, types=gtypes)

    # Cu FCC 256 (benchmark structure), 1-species model
    cu = W.cu_model()
    st = W.cu_fcc(seed=1)
    res = O.Oracle(M.load_model(cu)).system_efv(st)
    np.savez_compressed(os.path.join(HERE, "golden_cu256.npz"), model=np.array(json.dumps(cu)), box=st.box,
                        xyz=st.xyz, types=st.types, energies=res["energies"], forces=res["forces"],
                        virials=res["virials"])

    # glass, 2-species model (nmax 5, lmax 6, default wtype -> P = 12)
    m2 = M.random_model(["Cu", "Zr"], nmax=5, lmax=6, rcut=6.0, hidden=(32, 32), seed=21)
    gst = W.Structure(gbox, gxyz, gtypes)
    mdl = M.load_model(m2)
    fps = W.descriptors_np(mdl, gst, np.arange(0, 4000, 10))
    m2 = M.set_normalization(m2, fps)
    res = O.Oracle(M.load_model(m2)).system_efv(gst)
    np.savez_compressed(os.path.join(HERE, "golden_glass.npz"), model=np.array(json.dumps(m2)), box=gbox,
                        xyz=gxyz, types=gtypes, energies=res["energies"], forces=res["forces"],
                        virials=res["virials"])
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
