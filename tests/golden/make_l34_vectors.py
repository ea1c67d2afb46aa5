"""Writes tests/golden/l34_invariants.npz (run in the build container).

The reference holds its L3 / L4 invariants as sparse coefficient tables
(src/nnap/resources/assets/nnap/src/basis_SphericalUtil0.hpp: L3_SIZE,
L3_INDEX, L3_COEFF, L4_SIZE, L4_INDEX, L4_COEFF; evaluated term by term by
calL3SubSub_ / calL4SubSub_, basis_SphericalUtil.hpp:395-467).  This script
reads those tables as text, evaluates every invariant on a few seeded random
A rows (49 values, l <= 6) and stores the A rows and the values:

    A    [NS][49]   inputs
    l3   [NS][23]   the 23 L3 invariants (the reference's column order)
    l4   [NS][9]    the 9 L4 invariants

Inputs and expected outputs only: the tables themselves are not stored.
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/nnap/resources/assets/nnap/src/basis_SphericalUtil0.hpp"
NS = 6


def _block(text, name):
    """Numbers of the initialiser of `name`, grouped by its outermost
    second-level braces (one group per invariant)."""
    i = text.index(name)
    t = text[text.index("=", i) + 1:]
    depth, cur, groups = 0, "", []
    for ch in t:
        if ch == "{":
            depth += 1
            if depth == 2:
                cur = ""
            continue
        if ch == "}":
            depth -= 1
            if depth == 1:
                groups.append(cur)
            if depth == 0:
                break
            continue
        if depth >= 2:
            cur += ch
    return groups


_NUM = r"-?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?"


def _ints(s):
    return [int(v) for v in re.findall(r"-?\d+", s)]


def _floats(s):
    return [float(v) for v in re.findall(_NUM, s)]


def main():
    if not os.path.exists(SRC):
        sys.exit(f"{SRC} not found: the vectors are generated where the reference is present")
    text = open(SRC).read()
    sizes3 = _ints(text[text.index("=", text.index("L3_SIZE")) + 1:text.index("};", text.index("L3_SIZE"))])
    sizes4 = _ints(text[text.index("=", text.index("L4_SIZE")) + 1:text.index("};", text.index("L4_SIZE"))])
    idx3 = [_ints(g) for g in _block(text, "L3_INDEX")]
    co3 = [_floats(g) for g in _block(text, "L3_COEFF")]
    idx4 = [_ints(g) for g in _block(text, "L4_INDEX")]
    co4 = [_floats(g) for g in _block(text, "L4_COEFF")]
    assert len(sizes3) == 23 and len(sizes4) == 9
    assert [len(v) for v in idx3] == [3 * s for s in sizes3] and [len(v) for v in co3] == sizes3
    assert [len(v) for v in idx4] == [4 * s for s in sizes4] and [len(v) for v in co4] == sizes4
    rng = np.random.Generator(np.random.PCG64(20251203))
    A = rng.normal(size=(NS, 49))
    l3 = np.zeros((NS, 23))
    l4 = np.zeros((NS, 9))
    for s in range(NS):
        for k in range(23):
            I = np.array(idx3[k]).reshape(-1, 3)
            l3[s, k] = sum(c * A[s, a] * A[s, b] * A[s, d] for (a, b, d), c in zip(I, co3[k]))
        for k in range(9):
            I = np.array(idx4[k]).reshape(-1, 4)
            l4[s, k] = sum(c * A[s, a] * A[s, b] * A[s, d] * A[s, e] for (a, b, d, e), c in zip(I, co4[k]))
    out = os.path.join(HERE, "l34_invariants.npz")
    np.savez(out, A=A, l3=l3, l4=l4)
    print("wrote", out)


if __name__ == "__main__":
    main()
