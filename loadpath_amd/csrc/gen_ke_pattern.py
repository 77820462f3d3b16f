"""Generate ke_pattern.h: the structure the cube's symmetry group forces on a
hex8 element stiffness matrix of an isotropic material.

The 48 symmetries of the unit cube (3! axis permutations x 2^3 reflections)
map corner a to g(a) and displacement component i to pi(i) with a sign
s_i (-1 where the axis is reflected). An element matrix of a material with
that symmetry satisfies

    K[g(a), pi(i); g(b), pi(j)] = s_i s_j K[a, i; b, j],   K = K^T,

so its 576 entries fall into a few orbits: within an orbit every entry is
+-1 times one value, and an orbit that maps an entry onto itself with sign
-1 is identically zero. Both elements of this library (plain trilinear and
"im" with condensed bubbles) have the symmetry for an isotropic material;
the host checks an actual Ke against the pattern before the fine kernel
uses it (an anisotropic Ke falls back to reading all 576 values).

Output: KE_NV (number of orbit values), KE_IDX[576], KE_SGN[576] in
{-1, 0, +1}. Run by build.py; deterministic.
"""
from __future__ import annotations

import itertools
from pathlib import Path


def corner(bits):
    return bits[0] | (bits[1] << 1) | (bits[2] << 2)


def group():
    out = []
    for perm in itertools.permutations(range(3)):
        for refl in itertools.product((0, 1), repeat=3):
            out.append((perm, refl))
    return out


def apply_g(g, a, i):
    perm, refl = g
    bits = [(a >> d) & 1 for d in range(3)]
    nb = [0, 0, 0]
    for d in range(3):
        nb[perm[d]] = bits[d] ^ refl[d]
    sign = -1 if refl[i] else 1
    return corner(nb), perm[i], sign


def orbits():
    # union-find with parity (entry e = K[3a+i][3b+j], value = sign * root)
    n = 576
    parent = list(range(n))
    par = [1] * n          # sign relative to parent

    def find(x):
        if parent[x] == x:
            return x, 1
        r, s = find(parent[x])
        parent[x] = r
        par[x] *= s
        return r, par[x]

    zero = set()

    def union(x, y, s):     # K[y] = s K[x]
        rx, sx = find(x)
        ry, sy = find(y)
        if rx == ry:
            if sx * s != sy:
                zero.add(rx)
            return
        parent[ry] = rx
        par[ry] = s * sx * sy
    for a in range(8):
        for i in range(3):
            for b in range(8):
                for j in range(3):
                    e = (3 * a + i) * 24 + 3 * b + j
                    union(e, (3 * b + j) * 24 + 3 * a + i, 1)
                    for g in group():
                        ga, gi, si = apply_g(g, a, i)
                        gb, gj, sj = apply_g(g, b, j)
                        union(e, (3 * ga + gi) * 24 + 3 * gb + gj, si * sj)
    roots = {}
    idx, sgn = [], []
    for e in range(n):
        r, s = find(e)
        if r in zero:
            idx.append(0)
            sgn.append(0)
            continue
        if r not in roots:
            roots[r] = len(roots)
        idx.append(roots[r])
        sgn.append(s)
    return len(roots), idx, sgn


def write(path: Path):
    nv, idx, sgn = orbits()
    lines = [
        "// Generated by gen_ke_pattern.py: hex8 element stiffness structure under",
        "// the cube's symmetry group (isotropic material). Entry e = 24*row + col",
        "// of Ke equals KE_SGN[e] * value[KE_IDX[e]].",
        "#pragma once",
        "#include <cstdint>",
        "namespace lp {",
        f"constexpr int KE_NV = {nv};",
        "constexpr uint8_t KE_IDX[576] = {" + ", ".join(map(str, idx)) + "};",
        "constexpr int8_t KE_SGN[576] = {" + ", ".join(map(str, sgn)) + "};",
        "}  // namespace lp",
        "",
    ]
    text = "\n".join(lines)
    if not path.exists() or path.read_text() != text:
        path.write_text(text)
    return nv, sum(1 for s in sgn if s == 0)


if __name__ == "__main__":
    nv, nz = write(Path(__file__).resolve().parent / "ke_pattern.h")
    print(f"KE_NV = {nv}, structural zeros = {nz}")
