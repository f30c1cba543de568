"""Restatements of the reference's scene-side functions around the shooting
path: the OBJ face reader, Patch::divide with its neighbour links, and the
vertex-colour smoothing.  Pure Python / numpy float32, for small inputs.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
never imports it.
"""
import re

import numpy as np

f32 = np.float32


def _floats(s):
    """istringstream >> float, repeated: numbers until the first that fails."""
    out = []
    for tok in s.split():
        try:
            out.append(f32(float(tok)))
        except ValueError:
            break
    return out


_INT = re.compile(r"[+-]?\d+")


def obj_parse(text):
    """WaveFrontModel::parse (WaveFrontModel.cpp:15-136) on the text of a
    file.  Returns (faces n x 12 float32, complete)."""
    verts, faces = [], []
    for line in text.split("\n"):
        if line.startswith("v "):
            p = _floats(line[2:])
            if len(p) < 3:                       # WaveFrontModel.cpp:59-62
                return np.array(faces, f32).reshape(-1, 12), False
            verts.append([p[0] / f32(1000), p[1] / f32(1000), p[2] / f32(1000)])
        elif line.startswith("f "):
            rest, fv = line[2:], []
            while rest and len(fv) <= 4:         # WaveFrontModel.cpp:77-93
                pos = rest.find(" ")
                tok = rest if pos < 0 else rest[:pos]
                m = _INT.match(tok.lstrip(" \t\r\v\f"))
                if m and -2**31 <= int(m.group()) < 2**31:
                    fv.append((int(m.group()) - 1) & 0xFFFFFFFF)
                rest = "" if pos < 0 else rest[pos + 1:]
            if len(fv) == 3:                     # WaveFrontModel.cpp:96-97
                fv.append(fv[-1])
            if len(fv) != 4 or any(i >= len(verts) for i in fv):
                continue
            faces.append([c for i in fv for c in verts[i]])
    return np.array(faces, f32).reshape(-1, 12), True


def _len(a):
    return np.sqrt(f32(a[0] * a[0] + a[1] * a[1]) + f32(a[2] * a[2]), dtype=f32)


def divide_grid(v, area):
    """The split decision of Patch::divide (Patch.cpp:47-83): None when the
    quad is small enough, else (kx, ky)."""
    v = np.asarray(v, f32).reshape(4, 3)
    A, B, C, D = v
    u1, u2 = A - C, B - D
    l1, l2 = _len(u1), _len(u2)
    d = f32(f32(u1[0] * u2[0] + u1[1] * u2[1]) + f32(u1[2] * u2[2]))
    phi = float(np.arccos(f32(d / f32(l1 * l2)), dtype=f32))
    S = float(f32(f32(0.5) * l1) * l2) * np.sin(phi)
    if S <= area * 1.01:
        return None
    a = np.sqrt(area)
    return int(np.ceil(float(_len(B - A)) / a)), int(np.ceil(float(_len(C - B)) / a))


def divide_child(v, kx, ky, i):
    """Corners of child i of the kx x ky grid (Patch.cpp:85-117); division
    by a scalar multiplies by its float reciprocal (Vector.h:448-452)."""
    v = np.asarray(v, f32).reshape(4, 3)
    A, B, C, D = v
    rx, ry = f32(1) / f32(kx), f32(1) / f32(ky)
    pCD, pAB = (C - D) * rx, (B - A) * rx
    col, row = i % kx, i // kx
    bCD, bAB = pCD * f32(col) + D, pAB * f32(col) + A
    bCD1, bAB1 = pCD * f32(col + 1) + D, pAB * f32(col + 1) + A
    return np.array([bAB + ((bCD - bAB) * ry) * f32(row),
                     bAB1 + ((bCD1 - bAB1) * ry) * f32(row),
                     bAB1 + ((bCD1 - bAB1) * ry) * f32(row + 1),
                     bAB + ((bCD - bAB) * ry) * f32(row + 1)], f32)


def neighbours(kx, ky):
    """Patch.cpp:127-222: the 8 links of every child of a kx x ky grid, as
    child indices, slot by slot as the reference writes them."""
    out = np.zeros((kx * ky, 8), np.int64)
    for i in range(kx * ky):
        col, row = i % kx, i // kx
        at = lambda r, c: r * kx + c  # noqa: E731
        n = out[i]
        if row + 1 >= ky:
            n[0] = at(row, col - 1) if col > 0 else i
        elif col == 0:
            n[0] = at(row + 1, col) if row + 1 < ky else i
        else:
            n[0] = at(row + 1, col - 1)
        n[1] = i if row + 1 >= ky else at(row + 1, col)
        if row + 1 >= ky:
            n[2] = at(row, col + 1) if col + 1 < kx else i
        elif col + 1 >= kx:
            n[2] = at(row + 1, col) if row + 1 < ky else i
        else:
            n[2] = at(row + 1, col + 1)
        n[3] = i if col + 1 >= kx else at(row, col + 1)
        if row == 0:
            n[4] = at(row, col + 1) if col + 1 < kx else i
        elif col + 1 >= kx:
            n[4] = at(row - 1, col) if row > 0 else i
        else:
            n[4] = at(row - 1, col + 1)
        n[5] = i if row == 0 else at(row - 1, col)
        if row == 0:
            n[6] = at(row, col - 1) if col > 0 else i
        elif col == 0:
            n[6] = at(row - 1, col) if row > 0 else i
        else:
            n[6] = at(row - 1, col - 1)
        n[7] = i if col == 0 else at(row, col - 1)
    return out


def quad_values(patch_quad, nquad, refl, area, unshot, shot):
    """colour * (illumination + radiosity) per quad (Colors.cpp:198); a quad
    of several patches takes their area-weighted mean, summed in patch
    order."""
    refl, area = np.asarray(refl, f32), np.asarray(area, f32)
    unshot, shot = np.asarray(unshot, f32), np.asarray(shot, f32)
    v = refl * (shot + unshot)
    out = np.zeros((nquad, 3), f32)
    members = [[] for _ in range(nquad)]
    for p, q in enumerate(patch_quad):
        members[q].append(p)
    for q, ps in enumerate(members):
        if len(ps) == 1:
            out[q] = v[ps[0]]
        else:
            num, den = np.zeros(3, f32), f32(0)
            for p in ps:
                num = num + area[p] * v[p]
                den = f32(den + area[p])
            out[q] = num / den
    return out


def smooth_shade(val, nbr):
    """Colors::smoothShadePatch (Colors.cpp:195-262) for every quad:
    corners lb, rb, rt, lt."""
    val = np.asarray(val, f32)
    nbr = np.asarray(nbr)
    out = np.zeros((len(val), 4, 3), f32)
    quarter = f32(1) / f32(4)
    for q in range(len(val)):
        n = nbr[q]
        lt = val[q] + val[n[7]] + val[n[0]] + val[n[1]]
        rt = val[q] + val[n[1]] + val[n[2]] + val[n[3]]
        rb = val[q] + val[n[3]] + val[n[4]] + val[n[5]]
        lb = val[q] + val[n[5]] + val[n[6]] + val[n[7]]
        out[q] = [lb * quarter, rb * quarter, rt * quarter, lt * quarter]
    return out
