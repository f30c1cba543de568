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


# ---------------------------------------------------------------- checkpoints
# The portable solver-state file that succeeds the reference's .rr dump of
# raw Patch structs (Main.cpp:1460-1647, LoadingModel.cpp:4-26), written here
# independently of the library from its documented layout:
#   "RADCKPT1" | u32 version (1) | u32 npatch | u32 ntri | u32 mode |
#   u64 scene hash | f32 unshot[3][npatch] | f32 shot[3][npatch] |
#   f32 emission[3][npatch] | u64 FNV-1a of every byte before it
# little-endian, one plane per channel.  The scene hash is FNV-1a over ntri
# and npatch (u32 each), the triangles (ntri x 9 f32, triangle after
# triangle), the triangle patch ids (u32), the patch frames (npatch x 9 f32),
# the reflectances (npatch x 3 f32) and the areas (f32).

FNV_OFFSET = 0xCBF29CE484222325
FNV_PRIME = 0x100000001B3
_MASK64 = (1 << 64) - 1


def fnv1a64(data, h=FNV_OFFSET):
    """64-bit FNV-1a (Fowler/Noll/Vo) of a bytes-like object."""
    for b in bytes(data):
        h = ((h ^ b) * FNV_PRIME) & _MASK64
    return h


def _le(a, dtype):
    return np.ascontiguousarray(a, dtype=np.dtype(dtype).newbyteorder("<")).tobytes()


def scene_hash(tri, tri_patch, frame, refl, area):
    tri = np.asarray(tri, f32).reshape(-1, 9)
    ntri, npatch = tri.shape[0], np.asarray(area).shape[0]
    h = fnv1a64(_le([ntri, npatch], np.uint32))
    for part in (_le(tri, f32), _le(tri_patch, np.uint32), _le(frame, f32), _le(refl, f32),
                 _le(area, f32)):
        h = fnv1a64(part, h)
    return h


def ckpt_write(path, sc, mode, unshot, shot, emission, corrupt=None):
    """Writes a checkpoint of state (unshot, shot, emission: npatch x 3) for
    scene sc (tri, tri_patch, frame, refl, area)."""
    n = sc.area.shape[0]
    head = (b"RADCKPT1" + _le([1, n, sc.tri.shape[0], mode], np.uint32)
            + _le([scene_hash(sc.tri, sc.tri_patch, sc.frame, sc.refl, sc.area)], np.uint64))
    body = b"".join(_le(np.asarray(a, f32).reshape(n, 3).T, f32)
                    for a in (unshot, shot, emission))
    blob = head + body
    blob += _le([fnv1a64(blob)], np.uint64)
    with open(path, "wb") as f:
        f.write(blob)


def ckpt_read(path):
    """Parses a checkpoint; returns a dict of its fields and whether its
    checksum holds."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"RADCKPT1", "magic"
    ver, n, ntri, mode = np.frombuffer(raw[8:24], "<u4")
    (h,) = np.frombuffer(raw[24:32], "<u8")
    body = np.frombuffer(raw[32:32 + 36 * int(n)], "<f4").reshape(9, int(n))
    (cs,) = np.frombuffer(raw[32 + 36 * int(n):], "<u8")
    return {"version": int(ver), "npatch": int(n), "ntri": int(ntri), "mode": int(mode),
            "scene_hash": int(h), "unshot": body[0:3].T.copy(), "shot": body[3:6].T.copy(),
            "emission": body[6:9].T.copy(), "size": len(raw),
            "checksum_ok": int(cs) == fnv1a64(raw[:32 + 36 * int(n)])}
