"""The scene side of the path on the CPU: the OBJ face reader, Patch::divide
with neighbour links, the face-scene builder and the vertex-colour
smoothing restatement, each against hand-computed answers, closed forms
and the restatements in oracle/scene_io.py."""
import numpy as np
import pytest

from oracle import scene_io as S

F32 = np.float32


def _write(tmp_path, text, name="m.obj"):
    p = tmp_path / name
    p.write_bytes(text.encode())
    return str(p)


OBJ_EDGES = (
    "# comment line\n"
    "v 1000 0 0\n"
    "v 0 2000 0\n"
    "v 0 0 3000 1\n"          # w ignored
    "v 1000 2000 3000\n"
    "vn 0 0 1\n"
    "vt 0.5 0.5\n"
    "g group\n"
    "f 1 2 3\n"               # triangle -> A B C C
    "f 1/1/1 2/2/2 4/4/4 3/3/3\n"
    "f 1 2 3 4 1\n"           # five references: skipped
    "f 1 2 9\n"               # undefined vertex: skipped
    "f 0 1 2\n"               # index 0: skipped
    "f -1 2 3\n"              # negative: skipped
    " f 1 2 3\n"              # not at the line start: ignored
    "f 1 2\n"                 # two references: skipped
    "f 4  3 2\r\n"            # empty piece between spaces, CR at the end
    "f 2 3 4"                 # last line without newline
)


def test_obj_known_answer(rad, tmp_path):
    faces, complete = rad.obj_parse(_write(tmp_path, OBJ_EDGES))
    assert complete
    v = np.array([[1, 0, 0], [0, 2, 0], [0, 0, 3], [1, 2, 3]], F32)
    want = np.array([np.concatenate([v[0], v[1], v[2], v[2]]),
                     np.concatenate([v[0], v[1], v[3], v[2]]),
                     np.concatenate([v[3], v[2], v[1], v[1]]),
                     np.concatenate([v[1], v[2], v[3], v[3]])], F32)
    assert np.array_equal(faces, want)
    o, ok = S.obj_parse(OBJ_EDGES)
    assert ok and np.array_equal(o, want)


def test_obj_millimetres(rad, tmp_path):
    # coordinates are divided by 1000 in float
    txt = "v 1234.5 -0.25 7\nv 1 1 1\nv 2 2 3\nf 1 2 3\n"
    faces, _ = rad.obj_parse(_write(tmp_path, txt))
    assert faces[0, 0] == F32(1234.5) / F32(1000)
    assert faces[0, 1] == F32(-0.25) / F32(1000)
    assert np.array_equal(faces, S.obj_parse(txt)[0])


def test_obj_bad_vertex_keeps_earlier_faces(rad, tmp_path):
    txt = "v 0 0 0\nv 1000 0 0\nv 0 1000 0\nf 1 2 3\nv 5 6\nf 1 2 3\n"
    faces, complete = rad.obj_parse(_write(tmp_path, txt))
    assert not complete and faces.shape == (1, 12)
    o, ok = S.obj_parse(txt)
    assert not ok and np.array_equal(o, faces)


def test_obj_missing_file(rad, tmp_path):
    with pytest.raises(rad.RadError):
        rad.obj_parse(str(tmp_path / "nope.obj"))


def test_obj_random_file_matches_restatement(rad, tmp_path):
    rng = np.random.default_rng(5)
    lines = []
    nv = 0
    for _ in range(400):
        r = rng.random()
        if r < 0.45 or nv < 4:
            xyz = rng.uniform(-5000, 5000, 3).round(3)
            lines.append("v " + " ".join(f"{x:g}" for x in xyz))
            nv += 1
        elif r < 0.9:
            k = int(rng.integers(2, 6))
            idx = rng.integers(-1, nv + 3, k)
            style = rng.integers(3)
            toks = [f"{i}" if style == 0 else f"{i}/{i}" if style == 1 else f"{i}//{i}"
                    for i in idx]
            lines.append("f " + " ".join(toks))
        else:
            lines.append(rng.choice(["vn 0 1 0", "# x", "", "usemtl a", "s off"]))
    txt = "\n".join(lines) + "\n"
    faces, ok = rad.obj_parse(_write(tmp_path, txt))
    o, ok2 = S.obj_parse(txt)
    assert ok and ok2 and faces.shape[0] > 20
    assert np.array_equal(faces, o)


# ------------------------------------------------------ subdivision, neighbours

def test_faces_scene_equals_cornell(rad):
    q, c, e, s = rad.cornell_faces()
    assert q.shape == (17, 12)
    for tp in (False, True):
        a = rad.Scene.faces(q, 0.3127, rad.RAD_FACES_TRIANGLES if tp else rad.RAD_FACES_QUADS,
                            c, e, s)
        b = rad.Scene.cornell(0.3127, triangle_patches=tp, physical=False)
        for name in ("tri", "tri_patch", "frame", "refl", "area", "emission", "shot0",
                     "patch_quad", "quad_nbr"):
            x, y = getattr(a, name), getattr(b, name)
            assert x.shape == y.shape and np.array_equal(x.view(np.uint32), y.view(np.uint32)), name


def _random_quads(rng, n):
    out = []
    for _ in range(n):
        o = rng.uniform(-3, 3, 3)
        e1 = rng.normal(size=3) * rng.uniform(0.2, 2.0)
        e2 = np.cross(e1, rng.normal(size=3))
        e2 *= rng.uniform(0.2, 2.0) / np.linalg.norm(e2)
        skew = rng.normal(size=3) * 0.05
        out.append(np.concatenate([o, o + e1, o + e1 + e2 + skew, o + e2]))
    return np.array(out, F32)


def test_divide_children_match_restatement(rad):
    rng = np.random.default_rng(1)
    q = _random_quads(rng, 12)
    area = 0.05
    sc = rad.Scene.faces(q, area, rad.RAD_FACES_QUADS)
    t = 0
    for f in range(len(q)):
        g = S.divide_grid(q[f], area)
        kx, ky = g if g else (1, 1)
        for i in range(kx * ky):
            v = S.divide_child(q[f], kx, ky, i) if g else q[f].reshape(4, 3)
            want = np.concatenate([v[0], v[1], v[2], v[0], v[2], v[3]])
            got = sc.tri[t:t + 2].reshape(-1)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (f, i)
            t += 2
    assert t == sc.ntri


def _geometric_neighbours(kx, ky):
    """8 links by offsets around (col, row); a diagonal that leaves the grid
    falls back along the border it crosses, else to the cell itself."""
    off = [(-1, 1), (0, 1), (1, 1), (1, 0), (1, -1), (0, -1), (-1, -1), (-1, 0)]
    out = np.zeros((kx * ky, 8), np.int64)
    for i in range(kx * ky):
        c, r = i % kx, i // kx
        for s, (dc, dr) in enumerate(off):
            cin, rin = 0 <= c + dc < kx, 0 <= r + dr < ky
            if cin and rin:
                out[i, s] = (r + dr) * kx + c + dc
            elif dc and dr and cin:
                out[i, s] = r * kx + c + dc
            elif dc and dr and rin:
                out[i, s] = (r + dr) * kx + c
            else:
                out[i, s] = i
    return out


@pytest.mark.parametrize("kx,ky", [(1, 1), (1, 4), (3, 1), (2, 2), (3, 2), (5, 4)])
def test_neighbour_rule_pinned(kx, ky):
    assert np.array_equal(S.neighbours(kx, ky), _geometric_neighbours(kx, ky))


def test_scene_neighbours_match_restatement(rad):
    rng = np.random.default_rng(2)
    q = _random_quads(rng, 8)
    sc = rad.Scene.faces(q, 0.07, rad.RAD_FACES_QUADS)
    base = 0
    for f in range(len(q)):
        g = S.divide_grid(q[f], 0.07)
        kx, ky = g if g else (1, 1)
        want = S.neighbours(kx, ky) + base
        assert np.array_equal(sc.quad_nbr[base:base + kx * ky], want), f
        base += kx * ky
    assert base == sc.nquad


def test_native_triangle_grid(rad):
    a, b, c = np.array([0, 0, 0], F32), np.array([2, 0, 0.5], F32), np.array([0.3, 1.5, 0], F32)
    faces = np.array([np.concatenate([a, b, c, c])], F32)
    S0 = 0.5 * np.linalg.norm(np.cross((b - a).astype(float), (c - a).astype(float)))
    for area, n in ((10.0, 1), (S0 / 3.9, 2), (S0 / 50, 8)):
        sc = rad.Scene.faces(faces, area, rad.RAD_FACES_TRIANGLES)
        assert sc.npatch == sc.ntri == sc.nquad == n * n
        assert abs(sc.area.astype(float).sum() / S0 - 1) < 1e-5
        # every child keeps the face's orientation; frame normals follow the
        # reference's cross (a.v_Cross(b) = b x a): (C - A) x (B - A)
        nrm = np.cross(c - a, b - a).astype(float)
        assert np.all(sc.frame[:, 3:6].astype(float) @ nrm > 0)
        assert np.array_equal(sc.quad_nbr, np.repeat(np.arange(n * n), 8).reshape(-1, 8))
        # watertight inside the face: every vertex of the grid is shared
        # bit for bit by all the triangles that meet there
        verts = {tuple(v) for v in sc.tri.reshape(-1, 3).view(np.uint32)}
        assert len(verts) == (n + 1) * (n + 2) // 2
    # in quad mode a triangle is the degenerate quad of the reference
    sc = rad.Scene.faces(faces, 10.0, rad.RAD_FACES_QUADS)
    assert sc.npatch == 1 and sc.ntri == 2


# ------------------------------------------------------------- smoothing

def test_smooth_shade_constant_and_undivided():
    nbr = S.neighbours(4, 3)
    val = np.full((12, 3), F32(0.7), F32)
    assert np.all(S.smooth_shade(val, nbr) == F32(0.7) * F32(4) * F32(0.25))
    one = S.smooth_shade(np.array([[0.3, 1.7, 5.0]], F32), np.zeros((1, 8), np.int64))
    assert np.array_equal(one[0], np.repeat([[0.3, 1.7, 5.0]], 4, 0).astype(F32))


def test_smooth_shade_linear_field_interior():
    # a linear field on the cells: an interior corner takes the mean of its
    # four cells, i.e. the field's value at the corner point
    kx, ky = 5, 4
    nbr = S.neighbours(kx, ky)
    col, row = np.arange(kx * ky) % kx, np.arange(kx * ky) // kx
    val = np.stack([3 + 2 * col + 4 * row, 1 + col, 2 + row], 1).astype(F32)
    out = S.smooth_shade(val, nbr)
    for i in range(kx * ky):
        c, r = col[i], row[i]
        if 0 < c < kx - 1 and 0 < r < ky - 1:
            # corners lb, rb, rt, lt at (c -+ 1/2, r -+ 1/2)
            for k, (dc, dr) in enumerate([(-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)]):
                x, y = c + dc, r + dr
                assert np.allclose(out[i, k], [3 + 2 * x + 4 * y, 1 + x, 2 + y], atol=0), (i, k)


def test_fnv1a64_known_answers():
    """The published FNV-1a 64-bit test vectors pin the checksum and scene
    hash of the checkpoint format."""
    from oracle import scene_io as S
    assert S.fnv1a64(b"") == 0xCBF29CE484222325
    assert S.fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert S.fnv1a64(b"foobar") == 0x85944171F73967E8


def test_checkpoint_writer_reader_round_trip(rad, tmp_path):
    from oracle import scene_io as S
    sc = rad.Scene.cornell(0.5, triangle_patches=True, physical=True, seed=1)
    rng = np.random.default_rng(3)
    u = rng.random((sc.npatch, 3), dtype=np.float32)
    sh = rng.random((sc.npatch, 3), dtype=np.float32)
    path = str(tmp_path / "c.rck")
    S.ckpt_write(path, sc, 1, u, sh, sc.emission)
    raw = open(path, "rb").read()
    assert raw[:8] == b"RADCKPT1" and len(raw) == 40 + 36 * sc.npatch
    # the first unshot plane (red of every patch) follows the 32-byte header
    assert np.array_equal(np.frombuffer(raw[32:32 + 4 * sc.npatch], "<f4"), u[:, 0])
    d = S.ckpt_read(path)
    assert d["checksum_ok"] and d["ntri"] == sc.ntri and d["npatch"] == sc.npatch
    assert np.array_equal(d["unshot"], u) and np.array_equal(d["shot"], sh)
    bad = bytearray(raw)
    bad[40] ^= 1
    open(path, "wb").write(bytes(bad))
    assert not S.ckpt_read(path)["checksum_ok"]
