"""Pins the CPU oracle (oracle/oracle.c) against known answers, closed forms,
invariants and brute force on tiny inputs.  CPU only."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

F32 = np.float32


# ----------------------------------------------------------------- ΔFF table

def test_dff_sum_known_answer():
    # FormFactors.cpp:106: "must be 1, I get 1.00435" - printed by cout with 6
    # significant digits; the float sum in the reference's loop order at W=256.
    s = O.dff_sum_test(256)
    assert f"{s:.6g}" == "1.00435"
    assert abs(O.dff_sum_test(16) - 1.07146) < 1e-5


def test_dff_side_quirk_is_first_order():
    # The extra half pixel in the side numerator (dy + h) makes the excess
    # over 1 shrink like 1/W: halving the pixel halves the excess.
    e128 = O.dff_sum_test(128) - 1.0
    e256 = O.dff_sum_test(256) - 1.0
    assert 1.9 < e128 / e256 < 2.1


@pytest.mark.parametrize("W", [16, 32])
def test_dff_closed_form(W):
    top, side = O.dff_top_side(W)
    h = 1.0 / W
    area = (2.0 / W) ** 2
    for x in range(W):
        for y in range(W):
            dx = (x - W // 2) / (W / 2.0) + h
            dy = (y - W // 2) / (W / 2.0) + h
            ref = area / (math.pi * (dx * dx + dy * dy + 1) ** 2)
            assert top[x + y * W] == pytest.approx(ref, rel=2e-6)
    for x in range(W):
        for y in range(W // 2):
            dx = (x - W // 2) / (W / 2.0) + h
            dy = (W // 2 - 1 - y) / (W / 2.0) + h
            ref = area * (dy + h) / (math.pi * (dx * dx + dy * dy + 1) ** 2)
            assert side[x + y * W] == pytest.approx(ref, rel=2e-6)


@pytest.mark.parametrize("W", [16, 64])
def test_dff_cross_layout(W):
    TW, TH, RES = O.geometry(W)
    assert (TW, TH, RES) == (2 * W, W * 3 // 2, 3 * W * W)
    c = O.dff_cross(W).reshape(TH, TW)
    top, side = O.dff_top_side(W)
    top = top.reshape(W, W)  # [y][x]
    assert np.all(c > 0)  # every texel belongs to exactly one face region
    # front = top table
    assert np.array_equal(c[:W, W // 2:W + W // 2], top)
    # total equals the __SUM_TEST total up to float summation order
    assert c.sum(dtype=np.float64) == pytest.approx(float(O.dff_sum_test(W)), rel=1e-5)
    # mirror symmetry of the left/right halves and of the front
    assert np.array_equal(c[:W, :W // 2], c[:W, ::-1][:, :W // 2])
    assert np.array_equal(top, top[:, ::-1]) and np.array_equal(top, top[::-1, :])
    # geometric closed form per texel: every texel is a direction on the
    # unit hemicube; a side-face texel at height z above the patch plane and
    # lateral offset l has ΔFF = ΔA (z + 1/W) / (π (1 + z^2 + l^2)^2) (the
    # reference's numerator quirk), heights taken at pixel centres from the
    # viewport placement of each face (Main.cpp:316-389)
    hw = W / 2.0
    dA = (2.0 / W) ** 2
    exp = np.zeros((TH, TW))
    for y in range(TH):
        for x in range(TW):
            if y < W and W // 2 <= x < W + W // 2:        # front
                a, b = (x - W // 2 + 0.5 - hw) / hw, (y + 0.5 - hw) / hw
                exp[y, x] = dA / (math.pi * (1 + a * a + b * b) ** 2)
                continue
            if y < W and x < W // 2:                      # left: height grows with x
                z, lat = (x + 0.5) / hw, (y + 0.5 - hw) / hw
            elif y < W:                                   # right: height falls with x
                z, lat = (2 * W - 0.5 - x) / hw, (y + 0.5 - hw) / hw
            elif x < W:                                   # up: height falls with y
                z, lat = (1.5 * W - 0.5 - y) / hw, (x + 0.5 - hw) / hw
            else:                                         # down: height grows with y
                z, lat = (y - W + 0.5) / hw, (x - W + 0.5 - hw) / hw
            exp[y, x] = dA * (z + 1.0 / W) / (math.pi * (1 + z * z + lat * lat) ** 2)
    assert np.allclose(c, exp, rtol=3e-6, atol=0)
    # a transposed side index would break the closed form
    left = c[:W, :W // 2]
    assert not np.allclose(left, left[::-1, ::-1].T.reshape(left.shape) if W == 2 else left[:, ::-1])


# ------------------------------------------------------------------- camera

def _frame(eye, n, u):
    return np.array(list(eye) + list(n) + list(u), np.float32)


def _project(m, p):
    m = m.reshape(4, 4).astype(np.float64)  # m[col][row]
    v = np.array([p[0], p[1], p[2], 1.0])
    c = m.T @ v
    return c[:3] / c[3], c


def test_quad_frame_closed_form():
    v = np.array([[0, 0, 1], [2, 0, 1], [2, 3, 1], [0, 3, 1]], np.float32)
    fr = O.quad_frame(v)
    assert np.allclose(fr[:3], [1, 1.5, 1])
    # (v4-v1) x (v2-v1) = (0,3,0) x (2,0,0) = (0,0,-6)
    assert np.allclose(fr[3:6], [0, 0, -6])
    assert np.allclose(fr[6:9], [0, 3, 0])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_face_mvp_view_is_a_mirror(seed):
    rng = np.random.default_rng(seed)
    n = rng.normal(size=3)
    u = np.cross(n, rng.normal(size=3))
    fr = _frame(rng.normal(size=3), n, u)
    for face in range(5):
        m = O.face_mvp(fr, face).reshape(4, 4).astype(np.float64)
        # P only scales rows 0, 1 by 1 and builds rows 2, 3 from view row 2;
        # rows 0, 1 and -row 3 of P*V are the view rotation rows r, u', -d
        R = np.stack([m[:3, 0], m[:3, 1], -m[:3, 3]])
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
        assert np.linalg.det(R) == pytest.approx(-1.0, abs=1e-5)


def test_face_mvp_axes():
    eye = np.array([1.0, 2.0, 3.0])
    n = np.array([0.0, 0.0, 2.0])     # unnormalised, as getNormal
    u = np.array([0.0, 3.0, 0.0])     # getUp = v4 - v1
    fr = _frame(eye, n, u)
    nh, uh = n / 2, u / 3
    r_front = np.cross(uh, nh)        # right = up x dir (mirrored LookAt)
    # targets per face (Camera.cpp:26-51) and the camera up of each
    faces = {
        0: (uh, -nh),                 # UP
        1: (-uh, nh),                 # DOWN
        2: (-np.cross(uh, nh), uh),   # LEFT:  -(n.v_Cross(u)) = n x u ... = -(u x n)
        3: (np.cross(uh, nh), uh),    # RIGHT: n.v_Cross(u) = u x n
        4: (nh, uh),                  # FRONT
    }
    for face, (d, up) in faces.items():
        m = O.face_mvp(fr, face)
        ndc, clip = _project(m, eye + 2.0 * d)
        assert abs(ndc[0]) < 1e-6 and abs(ndc[1]) < 1e-6, face
        assert clip[3] == pytest.approx(2.0, rel=1e-6)  # w = distance along the view
        right = np.cross(up, d)
        ndc_up, _ = _project(m, eye + 2.0 * d + 1.0 * up)
        assert ndc_up[1] == pytest.approx(0.5, abs=1e-5) and abs(ndc_up[0]) < 1e-5
        ndc_r, _ = _project(m, eye + 2.0 * d + 1.0 * right)
        assert ndc_r[0] == pytest.approx(0.5, abs=1e-5) and abs(ndc_r[1]) < 1e-5
    assert np.allclose(r_front, np.cross(uh, nh))
    # depth: GL window z = (z_ndc + 1)/2 with near 0.01, far 1000
    m = O.face_mvp(fr, 4)
    ndc, _ = _project(m, eye + 1.0 * nh)
    f, nr = 1000.0, 0.01
    z_ndc = (-(f + nr) / (f - nr) * -1.0 - 2 * f * nr / (f - nr)) / 1.0
    assert ndc[2] == pytest.approx(z_ndc, abs=1e-6)


def test_viewports_tile_the_cross():
    W = 16
    TW, TH, _ = O.geometry(W)
    cover = np.zeros((TH, TW), int)
    for face in range(5):
        vp, sc = O.face_viewport(W, face)
        assert vp[2] == W and vp[3] == W
        x0, y0, w, h = sc
        # the scissor lies inside the viewport
        assert vp[0] <= x0 and x0 + w <= vp[0] + W and vp[1] <= y0 and y0 + h <= vp[1] + W
        cover[y0:y0 + h, x0:x0 + w] += 1
    assert np.all(cover == 1)


# --------------------------------------------------------------- rasterizer

def _snap(m, tri, W):
    """Spec of the vertex pipeline in float32 (no clipping needed)."""
    mm = m.reshape(4, 4)
    out = []
    halfW = F32(0.5) * F32(W)
    for k in range(3):
        x, y, z = (F32(t) for t in tri[3 * k:3 * k + 3])
        c = [((mm[0, r] * x + mm[1, r] * y) + mm[2, r] * z) + mm[3, r] for r in range(4)]
        iw = F32(1.0) / c[3]
        xl = (c[0] * iw + F32(1.0)) * halfW
        yl = (c[1] * iw + F32(1.0)) * halfW
        out.append((int(np.rint(xl * F32(256))), int(np.rint(yl * F32(256)))))
    return out


def _brute_cover(v, sc_local):
    """Pixel centres inside a snapped CCW triangle, exact integer arithmetic
    with a top-left tie rule; returns a set of (ix, iy)."""
    (x0, y0), (x1, y1), (x2, y2) = v
    area = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0)
    if area <= 0:
        return set()
    pts = set()
    edges = [(v[0], v[1]), (v[1], v[2]), (v[2], v[0])]
    for iy in range(sc_local[1], sc_local[3]):
        for ix in range(sc_local[0], sc_local[2]):
            px, py = ix * 256 + 128, iy * 256 + 128
            ok = True
            for (ax, ay), (bx, by) in edges:
                e = (bx - ax) * (py - ay) - (by - ay) * (px - ax)
                dx, dy = bx - ax, by - ay
                topleft = dy < 0 or (dy == 0 and dx < 0)
                if e < 0 or (e == 0 and not topleft):
                    ok = False
                    break
            if ok:
                pts.add((ix, iy))
    return pts


def _local_scissor(W, face):
    vp, sc = O.face_viewport(W, face)
    return vp, (sc[0] - vp[0], sc[1] - vp[1], sc[0] + sc[2] - vp[0], sc[1] + sc[3] - vp[1])


@pytest.mark.parametrize("seed", range(6))
def test_raster_matches_bruteforce(seed):
    W = 16
    rng = np.random.default_rng(seed)
    fr = _frame([0, 0, 0], [0, 0, 1], [0, 1, 0])
    # two triangles of a quad in front of the FRONT face (z in 0.5..2)
    q = np.array([[-0.6, -0.5, 1.0], [0.7, -0.4, 1.1], [0.5, 0.6, 0.9], [-0.4, 0.5, 1.0]])
    q = q + rng.normal(scale=0.1, size=q.shape)
    q[:, 2] = np.abs(q[:, 2]) + 0.5
    tri = np.array([np.r_[q[0], q[1], q[2]], np.r_[q[0], q[2], q[3]]], np.float32)
    # the face maps patch normal +z to the view direction; orient the quad
    # so it faces the camera (CCW on screen after the mirrored view)
    m = O.face_mvp(fr, 4)
    for t in range(2):
        v = _snap(m, tri[t], W)
        area = (v[1][0] - v[0][0]) * (v[2][1] - v[0][1]) - (v[2][0] - v[0][0]) * (v[1][1] - v[0][1])
        if area < 0:
            tri[t] = np.r_[tri[t][0:3], tri[t][6:9], tri[t][3:6]]
    vp, scl = _local_scissor(W, 4)
    TW, TH, RES = O.geometry(W)
    sets = []
    for t in range(2):
        item = O.render_hemicube(W, tri[t:t + 1], np.array([0], np.uint32), fr)
        got = {(i % TW - vp[0], i // TW - vp[1]) for i in np.nonzero(item)[0]}
        want = _brute_cover(_snap(m, tri[t], W), scl)
        assert got == want
        sets.append(got)
    # shared edge: no pixel centre drawn twice (watertight split of the quad)
    assert not (sets[0] & sets[1])


def _box_room(lo=0.0, hi=2.0):
    """Closed cube room, one quad per wall, inward normals, as triangles."""
    from itertools import product  # noqa: F401
    L, H = lo, hi
    quads = []

    def quad(o, e1, e2, n):
        o, e1, e2, n = (np.array(a, float) for a in (o, e1, e2, n))
        if np.dot(np.cross(e2, e1), n) < 0:
            e1, e2 = e2, e1
        return [o, o + e1, o + e1 + e2, o + e2]
    s = H - L
    quads.append(quad([L, L, L], [s, 0, 0], [0, 0, s], [0, 1, 0]))
    quads.append(quad([L, H, L], [s, 0, 0], [0, 0, s], [0, -1, 0]))
    quads.append(quad([L, L, L], [s, 0, 0], [0, s, 0], [0, 0, 1]))
    quads.append(quad([L, L, H], [s, 0, 0], [0, s, 0], [0, 0, -1]))
    quads.append(quad([L, L, L], [0, 0, s], [0, s, 0], [1, 0, 0]))
    quads.append(quad([H, L, L], [0, 0, s], [0, s, 0], [-1, 0, 0]))
    tri, tp = [], []
    for i, q in enumerate(quads):
        tri.append(np.r_[q[0], q[1], q[2]])
        tri.append(np.r_[q[0], q[2], q[3]])
        tp += [i, i]
    return np.array(tri, np.float32), np.array(tp, np.uint32)


@pytest.mark.parametrize("seed", range(4))
def test_closed_room_has_no_background(seed):
    # From any point inside a closed room every pixel of the hemicube shows a
    # wall, so the form-factor row sums to the ΔFF table sum.
    W = 32
    tri, tp = _box_room()
    rng = np.random.default_rng(seed)
    n = rng.normal(size=3)
    u = np.cross(n, rng.normal(size=3))
    fr = _frame(rng.uniform(0.4, 1.6, size=3), n, u)
    item = O.render_hemicube(W, tri, tp, fr)
    assert np.all(item > 0)
    dff = O.dff_cross(W)
    hits, ff = O.histogram(W, 1, item, dff, 6)
    assert ff.sum() == pytest.approx(dff.sum(dtype=np.float64), rel=1e-12)


def test_parallel_plates_form_factor():
    # Two coaxial unit squares at distance 1: F = 0.19982 (textbook value,
    # Howell catalogue C-11).  The hemicube estimate converges to it up to the
    # side-face numerator bias of the reference (a few 1e-3 at W=256).
    W = 256
    a = 0.5
    tri = np.array([[-a, -a, 1, a, -a, 1, a, a, 1], [-a, -a, 1, a, a, 1, -a, a, 1]], np.float32)
    # only the front face sees the plate; make it face the camera
    fr = _frame([0, 0, 0], [0, 0, 1], [0, 1, 0])
    m = O.face_mvp(fr, 4)
    for t in range(2):
        v = _snap(m, tri[t], W)
        area = (v[1][0] - v[0][0]) * (v[2][1] - v[0][1]) - (v[2][0] - v[0][0]) * (v[1][1] - v[0][1])
        if area < 0:
            tri[t] = np.r_[tri[t][0:3], tri[t][6:9], tri[t][3:6]]
    # average the point form factor over the emitting square (5x5 points)
    dff = O.dff_cross(W)
    est = []
    for px in np.linspace(-0.4, 0.4, 5):
        for py in np.linspace(-0.4, 0.4, 5):
            fr = _frame([px, py, 0], [0, 0, 1], [0, 1, 0])
            item = O.render_hemicube(W, tri, np.array([0, 0], np.uint32), fr)
            est.append(dff[item > 0].sum(dtype=np.float64))
    X = 1.0
    x2 = math.sqrt(1 + X * X)
    f_exact = (2 / (math.pi * X * X)) * (math.log(math.sqrt((1 + X * X) ** 2 / (1 + 2 * X * X)))
                                         + X * x2 * math.atan(X / x2) - X * math.atan(X)
                                         + X * x2 * math.atan(X / x2) - X * math.atan(X))
    assert f_exact == pytest.approx(0.19982, abs=1e-4)
    assert np.mean(est) == pytest.approx(f_exact, rel=0.03)


# ----------------------------------------------------------- ProcessHemicube

def test_run_form_equals_histogram(rad):
    sc = rad.Scene.cornell(0.3127, triangle_patches=False, physical=False)
    W = 32
    K = 3
    TW, TH, RES = O.geometry(W)
    dff = O.dff_cross(W)
    shooters = [int(np.argmax(sc.emission.sum(1))), 10, 500]
    item = np.concatenate([O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[s]) for s in shooters])
    ff_runs, emitted, unknown = O.process_runs(W, K, item, dff, sc.npatch)
    hits, ff64 = O.histogram(W, K, item, dff, sc.npatch)
    assert unknown == 0
    assert np.allclose(ff_runs, ff64, rtol=1e-5, atol=1e-9)
    assert hits.sum() == np.count_nonzero(item)
    # triplets = runs of equal non-zero ids inside each of the 4 spans of a row
    rows = item.reshape(K * TH, TW)
    span = TW // 4
    runs = 0
    for r in rows:
        for g in range(4):
            seg = r[g * span:(g + 1) * span]
            starts = np.r_[True, seg[1:] != seg[:-1]]
            runs += int(np.count_nonzero(starts & (seg > 0)))
    assert emitted == runs


def test_unknown_ids_are_counted():
    W = 16
    TW, TH, RES = O.geometry(W)
    item = np.zeros(RES, np.uint32)
    item[5] = 7      # patch 6
    item[20] = 99    # beyond npatch
    dff = O.dff_cross(W)
    ff, emitted, unknown = O.process_runs(W, 1, item, dff, 10)
    assert unknown == 1 and emitted == 2
    assert ff[0, 6] == dff[5]


# ---------------------------------------------------------------- selection

def _select_list(unshot, count):
    """Literal restatement of ModelContainer.cpp:263-298 with Python lists."""
    e = [float(np.float32(u[0]) * u[0] + np.float32(u[1]) * u[1] + np.float32(u[2]) * u[2])
         for u in unshot.astype(np.float32)]
    tops = []
    for pi in range(len(e)):
        if not tops or (e[pi] > 0 and e[tops[-1]] <= e[pi]):
            tops.append(pi)
            tops.sort(key=lambda i: e[i])     # stable, ascending
            tops.reverse()
            del tops[count:]
    return tops


@pytest.mark.parametrize("seed", range(8))
def test_select_reference_matches_list_rule(seed):
    rng = np.random.default_rng(seed)
    n = 60
    vals = rng.integers(0, 4, size=(n, 3)).astype(np.float32)  # many ties and zeros
    for count in (1, 3, 7, 64):
        ids, got_n = O.select_reference(vals, count)
        want = _select_list(vals, count)
        assert got_n == len(want)
        assert list(ids[:got_n]) == want and np.all(ids[got_n:] == 0)


def test_select_reference_quirks():
    z = np.zeros((5, 3), np.float32)
    ids, n = O.select_reference(z, 3)
    assert n == 1 and ids[0] == 0           # patch 0 is always taken first
    u = np.array([[np.sqrt(5), 0, 0], [1, 0, 0], [3, 0, 0]], np.float32)
    ids, n = O.select_reference(u, 3)
    # patch 1 (key 1) is below the list's last key (5) when it arrives and is
    # never added although the list has room: fewer than count are returned
    assert n == 2 and list(ids) == [2, 0, 0]
    u = np.array([[1, 0, 0], [3, 0, 0], [2, 0, 0], [3, 0, 0]], np.float32)
    ids, n = O.select_reference(u, 3)
    # equal keys: the later index ranks first
    assert list(ids[:n]) == [3, 1, 2]


def test_select_argmax_power():
    u = np.array([[1, 1, 1], [2, 0, 1], [0, 0, 0], [1, 2, 0]], np.float32)
    a = np.array([1, 1, 5, 1], np.float32)
    assert O.select_argmax_power(u, a) == 0   # ties -> lowest index


# ----------------------------------------------------------------- shooting

def test_shoot_reference_two_patches():
    # Main.cpp:1274: B_j += dB_i * F_ij * 0.3 * colour_i
    refl = np.array([[1.0, 0.5, 0.0], [0.2, 0.2, 0.2]], np.float32)
    area = np.ones(2, np.float32)
    u = np.array([[10, 20, 30], [0, 0, 0]], np.float64)
    s = np.zeros((2, 3))
    ff = np.array([[0.0, 0.25]])
    O.shoot(0, [0], ff, refl, area, u, s)
    c = float(np.float32(0.3))
    assert np.allclose(u[1], [10 * 0.25 * c * 1.0, 20 * 0.25 * c * 0.5, 0.0])
    assert np.allclose(u[0], 0) and np.allclose(s[0], [10, 20, 30])


@pytest.mark.parametrize("seed", range(3))
def test_shoot_physical_conserves_energy(seed):
    rng = np.random.default_rng(seed)
    n, k = 40, 5
    refl = rng.uniform(0.1, 0.8, size=(n, 3)).astype(np.float32)
    area = rng.uniform(0.5, 2.0, size=n).astype(np.float32)
    u = rng.uniform(0, 1, size=(n, 3))
    s = np.zeros((n, 3))
    shooters = rng.choice(n, size=k, replace=False).astype(np.uint32)
    ff = rng.uniform(0, 1, size=(k, n)) / n
    ff[np.arange(k), shooters] = 0
    snap = u[shooters].copy()
    u0 = u.copy()
    O.shoot(1, shooters, ff, refl, area, u, s)
    # received power: sum_j A_j dB_j = sum_i A_i dB_i sum_j rho_j F_ij
    recv = (u - u0) * area[:, None]
    recv[shooters] += snap * area[shooters, None]
    want = np.zeros(3)
    for si, i in enumerate(shooters):
        want += area[i] * snap[si] * (refl * ff[si][:, None]).sum(0)
    assert np.allclose(recv.sum(0), want, rtol=1e-12)
    assert np.allclose(s[shooters], snap)


def test_residual():
    u = np.array([[1, 1, 1], [0, 0, 0]], np.float64)
    e = np.array([[2, 2, 2], [1, 1, 1]], np.float32)
    a = np.array([1.0, 2.0], np.float32)
    assert O.residual(u, e, a) == pytest.approx(3.0 / (6.0 + 6.0))


# ------------------------------------------------------------------ gather

def _random_system(seed, n, R):
    rng = np.random.default_rng(seed)
    M = rng.random((n, n)).astype(np.float32)
    np.fill_diagonal(M, 0.0)
    M /= (M.sum(1, keepdims=True) * 1.02).astype(np.float32)
    refl = rng.uniform(0.1, 0.8, (n, 3)).astype(np.float32)
    area = rng.uniform(0.5, 2.0, n).astype(np.float32)
    E = rng.random((R, n)) * (rng.random((R, n)) < 0.2)
    return M, refl, area, E


def test_gather_jacobi_equals_direct_solve(orc):
    """The FP64 Jacobi fixed point equals LAPACK's solve of
    (I - diag(rho_c) M) B = E per column (channel c = r % 3)."""
    M, refl, area, E = _random_system(5, 60, 5)
    B, it = orc.gather_jacobi(M, refl, area, E, tol=1e-14)
    assert 5 < it < 2000
    for r in range(E.shape[0]):
        A = np.eye(60) - refl[:, r % 3, None].astype(np.float64) * M.astype(np.float64)
        assert np.allclose(B[r], np.linalg.solve(A, E[r]), rtol=1e-11, atol=1e-13)


def test_gather_jacobi_residual_measure(orc):
    """tol is the area-weighted residual sum_A|B_k+1 - B_k| / sum_A|E|: a
    loose tolerance stops early, and the stopped iterate's true residual is
    within the bound the contraction gives."""
    M, refl, area, E = _random_system(6, 40, 3)
    B1, it1 = orc.gather_jacobi(M, refl, area, E, tol=1e-3)
    B2, it2 = orc.gather_jacobi(M, refl, area, E, tol=1e-14)
    assert it1 < it2
    den = float((area[None, :] * np.abs(E)).sum())
    err = float((area[None, :] * np.abs(B1 - B2)).sum()) / den
    # ||B* - B_k|| <= q/(1-q) ||B_k - B_k-1||, q = max rho * max row sum < 0.8
    assert err <= 1e-3 * 0.8 / 0.2 * 1.5


# ------------------------------------------ item buffer vs a GL driver's arithmetic

def _on_boundary(a, b, W):
    """Every texel where a and b differ has, in a 3 x 3 neighbourhood of one
    of the two buffers, both of its ids: it lies on the edge between them."""
    TW, TH, _ = O.geometry(W)
    A, B = a.reshape(TH, TW), b.reshape(TH, TW)
    for y, x in zip(*np.nonzero(A != B)):
        ys, xs = slice(max(0, y - 1), y + 2), slice(max(0, x - 1), x + 2)
        ids = {int(A[y, x]), int(B[y, x])}
        if not (ids <= set(A[ys, xs].ravel()) or ids <= set(B[ys, xs].ravel())):
            return False
    return True


@pytest.mark.parametrize("flags", [O.VAR_DEPTH24, O.VAR_FMA, O.VAR_DIVIDE,
                                   O.VAR_DEPTH24 | O.VAR_FMA | O.VAR_DIVIDE])
def test_item_buffer_sensitivity_to_driver_arithmetic(rad, flags):
    """The GL item buffer of the reference is not reproducible here, so bound
    how much it could differ: render with the arithmetic a driver may use
    instead (24-bit depth as the reference's GL_DEPTH_COMPONENT24, FMA
    transforms, a divide by w) and count the changed texels.  They stay
    below 0.01 % and lie on triangle edges only."""
    sc = rad.Scene.cornell(0.3127, triangle_patches=False, physical=False)
    rng = np.random.default_rng(1)
    diff = total = 0
    for p in rng.choice(sc.npatch, 40, replace=False):
        a = O.render_hemicube(128, sc.tri, sc.tri_patch, sc.frame[p])
        b = O.render_hemicube_variant(128, sc.tri, sc.tri_patch, sc.frame[p], flags)
        diff += int((a != b).sum())
        total += a.size
        assert _on_boundary(a, b, 128)
    print(f"flags {flags}: {diff} of {total} texels differ ({100 * diff / total:.4f} %)")
    assert diff / total < 1e-4
    if flags == O.VAR_DEPTH24:
        # float depth in [0.5, 1) is coarser than 24-bit fixed point, so the
        # rounding keeps every order and every tie
        assert diff == 0


# --------------------------------------------- multi-core CPU baseline leg

@pytest.mark.parametrize("threads", [2, 5, 8])
def test_mt_render_equals_single_thread(rad, threads):
    """The OpenMP render (triangle slices merged in slice order with the
    strict depth test) equals the single-thread GL_LESS render bit for bit,
    and so does a plain pass built on it."""
    sc = rad.Scene.boxes(nboxes=30, room=(8.0, 3.0, 8.0), target_tris=6000, nlights=4, seed=3)
    rng = np.random.default_rng(threads)
    for p in rng.choice(sc.npatch, 3, replace=False):
        a, fa = O.render_hemicube(64, sc.tri, sc.tri_patch, sc.frame[p], depth=True)[0::2]
        b, fb = O.render_hemicube_mt(64, sc.tri, sc.tri_patch, sc.frame[p], threads)
        assert np.array_equal(a, b) and fa == fb
    dff = O.dff_cross(64)
    u1, s1 = sc.emission.copy(), sc.shot0.copy()
    u2, s2 = sc.emission.copy(), sc.shot0.copy()
    for _ in range(4):
        i1 = O.plain_pass(64, 1, 1, sc, dff, u1, s1)
        i2 = O.plain_pass_mt(64, 1, 1, sc, dff, u2, s2, threads)
        assert i1 == i2
    assert np.array_equal(u1.view(np.uint32), u2.view(np.uint32))
    assert np.array_equal(s1.view(np.uint32), s2.view(np.uint32))



@pytest.mark.parametrize("wx", [4, 3, 1])
def test_triplet_stream_from_run_boundaries(wx):
    """orc_process_hemicube_triplets against run boundaries found with
    numpy (np.diff of each span): one triplet per run of a nonzero id, its
    ΔFF summed in texel order (a float32 cumulative sum), each work-item's
    triplets padded to a multiple of 4 with zeros, work-items in order; the
    triplets accumulated per id give the histogram's form factors."""
    W, K = 16, 3
    TW, TH, RES = O.geometry(W)
    rng = np.random.default_rng(wx)
    item = np.repeat(rng.integers(0, 6, size=K * RES), rng.integers(1, 9, size=K * RES))
    item = item[:K * RES].astype(np.uint32)
    dff = O.dff_cross(W)
    h, i, e = O.process_triplets(W, K, item, dff, workitems_x=wx)
    nx = min(wx, TW)
    span = TW // nx
    wh, wi, we = [], [], []
    for y in range(TH * K):
        for g in range(nx):
            x0 = min(g * span, TW - 1)
            x1 = min(x0 + span, TW)
            seg = item[y * TW + x0:y * TW + x1]
            cuts = np.r_[0, np.flatnonzero(np.diff(seg)) + 1, len(seg)]
            n0 = len(wh)
            for a, b in zip(cuts[:-1], cuts[1:]):
                if seg[a] == 0:
                    continue
                wh.append(y // TH)
                wi.append(int(seg[a]) - 1)
                we.append(np.cumsum(dff[(y % TH) * TW + x0 + a:(y % TH) * TW + x0 + b],
                                    dtype=np.float32)[-1])
            while (len(wh) - n0) % 4:
                wh.append(0)
                wi.append(0)
                we.append(np.float32(0))
    assert np.array_equal(h, np.array(wh, np.uint32))
    assert np.array_equal(i, np.array(wi, np.uint32))
    assert np.array_equal(e.view(np.uint32), np.array(we, np.float32).view(np.uint32))
    hits, ff64 = O.histogram(W, K, item, dff, 6)
    acc = np.zeros((K, 6))
    np.add.at(acc, (h, i), e.astype(np.float64))
    if wx in (4, 1):  # spans cover every texel
        assert np.allclose(acc, ff64, rtol=1e-5, atol=1e-12)


# ------------------------------------------------- top-K and the K-shooter pass

@pytest.mark.parametrize("seed", range(3))
def test_select_topk_power_is_a_stable_sort(seed):
    # POWER rule: key = area * (r + g + b) in float; patches without positive
    # power are never taken; equal keys keep index order (numpy lexsort)
    rng = np.random.default_rng(seed)
    n = 500
    u = rng.uniform(0.0, 1.0, size=(n, 3)).astype(np.float32)
    u[rng.random(n) < 0.3] = 0.0
    u[rng.random(n) < 0.05] *= -1.0
    a = rng.choice([0.5, 1.0, 2.0], n).astype(np.float32)
    u[10:20] = u[5]
    a[10:20] = a[5]
    key = a * ((u[:, 0] + u[:, 1]) + u[:, 2])
    pos = np.nonzero(key > 0)[0]
    order = pos[np.lexsort((pos, -key[pos]))]
    for count in (1, 7, 64, 600):
        ids, m = O.select_topk_power(u, a, count)
        want = order[:count]
        assert m == len(want)
        assert np.array_equal(ids[:m], want)
        assert not ids[m:].any()


@pytest.mark.parametrize("mode,rule", [(0, 0), (1, 1)])
def test_pass_k_single_shooter_equals_plain_pass(rad, mode, rule):
    # orc_pass_k at K = 1 and orc_plain_pass are separate code: same shooter,
    # same float state, and lastEnergy = the shooter's unshot before its
    # snapshot is taken away
    sc = rad.Scene.cornell(0.3127, triangle_patches=False, physical=bool(mode))
    W = 32
    dff = O.dff_cross(W)
    u1, s1 = sc.emission.copy(), sc.shot0.copy()
    u2, s2 = sc.emission.copy(), sc.shot0.copy()
    for _ in range(4):
        i = O.plain_pass(W, mode, rule, sc, dff, u1, s1)
        snap = u2.copy()
        ids, le = O.pass_k(W, mode, rule, 1, sc, dff, u2, s2)
        assert list(ids) == [i]
        assert np.array_equal(u1, u2) and np.array_equal(s1, s2)
        assert np.array_equal(u2[i], le - snap[i])


@pytest.mark.parametrize("mode", [0, 1])
def test_pass_k_against_fp64_shoot(rad, mode):
    # K shooters of one pass: F rows from the oracle's render and FP64
    # histogram, applied by orc_shoot in FP64 (pinned by the hand-computed
    # two-patch case and energy conservation) from the same start state
    sc = rad.Scene.cornell(0.3127, triangle_patches=False, physical=bool(mode))
    W, K = 32, 4
    dff = O.dff_cross(W)
    u, s = sc.emission.copy(), sc.shot0.copy()
    O.pass_k(W, mode, mode, K, sc, dff, u, s)  # a first pass spreads the energy
    u64, s64 = u.astype(np.float64), s.astype(np.float64)
    ids, le = O.pass_k(W, mode, mode, K, sc, dff, u, s)
    assert len(ids) == K
    item = np.concatenate([O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[i]) for i in ids])
    _, ff64 = O.histogram(W, K, item, dff, sc.npatch)
    last_before = u64[ids[-1]].copy()
    O.shoot(mode, ids, ff64, sc.refl, sc.area, u64, s64)
    scale = np.abs(u64).max()
    assert np.abs(u - u64).max() <= 1e-5 * scale
    assert np.abs(s - s64).max() <= 1e-5 * np.abs(s64).max()
    # lastEnergy: the last shooter after the pass's contributions, before its
    # own snapshot is taken away
    assert np.allclose(le, u64[ids[-1]] + last_before, rtol=1e-5, atol=1e-6 * scale)
