"""An FP64 ray-cast checker for the oracle's item-buffer render.

The oracle rasterizes each face of a hemicube under GL 3.0 rules (clipping,
8-bit snapping, top-left fill, GL_LESS on float depth; Main.cpp:1145-1200,
689-725, 316-389).  This checker asks the same question without any of that
machinery: for the centre of every texel, which front-facing triangle does
the ray through it hit first?  It works in homogeneous 2D (one line per
edge), in FP64, from the face MVPs:

  - C_i = MVP (v_i, 1), taken as (x, y, w) per vertex; det = det[C_0 C_1 C_2]
    is positive for a counter-clockwise (front-facing) triangle in window
    space, negative for a back face (culled);
  - a texel centre p = (x_ndc, y_ndc, 1) lies on the visible part of the
    triangle iff every edge function e_k(p) = (C_{k+1} x C_{k+2}) . p has the
    sign of det (then w > 0 at the hit; all opposite is the triangle's image
    behind the eye);
  - its depth is z_ndc = sum_k e_k(p) z_k / det (since sum_k e_k w_k = det),
    kept when -1 <= z_ndc (near plane) and the window depth < 1 (far plane
    left to the depth test, clear value 1.0).

The nearest such triangle wins, equal depths going to the one drawn first.
A texel is decided when the winner's edges, the near-plane line of every
candidate and the depth gap to the runner-up all clear a margin: then the
oracle's texel must equal the checker's.  Texels closer than the margins
(pixel centres within EDGE_EPS pixels of an edge, depth gaps below DEPTH_EPS)
may differ: there the snapping to 1/256 pixel and float arithmetic decide,
as in any GL implementation.
"""
import numpy as np
import pytest

from oracle import oracle as O

EDGE_EPS = 0.01      # pixels
DEPTH_EPS = 2e-6     # window depth
MIN_DECIDED = 0.97   # fraction of texels the margins must leave decided


def _face_texels(W, face):
    """Texels (tx, ty) of the cross layout that face owns (its scissor) and
    their pixel-centre NDC coordinates."""
    vp, sc = O.face_viewport(W, face)
    xs = np.arange(sc[0], sc[0] + sc[2])
    ys = np.arange(sc[1], sc[1] + sc[3])
    tx, ty = np.meshgrid(xs, ys)
    tx, ty = tx.ravel(), ty.ravel()
    xn = 2.0 * (tx + 0.5 - vp[0]) / W - 1.0
    yn = 2.0 * (ty + 0.5 - vp[1]) / W - 1.0
    return tx, ty, xn, yn


def raycast(W, tri, tri_patch, frame, chunk=128):
    """Returns (item, decided) over the TEX_W x TEX_H cross: the id + 1 the
    rays see (0: background) and whether the texel is decided."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return _raycast(W, tri, tri_patch, frame, chunk)


def _raycast(W, tri, tri_patch, frame, chunk):
    TW, TH, RES = O.geometry(W)
    item = np.zeros(RES, np.uint32)
    decided = np.zeros(RES, bool)
    tri = np.asarray(tri, np.float64).reshape(-1, 3, 3)
    ids = np.asarray(tri_patch, np.int64) + 1
    for face in range(5):
        m = O.face_mvp(frame, face).astype(np.float64).reshape(4, 4).T   # row-major M[row][col]
        tx, ty, xn, yn = _face_texels(W, face)
        n = tx.size
        P = np.stack([xn, yn, np.ones(n)], 1)                          # n x 3
        best = np.full(n, np.inf)
        second = np.full(n, np.inf)
        win = np.zeros(n, np.int64)
        win_edge = np.full(n, np.inf)
        for t0 in range(0, tri.shape[0], chunk):
            T = tri[t0:t0 + chunk]                                     # c x 3 x 3
            c = T.shape[0]
            Vh = np.concatenate([T, np.ones((c, 3, 1))], 2)           # c x 3 x 4
            Cl = Vh @ m.T                                              # clip coords, c x 3 x 4
            Cxyw = Cl[:, :, [0, 1, 3]]                                 # c x 3 x 3 (rows: vertices)
            det = np.linalg.det(Cxyw)                                  # c
            front = det > 0
            # edge k is opposite vertex k: the line through C_{k+1}, C_{k+2}
            L = np.stack([np.cross(Cxyw[:, (k + 1) % 3], Cxyw[:, (k + 2) % 3])
                          for k in range(3)], 1)                       # c x 3 (edges) x 3
            E = np.einsum("nj,ckj->nck", P, L)                         # n x c x 3
            s = np.sign(det)[None, :, None]
            ab = np.sqrt(L[:, :, 0] ** 2 + L[:, :, 1] ** 2)[None]      # 1 x c x 3
            dist = E * s / np.maximum(ab, 1e-300) * (W / 2.0)          # signed pixels, inside > 0
            zc = Cl[:, :, 2]                                           # c x 3
            zn = np.einsum("nck,ck->nc", E, zc) / det[None, :]         # NDC depth at p
            zw = 0.5 * zn + 0.5
            # the near-plane line: z_ndc + 1 = sum_k e_k (z_k + w_k) / det
            g = (zc + Cl[:, :, 3])                                     # c x 3
            nl = np.einsum("ckj,ck->cj", L, g) / det[:, None]          # c x 3
            nd = (zn + 1.0) / np.maximum(np.sqrt(nl[:, 0] ** 2 + nl[:, 1] ** 2), 1e-300)[None] \
                * (W / 2.0)                                            # pixels past the near line
            mind = dist.min(2)                                         # n x c
            inside = front[None] & (mind > 0) & (zn >= -1.0) & (zw < 1.0)
            zin = np.where(inside, zw, np.inf)
            # running best / second best over the chunks; ties to the first
            # triangle drawn (lowest index)
            order = np.argsort(zin, axis=1, kind="stable")
            z1 = np.take_along_axis(zin, order[:, :1], 1)[:, 0]
            z2 = np.take_along_axis(zin, order[:, 1:2], 1)[:, 0] if c > 1 else np.full(n, np.inf)
            i1 = order[:, 0] + t0
            better = z1 < best
            second = np.where(better, np.minimum(best, z2), np.minimum(second, z1))
            e1 = np.take_along_axis(np.minimum(mind, np.abs(nd)), order[:, :1], 1)[:, 0]
            win = np.where(better, i1, win)
            win_edge = np.where(better, e1, win_edge)
            best = np.where(better, z1, best)
        # decided: a winner clear of its edges and of the runner-up, and no
        # boundary-near candidate in front of or at its depth
        has = np.isfinite(best)
        got = np.where(has, ids[win], 0)
        clear = (~has | ((win_edge >= EDGE_EPS) & (second - best > DEPTH_EPS))) & \
            (~has | (np.abs(best) > DEPTH_EPS))
        t = ty * TW + tx
        item[t] = got
        # second sweep: a front-facing triangle whose edge or near-plane line
        # passes within EDGE_EPS of the centre, at or in front of the
        # winner's depth, leaves the texel undecided
        dec = clear.copy()
        for t0 in range(0, tri.shape[0], chunk):
            T = tri[t0:t0 + chunk]
            c = T.shape[0]
            Vh = np.concatenate([T, np.ones((c, 3, 1))], 2)
            Cl = Vh @ m.T
            Cxyw = Cl[:, :, [0, 1, 3]]
            det = np.linalg.det(Cxyw)
            front = det > 0
            L = np.stack([np.cross(Cxyw[:, (k + 1) % 3], Cxyw[:, (k + 2) % 3])
                          for k in range(3)], 1)
            E = np.einsum("nj,ckj->nck", P, L)
            s = np.sign(det)[None, :, None]
            ab = np.sqrt(L[:, :, 0] ** 2 + L[:, :, 1] ** 2)[None]
            dist = E * s / np.maximum(ab, 1e-300) * (W / 2.0)
            zc = Cl[:, :, 2]
            zn = np.einsum("nck,ck->nc", E, zc) / det[None, :]
            zw = 0.5 * zn + 0.5
            g = zc + Cl[:, :, 3]
            nl = np.einsum("ckj,ck->cj", L, g) / det[:, None]
            nd = (zn + 1.0) / np.maximum(np.sqrt(nl[:, 0] ** 2 + nl[:, 1] ** 2), 1e-300)[None] \
                * (W / 2.0)
            mind = dist.min(2)
            near_b = front[None] & (mind > -EDGE_EPS) & \
                (np.minimum(mind, np.abs(nd)) < EDGE_EPS) & (zw <= best[:, None] + DEPTH_EPS)
            dec &= ~near_b.any(1)
        decided[t] = dec
    return item, decided


def _compare(W, tri, tp, frame):
    want = O.render_hemicube(W, tri, tp, frame)
    got, dec = raycast(W, tri, tp, frame)
    bad = dec & (got != want)
    return int(bad.sum()), float(dec.mean()), int((got != want).sum())


def _soup(seed, n=400):
    """Random triangles around the origin: large ones spanning faces, ones
    crossing the near plane and the shooter's plane, tiny ones, both
    windings; a few reach far outside the viewport (the guard band)."""
    rng = np.random.default_rng(seed)
    centre = rng.uniform(-2.0, 2.0, (n, 3))
    big = rng.random(n) < 0.08
    scale = np.where(big, 10 ** rng.uniform(-0.3, 0.5, n), 10 ** rng.uniform(-1.5, -0.5, n))
    tri = centre[:, None, :] + rng.normal(size=(n, 3, 3)) * scale[:, None, None]
    return tri.astype(np.float32).reshape(n, 9), np.arange(n, dtype=np.uint32)


def _frame(eye, n, u):
    n = np.asarray(n, float)
    u = np.asarray(u, float)
    u = u - np.dot(u, n) / np.dot(n, n) * n
    return np.r_[eye, n, u].astype(np.float32)


@pytest.mark.parametrize("W", [32, 64])
def test_raycast_cornell_quads(rad, W):
    sc = rad.Scene.cornell(0.3127, triangle_patches=False, physical=False)
    light = int(np.argmax(sc.emission.sum(1)))
    for p in (light, 5, sc.npatch // 2):
        bad, frac, diff = _compare(W, sc.tri, sc.tri_patch, sc.frame[p])
        assert bad == 0, (p, bad)
        assert frac >= MIN_DECIDED, (p, frac)


@pytest.mark.parametrize("seed,W", [(0, 32), (1, 64), (2, 128)])
def test_raycast_random_soup(seed, W):
    """Near-plane crossings (the soup surrounds the eye at the origin),
    triangles behind the eye, guard-band vertices, every face."""
    tri, tp = _soup(seed)
    rng = np.random.default_rng(50 + seed)
    fr = _frame([0.0, 0.0, 0.0], rng.normal(size=3), rng.normal(size=3))
    bad, frac, diff = _compare(W, tri, tp, fr)
    assert bad == 0, bad
    assert frac >= 0.9, frac


def test_raycast_covers_near_plane_and_guard_band():
    """The soup cases above do reach the clipper: some front-facing
    triangles cross the near plane (a vertex with w < 0.01 and one beyond)
    and some have a vertex outside the guard band |x|, |y| > 4096 w."""
    tri, tp = _soup(0)
    fr = _frame([0.0, 0.0, 0.0], *np.random.default_rng(50).normal(size=(2, 3)))
    near = guard = 0
    for face in range(5):
        m = O.face_mvp(fr, face).astype(np.float64).reshape(4, 4).T
        C = np.concatenate([tri.reshape(-1, 3, 3), np.ones((len(tri), 3, 1))], 2) @ m.T
        z, w = C[:, :, 2], C[:, :, 3]
        near += int(((z < -w).any(1) & (z > -w).any(1)).sum())
        guard += int(((np.abs(C[:, :, 0]) > 4096 * w) & (w > 0)).any(1).sum())
    assert near > 0 and guard > 0


def test_raycast_detects_a_planted_mistake():
    """A checker that cannot fail pins nothing: with the oracle's item
    buffer shifted by one texel, decided texels disagree."""
    tri, tp = _soup(1)
    fr = _frame([0.0, 0.0, 0.0], [0.3, 0.2, 1.0], [0.0, 1.0, 0.0])
    W = 32
    want = O.render_hemicube(W, tri, tp, fr)
    got, dec = raycast(W, tri, tp, fr)
    assert np.count_nonzero(dec & (got != want)) == 0
    shifted = np.roll(want, 1)
    assert np.count_nonzero(dec & (got != shifted)) > 10
