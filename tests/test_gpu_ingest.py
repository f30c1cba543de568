"""GPU scene ingest, vertex colours and viewpoint renders through the C ABI:
the device-built scene equals the host builder bit for bit, the smoothed
vertex colours equal the restatement of Colors::smoothShadePatch, and a
viewpoint item buffer equals the oracle's render."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from oracle import scene_io as S

pytestmark = pytest.mark.gpu

F32 = np.float32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32) if a.dtype == np.float32 else a


def _mixed_faces(rad, rng):
    """Cornell faces plus random quads and triangles (C == D)."""
    q, c, e, s = rad.cornell_faces()
    extra = []
    for k in range(40):
        o = rng.uniform(0.5, 4.5, 3)
        e1 = rng.normal(size=3) * rng.uniform(0.1, 1.0)
        e2 = np.cross(e1, rng.normal(size=3))
        e2 *= rng.uniform(0.1, 1.0) / np.linalg.norm(e2)
        if k % 2:
            extra.append(np.concatenate([o, o + e1, o + e2, o + e2]))
        else:
            extra.append(np.concatenate([o, o + e1, o + e1 + e2, o + e2]))
    extra = np.array(extra, F32)
    n = len(extra)
    col = rng.uniform(0.1, 0.9, (n, 3)).astype(F32)
    em = np.zeros((n, 3), F32)
    em[3] = [5, 4, 3]
    return (np.concatenate([q, extra]), np.concatenate([c, col]), np.concatenate([e, em]),
            np.concatenate([s, np.zeros((n, 3), F32)]))


@pytest.mark.parametrize("mode", [0, 1])
def test_ingest_equals_host_builder(rad, mode):
    say(f"ingest mode {mode}")
    q, c, e, s = _mixed_faces(rad, np.random.default_rng(3))
    host = rad.Scene.faces(q, 0.05, mode, c, e, s)
    sol = rad.Solver(hemicube=32, hemicubes=4)
    nt, np_, nq = sol.ingest(q, 0.05, mode, c, e, s)
    assert (nt, np_, nq) == (host.ntri, host.npatch, host.nquad)
    dev = sol.download()
    for name, want in (("tri", host.tri), ("tri_patch", host.tri_patch), ("frame", host.frame),
                       ("refl", host.refl), ("area", host.area), ("emission", host.emission),
                       ("shot0", host.shot0), ("patch_quad", host.patch_quad),
                       ("quad_nbr", host.quad_nbr)):
        got = getattr(dev, name)
        assert got.shape == want.shape and np.array_equal(_bits(got), _bits(want)), name


def test_ingested_scene_shoots_like_uploaded(rad):
    q, c, e, s = _mixed_faces(rad, np.random.default_rng(4))
    host = rad.Scene.faces(q, 0.08, 1, c, e, s)
    a = rad.Solver(hemicube=64, hemicubes=8)
    a.ingest(q, 0.08, 1, c, e, s)
    b = rad.Solver(hemicube=64, hemicubes=8)
    b.upload(host)
    for sol in (a, b):
        sol.step(6)
        sol.finish()
    ua, sa = a.state()
    ub, sb = b.state()
    # same arrays, same passes; F rows are summed with float atomics, so the
    # last bits may differ between any two runs
    assert np.allclose(ua, ub, rtol=1e-5, atol=1e-7) and np.allclose(sa, sb, rtol=1e-5, atol=1e-7)
    assert np.abs(sa).sum() > 0


@pytest.mark.parametrize("mode", [0, 1])
def test_vertex_colors_match_restatement(rad, mode):
    say(f"vertex colours mode {mode}")
    q, c, e, s = _mixed_faces(rad, np.random.default_rng(5))
    sol = rad.Solver(hemicube=64, hemicubes=8)
    sol.ingest(q, 0.06, mode, c, e, s)
    sol.step(5)
    sol.finish()
    got = sol.vertex_colors()
    dev = sol.download()
    u, sh = sol.state()
    val = S.quad_values(dev.patch_quad, len(dev.quad_nbr), dev.refl, dev.area, u, sh)
    want = S.smooth_shade(val, dev.quad_nbr)
    assert got.shape == want.shape
    assert np.array_equal(_bits(got), _bits(want))
    assert np.abs(got).max() > 0


def test_vertex_colors_of_uploaded_cornell(rad):
    sc = rad.Scene.cornell(0.3127, triangle_patches=False, physical=False)
    sol = rad.Solver(hemicube=64, hemicubes=4, mode=rad.RAD_MODE_REFERENCE,
                     select=rad.RAD_SELECT_REFERENCE)
    sol.upload(sc)
    with pytest.raises(rad.RadError):
        sol.vertex_colors()                   # no quad topology yet
    sol.set_quads(sc.patch_quad, sc.quad_nbr)
    sol.step(3)
    sol.finish()
    got = sol.vertex_colors()
    u, sh = sol.state()
    want = S.smooth_shade(S.quad_values(sc.patch_quad, sc.nquad, sc.refl, sc.area, u, sh),
                          sc.quad_nbr)
    assert np.array_equal(_bits(got), _bits(want))


def test_render_view_equals_oracle(rad):
    sc = rad.Scene.cornell(0.3127, triangle_patches=True, physical=True, seed=1)
    sol = rad.Solver(hemicube=128, hemicubes=2)
    sol.upload(sc)
    rng = np.random.default_rng(6)
    for _ in range(3):
        eye = rng.uniform(1.0, 4.5, 3)
        d = rng.normal(size=3)
        up = np.cross(d, rng.normal(size=3))
        fr = np.concatenate([eye, d, up]).astype(F32)
        say(f"render_view eye {eye}")
        got = sol.render_view(fr)
        want = O.render_hemicube(128, sc.tri, sc.tri_patch, fr)
        assert np.array_equal(got, want)
    # the shooters' renders are unaffected by a viewpoint render
    it = sol.render([5])
    assert np.array_equal(it, O.render_hemicube(128, sc.tri, sc.tri_patch, sc.frame[5]))


def test_checkpoint_across_ingest_and_upload(rad, tmp_path):
    q, c, e, s = _mixed_faces(rad, np.random.default_rng(7))
    host = rad.Scene.faces(q, 0.1, 0, c, e, s)
    a = rad.Solver(hemicube=32, hemicubes=4)
    a.ingest(q, 0.1, 0, c, e, s)
    a.step(3)
    a.finish()
    path = str(tmp_path / "ck.bin")
    a.checkpoint_save(path)
    b = rad.Solver(hemicube=32, hemicubes=4)
    b.upload(host)
    b.checkpoint_load(path)
    ua, sa = a.state()
    ub, sb = b.state()
    assert np.array_equal(_bits(ua), _bits(ub)) and np.array_equal(_bits(sa), _bits(sb))


def test_ingest_rejects_degenerate_faces(rad):
    sol = rad.Solver(hemicube=16)
    flat = np.zeros((1, 12), F32)             # all four corners at the origin
    with pytest.raises(rad.RadError):
        sol.ingest(flat, 0.0, 1)
    with pytest.raises(rad.RadError):
        sol.ingest(flat, 0.1, 0)              # Patch::divide gives it no children
    with pytest.raises(rad.RadError):
        sol.ingest(np.full((1, 12), np.nan, F32), 0.1, 0)
    # the context stays usable
    q, c, e, s = rad.cornell_faces()
    assert sol.ingest(q, 0.5, 0, c, e, s)[1] > 17
