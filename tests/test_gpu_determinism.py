"""Deterministic F rows (rad_set_deterministic): the form-factor entries are
accumulated in 64-bit fixed point with integer atomics, so a pass gives the
same bits whatever order its atomics land in.  Two independent runs must
agree bit for bit; both must still match the oracle (1e-4) and the default
float-atomic mode (1e-5)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cornell_t(rad):
    return rad.Scene.cornell(0.3127, triangle_patches=True, physical=True, seed=1)


@pytest.fixture(scope="module")
def boxes(rad):
    return rad.Scene.boxes(nboxes=30, room=(8.0, 3.0, 8.0), target_tris=6000, nlights=4, seed=3)


def _run(rad, sc, W, K, passes, det, graph=0):
    s = rad.Solver(hemicube=W, hemicubes=K, graph_passes=graph)
    if det:
        s.set_deterministic(True)
    s.upload(sc)
    s.step(passes)
    s.finish()
    return s.state(), s


@pytest.mark.parametrize("graph", [0, 4])
def test_passes_identical_run_to_run(rad, cornell_t, boxes, graph):
    for sc, W, K, n in ((cornell_t, 32, 4, 16), (boxes, 64, 16, 8)):
        say(f"deterministic {sc.npatch} patches W={W} K={K} passes={n} graph={graph}")
        (u1, s1), _ = _run(rad, sc, W, K, n, True, graph)
        (u2, s2), _ = _run(rad, sc, W, K, n, True, graph)
        assert np.array_equal(u1.view(np.uint32), u2.view(np.uint32))
        assert np.array_equal(s1.view(np.uint32), s2.view(np.uint32))
        (uf, sf), _ = _run(rad, sc, W, K, n, False, graph)
        a = sc.area[:, None].astype(np.float64)
        tot_d, tot_f = (u1 + s1).astype(np.float64), (uf + sf).astype(np.float64)
        rel = np.abs(a * (tot_d - tot_f)).sum() / np.abs(a * tot_f).sum()
        say(f"  deterministic vs float atomics: {rel:.2e}")
        assert rel < 1e-5


def test_reduction_identical_and_equal_to_oracle(rad, boxes):
    W, K = 64, 8
    sc = boxes
    sh = [int(x) for x in np.random.default_rng(5).choice(sc.npatch, K, replace=False)]
    outs = []
    for rep in range(2):
        s = rad.Solver(hemicube=W, hemicubes=K)
        s.set_deterministic(True)
        s.upload(sc)
        item = s.render(sh)
        hits, ff = s.process(K)                  # standalone reduction (k_reduce)
        outs.append((item, hits, ff))
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    item, hits, ff = outs[0]
    h_o, ff_o = O.histogram(W, K, item, O.dff_cross(W), sc.npatch)
    assert np.array_equal(hits, h_o)
    assert np.allclose(ff, ff_o, rtol=1e-4, atol=1e-9)
