"""GPU: scene packs (many scenes, one launch per kernel) against the same
scenes solved one context each, through the C ABI."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scenes(rad):
    return [rad.Scene.cornell(0.5, triangle_patches=True, physical=True, seed=1),
            rad.Scene.boxes(nboxes=6, room=(5.0, 3.0, 5.0), target_tris=1500, nlights=3, seed=5),
            rad.Scene.cornell(1.0, triangle_patches=True, physical=True, seed=2),
            rad.Scene.boxes(nboxes=10, room=(6.0, 3.0, 6.0), target_tris=3000, nlights=4, seed=6)]


def _alone(rad, sc, W, k, target):
    s = rad.Solver(hemicube=W, hemicubes=k)
    s.upload(sc)
    st = s.run(100000, target, check_every=1)
    u, sh = s.state()
    return u + sh, st.passes


@pytest.mark.parametrize("graph", [0, 4])
def test_pack_matches_separate_solves(rad, scenes, graph):
    W, slots, target = 32, [3, 4, 2, 5], 1e-3
    p = rad.Solver(hemicube=W, hemicubes=16, graph_passes=graph)
    p.pack_upload(scenes, slots)
    p.pack_begin(target)
    for it in range(5000):
        p.step(8)
        done, res, passes, rem = p.pack_status()
        if rem == 0:
            break
    say(f"graph {graph}: pack done after {passes} passes; per scene {list(done)}")
    assert rem == 0 and np.all(res <= target)
    u, sh = p.state()
    tot = u + sh
    _, pb = p.pack_offsets
    for k, sc in enumerate(scenes):
        want, npass = _alone(rad, sc, W, slots[k], target)
        got = tot[pb[k]:pb[k + 1]]
        a = sc.area[:, None].astype(np.float64)
        rel = np.abs(a * (got - want)).sum() / np.abs(a * want).sum()
        say(f"  scene {k}: {sc.npatch} patches, done at pass {done[k]} (alone {npass}), rel {rel:.1e}")
        # the pass count at which the scene's residual was first <= target
        assert abs(int(done[k]) - npass) <= 1
        assert rel < 1e-4


def test_pack_rejects_bad_layouts(rad, scenes):
    p = rad.Solver(hemicube=16, hemicubes=4)
    with pytest.raises(rad.RadError):
        p.pack_upload(scenes, [2, 2, 2, 2])           # 8 shooters per pass > K = 4
    with pytest.raises(rad.RadError):
        p.pack_upload(scenes[:2], [0, 1])             # a scene without shooters
    p.pack_upload(scenes[:2], [2, 2])
    with pytest.raises(rad.RadError):
        p.select(2)                                   # not with a pack loaded
    with pytest.raises(rad.RadError):
        p.dense_build()
