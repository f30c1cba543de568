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
        p.pack_upload(scenes[:2], [5, 1])             # a quota above K = 4
    with pytest.raises(rad.RadError):
        p.pack_upload(scenes[:2], [0, 1])             # a scene without shooters
    p.pack_upload(scenes[:2], [2, 2])
    with pytest.raises(rad.RadError):
        p.select(2)                                   # not with a pack loaded
    with pytest.raises(rad.RadError):
        p.dense_build()
    with pytest.raises(rad.RadError):
        p.pack_hemicubes([16, 18])                    # above the context's side
    with pytest.raises(rad.RadError):
        p.pack_hemicubes([16, 7])                     # odd


def _mixed_scenes(rad, n):
    """n small scenes of three sizes, hemicube 64 / 128 / 256 by size."""
    out, ws = [], []
    for i in range(n):
        kind = i % 3
        if kind == 0:
            sc = rad.Scene.cornell(1.0 if i % 2 else 0.5, triangle_patches=True, physical=True,
                                   seed=1 + i)
        else:
            sc = rad.Scene.boxes(nboxes=4 + 3 * kind, room=(4.0 + kind, 3.0, 4.0 + kind),
                                 target_tris=800 * kind, nlights=2 + kind, seed=20 + i)
        out.append(sc)
        ws.append((64, 128, 256)[kind])
    return out, ws


def test_pack_mixed_resolutions_first_pass_against_oracle(rad):
    """18 scenes at hemicube 64, 128 and 256 in one pack (one launch per
    kernel): after the first pass every scene's state equals the oracle's
    pass at the scene's own resolution (POWER top-k in the scene, FP64
    shooting of the oracle's histogram rows), 1e-4."""
    from oracle import oracle as O
    scenes, ws = _mixed_scenes(rad, 18)
    quotas = [1 + i % 4 for i in range(18)]          # 45 shooters per pass <= K
    p = rad.Solver(hemicube=256, hemicubes=64)
    p.pack_upload(scenes, quotas)
    p.pack_hemicubes(ws)
    p.pack_begin(1e-3)
    p.step(1)
    p.finish()
    u, sh = p.state()
    _, pb = p.pack_offsets
    for k, sc in enumerate(scenes):
        ids, n = O.select_topk_power(sc.emission, sc.area, quotas[k])
        ids = ids[:n]
        W = ws[k]
        item = np.concatenate([O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[i])
                               for i in ids])
        _, ff = O.histogram(W, n, item, O.dff_cross(W), sc.npatch)
        u64, s64 = sc.emission.astype(np.float64), sc.shot0.astype(np.float64)
        O.shoot(1, ids, ff, sc.refl, sc.area, u64, s64)
        gu, gs = u[pb[k]:pb[k + 1]], sh[pb[k]:pb[k + 1]]
        scale = float(np.abs(u64).max())
        ok = np.all(np.abs(gu - u64) <= 1e-4 * np.maximum(np.abs(u64), 1e-2 * scale))
        ok &= np.all(np.abs(gs - s64) <= 1e-4 * np.maximum(np.abs(s64), 1e-2 * scale))
        say(f"  scene {k}: W={W}, {sc.npatch} patches, {n} shooters, match {bool(ok)}")
        assert ok, k


def test_pack_mixed_resolutions_match_separate_solves(rad):
    """The same mixed pack run to residual 1e-2 against each scene solved
    alone in its own context at its own resolution with its quota: same
    finishing pass (+-1) and radiosity within 1e-4 (area-weighted L1)."""
    scenes, ws = _mixed_scenes(rad, 18)
    quotas = [1 + i % 4 for i in range(18)]
    target = 1e-2
    p = rad.Solver(hemicube=256, hemicubes=64, graph_passes=4)
    p.pack_upload(scenes, quotas)
    p.pack_hemicubes(ws)
    p.pack_begin(target)
    for _ in range(5000):
        p.step(8)
        done, res, passes, rem = p.pack_status()
        if rem == 0:
            break
    assert rem == 0 and np.all(res <= target)
    u, sh = p.state()
    tot = u + sh
    _, pb = p.pack_offsets
    for k, sc in enumerate(scenes):
        want, npass = _alone(rad, sc, ws[k], quotas[k], target)
        got = tot[pb[k]:pb[k + 1]]
        a = sc.area[:, None].astype(np.float64)
        rel = np.abs(a * (got - want)).sum() / np.abs(a * want).sum()
        say(f"  scene {k}: W={ws[k]} done at pass {done[k]} (alone {npass}), rel {rel:.1e}")
        assert abs(int(done[k]) - npass) <= 1
        assert rel < 1e-4


def test_pack_water_filling(rad):
    """Quotas summing to more than K: every pass shares the K slots out by
    water-filling (each active scene min(wanted, L), the leftovers one each
    from a start scene that rotates), and the pack still solves every
    scene.  The first pass's shooter counts against a host restatement."""
    from oracle import oracle as O
    scenes, ws = _mixed_scenes(rad, 12)
    quotas = [8] * 6 + [1, 2, 3, 4, 5, 6]            # 69 wanted, K = 24
    K = 24
    p = rad.Solver(hemicube=256, hemicubes=K)
    p.pack_upload(scenes, quotas)
    p.pack_hemicubes(ws)
    p.pack_begin(1e-2)
    p.step(1)
    p.pack_status()
    got = p.pack_shooters.astype(int)
    wanted = [int(O.select_topk_power(sc.emission, sc.area, q)[1]) for sc, q in zip(scenes, quotas)]
    L = max(l for l in range(0, max(wanted) + 1) if sum(min(w, l) for w in wanted) <= K)
    left = K - sum(min(w, L) for w in wanted)
    more = [k for k, w in enumerate(wanted) if w > L]       # rotation starts at scene 0
    want = [min(w, L) + (1 if k in more[:left] else 0) for k, w in enumerate(wanted)]
    say(f"wanted {wanted}, level {L}, leftover {left}: {list(got)} vs {want}")
    assert list(got) == want and sum(want) == min(K, sum(wanted))
    for _ in range(5000):
        p.step(8)
        done, res, passes, rem = p.pack_status()
        if rem == 0:
            break
    assert rem == 0 and np.all(res <= 1e-2)
    u, sh = p.state()
    _, pb = p.pack_offsets
    for k, sc in enumerate(scenes):
        e = (sc.area[:, None] * sc.emission).sum()
        t = (sc.area[:, None] * (u[pb[k]:pb[k + 1]] + sh[pb[k]:pb[k + 1]])).sum()
        assert t >= e * 0.999
