"""Parity at the benchmark's own workloads, full size.

  room200k   the 20 x 5 x 20 room with 400 boxes (seed 3), 199,936 triangle
             patches, hemicube 256, K = 64 shooters per pass, 32-texel tiles,
             in-lane limit 64 px: the configuration every headline number is
             measured on.  Item buffers bit-exact against the oracle's render
             (Main.cpp:1145-1200), hit counts bit-exact and F rows within 1e-4
             against the oracle's histogram (Kernel_ProcessHemicube.h:23-69),
             the fused pass against the FP64 shooting update (Main.cpp:1251-1295).
  cornell20k the Cornell layout split to 19,906 triangle patches, hemicube 256:
             plain single-shooter passes in lockstep with the oracle's pass.
  dense32k   32,760 patches, hemicube 128, the dense matrix in the rows
             layout and the 192-column gather: the FP64 residual of the
             solution on sampled rows of the FP32 form factors.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _close(a, b, scale):
    # relative to each entry, with a floor at 1e-4 of the field's scale
    return np.all(np.abs(a - b) <= RTOL * np.maximum(np.abs(b), scale * 1e-2))


@pytest.fixture(scope="module")
def room200k(rad):
    return rad.Scene.boxes(nboxes=400, room=(20.0, 5.0, 20.0), target_tris=200000, nlights=16,
                           seed=3)


def test_room200k_pass_against_oracle(rad, room200k):
    sc = room200k
    W, K = 256, 64
    say(f"room200k: {sc.npatch} patches, W={W}, K={K}")
    s = rad.Solver(hemicube=W, hemicubes=K, tile_px=32, small_tri_px=64, graph_passes=8)
    s.upload(sc)
    s.step(8)                                  # one captured graph of 8 passes
    s.finish()
    u0, sh0 = s.state()
    ids, n = s.select(K)
    assert n == K
    say("  render the 64 selected hemicubes")
    item = s.render(ids)
    c = s.debug_counters()
    bins, tile = s.debug_bins()
    say(f"  largest bin {int(bins.max())} records ({tile}-texel tiles); counters {c}")
    assert tile == 32
    assert bins.max() > 4096                    # multi-chunk bins are exercised
    assert c["overflow"] == 0 and c["sticky"] == 0 and c["unknown"] == 0
    assert c["spill"] <= c["spill_cap"] and c["entries"] <= c["entry_cap"]
    assert c["clip"] <= c["clip_cap"]
    res = s.tex_res
    # the 8 slots with the most binned records, and the first and last slot
    heavy = list(np.argsort(-bins.sum(1))[:6]) + [0, K - 1]
    for k in heavy:
        want, _ = O.render_hemicube_mt(W, sc.tri, sc.tri_patch, sc.frame[ids[k]], threads())
        mism = int(np.count_nonzero(item[k * res:(k + 1) * res] != want))
        say(f"  slot {k} shooter {ids[k]}: {mism} mismatching texels")
        assert mism == 0
    say("  ProcessHemicube over the 64 item buffers")
    hits, ff = s.process(K)
    h_o, ff_o = O.histogram(W, K, item, O.dff_cross(W), sc.npatch)
    assert np.array_equal(hits, h_o)
    assert np.allclose(ff, ff_o, rtol=RTOL, atol=1e-9)
    say("  the fused pass (tile raster + F rows + shoot) from the same state")
    s.set_state(u0, sh0, sc.emission)
    s.shoot(ids)
    c = s.debug_counters()
    assert c["overflow"] == 0 and c["sticky"] == 0
    u1, s1 = s.state()
    u64, s64 = u0.astype(np.float64), sh0.astype(np.float64)
    O.shoot(1, ids, ff_o, sc.refl, sc.area, u64, s64)
    scale = float(np.abs(u64).max())
    assert _close(u1, u64, scale) and _close(s1, s64, float(np.abs(s64).max()))


def test_cornell20k_plain_lockstep(rad):
    """20 plain passes (K = 1, POWER argmax, PHYSICAL) at hemicube 256 on the
    19,906-patch Cornell scene: the state after every pass against the
    oracle's CPU pass on float state."""
    sc = rad.Scene.cornell(0.02314, triangle_patches=True, physical=True, seed=1)
    W = 256
    say(f"cornell {sc.npatch} patches, W={W}")
    assert sc.npatch == 19906
    s = rad.Solver(hemicube=W, hemicubes=1)
    s.upload(sc)
    dff = O.dff_cross(W)
    u, sh = sc.emission.copy(), sc.shot0.copy()
    item = np.zeros(3 * W * W, np.uint32)
    ffs = np.zeros(sc.npatch, np.float32)
    for p in range(20):
        ids, n = s.select(1)
        want = O.plain_pass(W, 1, 1, sc, dff, u, sh, item, ffs)
        assert n == 1 and int(ids[0]) == int(want), (p, ids[0], want)
        s.step(1)
        gu, gs = s.state()
        assert _close(gu, u, float(np.abs(u).max())), p
        assert _close(gs, sh, float(np.abs(sh).max())), p
    say(f"  20 passes in lockstep, residual {s.residual():.4e}")


def test_dense32k_rows_gather_fp64_residual(rad):
    """The 32,760-patch room (seed 4) at hemicube 128: the dense matrix in
    the rows layout (M = F, row i = patch i's hemicube) and the gather of
    64 lighting conditions x RGB (192 columns) over 32 candidate lights to a
    residual of 1e-4.  On 512 sampled rows the FP64 residual
    E + diag(rho) F B - B of the answer, with the FP32 form-factor rows
    rendered again, is at most 1e-3 of |B| (area-weighted); the stored BF16
    halves of those rows equal the FP32 rows to 2^-15; two of the rows equal
    the oracle's histogram of the oracle's render."""
    sc = rad.Scene.boxes(nboxes=60, room=(12.0, 4.0, 12.0), target_tris=32768, nlights=16,
                         seed=4)
    n = sc.npatch
    say(f"dense32k: {n} patches, W=128")
    assert n == 32760
    s = rad.Solver(hemicube=128, hemicubes=64)
    s.upload(sc)
    s.dense_build(rad.RAD_DENSE_ROWS)
    rng = np.random.default_rng(4)
    cand = rng.choice(n, 32, replace=False)
    E = np.zeros((192, n), np.float32)
    for c in range(64):
        E[3 * c:3 * c + 3, cand] = rng.uniform(0.0, 20.0, (3, 32)) * (rng.random(32) < 0.5)
    B, st = s.gather(E, target=1e-4)
    say(f"  gather: {st.outer} steps, {st.sweeps} sweeps, residual {st.residual:.2e}")
    assert st.residual <= 1e-4
    rows = np.sort(np.random.default_rng(44).choice(n, 512, replace=False)).astype(np.uint32)
    F = np.concatenate([s.ff_rows(rows[b:b + 64]) for b in range(0, 512, 64)])
    stored = np.concatenate([s.dense_download(int(r), 1) for r in rows])
    scale = np.maximum(np.abs(F), 1e-30)
    assert np.all(np.abs(stored - F) <= scale * 2.0 ** -15 + 1e-12)
    B64 = B.astype(np.float64)                          # 192 x n
    rho = sc.refl.astype(np.float64)[rows][:, np.arange(192) % 3]   # 512 x 192
    r = E.T.astype(np.float64)[rows] + rho * (F.astype(np.float64) @ B64.T) - B64.T[rows]
    A = sc.area.astype(np.float64)[rows][:, None]
    rel = float((A * np.abs(r)).sum() / (A * np.abs(B64.T[rows])).sum())
    say(f"  FP64 residual on 512 sampled rows: {rel:.3e} (area-weighted)")
    assert rel <= 1e-3
    for k in (0, 511):
        p = int(rows[k])
        it = O.render_hemicube(128, sc.tri, sc.tri_patch, sc.frame[p])
        _, ff_o = O.histogram(128, 1, it, O.dff_cross(128), n)
        assert np.allclose(F[k], ff_o[0], rtol=RTOL, atol=1e-9), p
    s.dense_release()
