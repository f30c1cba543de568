"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.  Item buffers and hit counts bit-exact; form factors and
radiosities within 1e-4 relative (FP32 against the FP64 oracle)."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

RTOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cornell_q(rad):
    return rad.Scene.cornell(0.3127, triangle_patches=False, physical=False)


@pytest.fixture(scope="module")
def cornell_t(rad):
    return rad.Scene.cornell(0.3127, triangle_patches=True, physical=True, seed=1)


@pytest.fixture(scope="module")
def boxes_small(rad):
    return rad.Scene.boxes(nboxes=30, room=(8.0, 3.0, 8.0), target_tris=6000, nlights=4, seed=3)


def _solver(rad, sc, W, K=1, mode=None, select=None):
    mode = rad.RAD_MODE_PHYSICAL if mode is None else mode
    select = rad.RAD_SELECT_POWER if select is None else select
    s = rad.Solver(hemicube=W, hemicubes=K, mode=mode, select=select)
    s.upload(sc)
    return s


@pytest.mark.parametrize("W", [16, 64, 256])
def test_dff_table_bit_exact(rad, W):
    say(f"dff W={W}")
    s = rad.Solver(hemicube=W)
    assert np.array_equal(s.dff().view(np.uint32), O.dff_cross(W).view(np.uint32))


def test_face_mvps_bit_exact(rad, cornell_q, cornell_t):
    for sc in (cornell_q, cornell_t):
        s = _solver(rad, sc, 16)
        for p in (0, 7, sc.npatch // 2, sc.npatch - 1):
            say(f"mvp patch={p}")
            m = s.face_mvps(p)
            for f in range(5):
                assert np.array_equal(m[f].view(np.uint32),
                                      O.face_mvp(sc.frame[p], f).view(np.uint32)), (p, f)


def _render_case(rad, sc, W, shooters):
    s = _solver(rad, sc, W, K=len(shooters))
    item, z = s.render(shooters, depth=True)
    res = s.tex_res
    for k, p in enumerate(shooters):
        it_o, z_o, _ = O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[p], depth=True)
        got = item[k * res:(k + 1) * res]
        mism = int(np.count_nonzero(got != it_o))
        say(f"  shooter {p}: {mism} mismatching texels of {res}")
        assert mism == 0
        gz = z[k * res:(k + 1) * res]
        covered = it_o > 0
        assert np.array_equal(gz[covered].view(np.uint32), z_o[covered].view(np.uint32))
    return s, item


@pytest.mark.parametrize("W", [16, 64, 128])
def test_itembuf_bit_exact_cornell_quads(rad, cornell_q, W):
    say(f"itembuf cornell quads W={W}")
    light = int(np.argmax(cornell_q.emission.sum(1)))
    _render_case(rad, cornell_q, W, [light, 0, 3, cornell_q.npatch - 1, 200])


@pytest.mark.parametrize("W", [32, 256])
def test_itembuf_bit_exact_cornell_triangles(rad, cornell_t, W):
    say(f"itembuf cornell triangles W={W}")
    _render_case(rad, cornell_t, W, [int(np.argmax(cornell_t.emission.sum(1))), 11, 700, 1500])


def test_itembuf_bit_exact_boxes(rad, boxes_small):
    say("itembuf boxes W=128")
    rng = np.random.default_rng(0)
    sh = [int(x) for x in rng.choice(boxes_small.npatch, 6, replace=False)]
    _render_case(rad, boxes_small, 128, sh)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_itembuf_bit_exact_w256_every_tile_side(rad, boxes_small, tile):
    # W=256 gives 3,072, 768 and 192 tiles per hemicube: the three ways the
    # bin fill runs (one tile per thread, four per thread, and the plain fill
    # beyond 1,024 tiles)
    say(f"itembuf boxes W=256 tile {tile}")
    rng = np.random.default_rng(5)
    shooters = [int(x) for x in rng.choice(boxes_small.npatch, 3, replace=False)]
    s = rad.Solver(hemicube=256, hemicubes=len(shooters), tile_px=tile, small_tri_px=64)
    s.upload(boxes_small)
    item = s.render(shooters)
    res = s.tex_res
    for k, p in enumerate(shooters):
        it_o = O.render_hemicube(256, boxes_small.tri, boxes_small.tri_patch, boxes_small.frame[p])
        assert np.array_equal(item[k * res:(k + 1) * res], it_o)


def test_itembuf_tuning_knobs_invariant(rad, boxes_small):
    # tile size and the in-lane pixel limit are tuning knobs: the item buffer
    # must not depend on them
    say("itembuf tile_px / small_tri_px sweep")
    outs = []
    for tile, px in ((8, 1), (16, 16), (32, 16), (32, 1024), (16, 4)):
        s = rad.Solver(hemicube=64, hemicubes=2, small_tri_px=px, tile_px=tile)
        s.upload(boxes_small)
        outs.append(s.render([5, 4000]))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    it_o = O.render_hemicube(64, boxes_small.tri, boxes_small.tri_patch, boxes_small.frame[5])
    assert np.array_equal(outs[0][:len(it_o)], it_o)


def test_many_bins_itembuf_and_fused_shoot(rad, boxes_small):
    # K=64 shooters with 8-pixel tiles: 64 x 192 = 12,288 (shooter, tile)
    # bins, so the bin scan runs over more than one LDS chunk, the bin fill
    # walks full groups of 8 shooter slots, and the clipper sees the near
    # plane of many shooters.  Item buffers bit-exact; the fused pass's
    # shooting update against the oracle's F rows.
    say("many bins: W=64 K=64 tile 8")
    W, K = 64, 64
    sc = boxes_small
    rng = np.random.default_rng(7)
    sh = [int(x) for x in rng.choice(sc.npatch, K, replace=False)]
    s = rad.Solver(hemicube=W, hemicubes=K, tile_px=8, small_tri_px=16)
    s.upload(sc)
    item = s.render(sh)
    res = s.tex_res
    for k, p in enumerate(sh):
        it_o = O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[p])
        assert np.array_equal(item[k * res:(k + 1) * res], it_o), (k, p)
    u0 = rng.random((sc.npatch, 3)).astype(np.float32)
    s.set_state(u0, np.zeros_like(u0), sc.emission)
    u0, _ = s.state()
    s.shoot(sh)
    u1, _ = s.state()
    _, ff_o = O.histogram(W, K, item, O.dff_cross(W), sc.npatch)
    snap = u0[sh].astype(np.float64)
    d = np.zeros_like(u0, dtype=np.float64)
    for k, i in enumerate(sh):
        d += sc.refl * snap[k] * ff_o[k][:, None] * (sc.area[i] / sc.area)[:, None]
    d[sh] -= snap
    want = u0.astype(np.float64) + d
    assert _close(u1, want, float(np.abs(want).max()))


def test_fused_reduce_matches_item_buffer_reduce(rad, cornell_t):
    # the shooting pass reduces tiles in the raster kernel; its F rows must
    # equal the reduction of the written item buffer
    say("fused vs item-buffer reduction")
    W, K = 64, 3
    sc = cornell_t
    s = _solver(rad, sc, W, K=K)
    sh = [int(np.argmax(sc.emission.sum(1))), 100, 900]
    item = s.render(sh)
    h_o, ff_o = O.histogram(W, K, item, O.dff_cross(W), sc.npatch)
    u0, s0 = s.state()
    s.shoot(sh)
    u1, _ = s.state()
    snap = u0[sh].astype(np.float64)
    d = np.zeros_like(u0, dtype=np.float64)
    for k, i in enumerate(sh):
        d += sc.refl * snap[k] * ff_o[k][:, None] * (sc.area[i] / sc.area)[:, None]
    d[sh] -= snap
    want = u0.astype(np.float64) + d
    scale = float(np.abs(want).max())
    assert _close(u1, want, scale)


def test_process_hemicube_hits_and_ff(rad, cornell_q):
    say("process hemicube K=3 W=64")
    W = 64
    sh = [int(np.argmax(cornell_q.emission.sum(1))), 10, 300]
    s, item = _render_case(rad, cornell_q, W, sh)
    hits, ff = s.process(len(sh))
    dff = O.dff_cross(W)
    h_o, ff_o = O.histogram(W, len(sh), item, dff, cornell_q.npatch)
    assert np.array_equal(hits, h_o)
    assert np.allclose(ff, ff_o, rtol=RTOL, atol=1e-9)
    # closed room: each row sums to the ΔFF table sum minus the cracks
    assert np.all(ff.sum(1) <= dff.sum(dtype=np.float64) * (1 + 1e-5))


def test_process_itembuf_given_buffer(rad, cornell_q):
    say("process caller item buffer with unknown ids")
    W = 32
    s = _solver(rad, cornell_q, W, K=2)
    rng = np.random.default_rng(1)
    item = rng.integers(0, cornell_q.npatch + 1, size=2 * s.tex_res).astype(np.uint32)
    item[:500] = 3                     # long runs
    item[1000] = cornell_q.npatch + 5  # unknown id: skipped (Main.cpp:1259)
    hits, ff = s.process_itembuf(item, 2)
    dff = O.dff_cross(W)
    h_o, ff_o = O.histogram(W, 2, item, dff, cornell_q.npatch)
    assert np.array_equal(hits, h_o)
    assert np.allclose(ff, ff_o, rtol=RTOL, atol=1e-9)


def _oracle_pass(sc, W, mode, shooters, u64, s64):
    dff = O.dff_cross(W)
    item = np.concatenate([O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[p]) for p in shooters])
    _, ff = O.histogram(W, len(shooters), item, dff, sc.npatch)
    O.shoot(mode, shooters, ff, sc.refl, sc.area, u64, s64)


def _close(a, b, scale):
    # relative to each entry, with a floor at 1e-4 of the field's scale
    return np.all(np.abs(a - b) <= RTOL * np.maximum(np.abs(b), scale * 1e-2))


@pytest.mark.parametrize("mode", [0, 1])
def test_shoot_lockstep(rad, cornell_t, cornell_q, mode):
    sc = cornell_q if mode == 0 else cornell_t
    W = 64
    s = _solver(rad, sc, W, K=1, mode=mode, select=0 if mode == 0 else 1)
    u64 = sc.emission.astype(np.float64).copy()
    s64 = sc.shot0.astype(np.float64).copy()
    for p in range(8):
        u, sh = s.state()
        ids, n = s.select(1)
        if mode == 0:
            want, _ = O.select_reference(u, 1)
        else:
            want = [O.select_argmax_power(u, sc.area)]
        say(f"mode {mode} pass {p}: shooter {ids[0]}")
        assert n == 1 and ids[0] == want[0]
        s.shoot(ids[:1])
        # oracle from the GPU's float state, in FP64
        u64 = u.astype(np.float64)
        s64 = sh.astype(np.float64)
        _oracle_pass(sc, W, mode, ids[:1], u64, s64)
        gu, gs = s.state()
        scale = float(np.abs(u64).max())
        assert _close(gu, u64, scale) and _close(gs, s64, scale)


def test_batched_topk_and_shoot(rad, cornell_t):
    W, K = 32, 8
    sc = cornell_t
    s = _solver(rad, sc, W, K=K)
    # a few single passes spread the energy, then one batched pass
    s.step(1)
    u, sh = s.state()
    ids, n = s.select(K)
    key = sc.area * u.astype(np.float32).sum(1)
    order = np.lexsort((np.arange(sc.npatch), -key))
    want = [int(i) for i in order[:K] if key[i] > 0]
    say(f"top-{K}: {list(ids[:n])}")
    assert n == len(want) and list(ids[:n]) == want
    s.shoot(ids[:n])
    u64, s64 = u.astype(np.float64), sh.astype(np.float64)
    _oracle_pass(sc, W, 1, ids[:n], u64, s64)
    gu, gs = s.state()
    scale = float(np.abs(u64).max())
    assert _close(gu, u64, scale) and _close(gs, s64, scale)


@pytest.mark.parametrize("K", [2, 33, 64, 100, 256])
def test_power_topk_with_ties(rad, boxes_small, K):
    """POWER top-K (wave lists for K <= 64, bitonic chunks above) against a
    lexicographic sort by (-A (r+g+b), index) on a state with many ties and
    zeros."""
    sc = boxes_small
    s = _solver(rad, sc, 16, K=K)
    rng = np.random.default_rng(K)
    u = rng.integers(0, 3, size=(sc.npatch, 3)).astype(np.float32)
    u[rng.random(sc.npatch) < 0.3] = 0
    s.set_state(u, np.zeros_like(u), u)
    ids, n = s.select(K)
    key = sc.area * u.sum(1, dtype=np.float32)
    order = np.lexsort((np.arange(sc.npatch), -key))
    want = [int(i) for i in order[:K] if key[i] > 0]
    assert n == len(want) and list(ids[:n]) == want
    assert np.all(ids[n:] == 0)


def test_power_topk_large_scene(rad):
    """More than 524,288 patches: the scan (8 wavefronts of 256 patches per
    workgroup) leaves more than 256 workgroup lists, so the top-64 path runs
    an intermediate merge level before the final one.  Same lexicographic
    reference as above."""
    sc = rad.Scene.boxes(nboxes=60, room=(12.0, 4.0, 12.0), target_tris=600000, nlights=4, seed=5)
    assert sc.npatch > 524288
    for K in (7, 64):
        say(f"top-{K} of {sc.npatch} patches")
        s = _solver(rad, sc, 16, K=K)
        rng = np.random.default_rng(100 + K)
        u = rng.integers(0, 4, size=(sc.npatch, 3)).astype(np.float32)
        u[rng.random(sc.npatch) < 0.5] = 0
        s.set_state(u, np.zeros_like(u), u)
        ids, n = s.select(K)
        key = sc.area * u.sum(1, dtype=np.float32)
        order = np.lexsort((np.arange(sc.npatch), -key))
        want = [int(i) for i in order[:K] if key[i] > 0]
        assert n == len(want) and list(ids[:n]) == want
        del s


def test_reference_rule_batched_matches_list(rad, cornell_q):
    say("reference selection K=10")
    s = _solver(rad, cornell_q, 16, K=10, mode=0, select=0)
    rng = np.random.default_rng(2)
    u = rng.integers(0, 3, size=(cornell_q.npatch, 3)).astype(np.float32)
    s.set_state(u, np.zeros_like(u), u)
    ids, n = s.select(10)
    want, wn = O.select_reference(u, 10)
    assert n == wn and np.array_equal(ids, want)


def test_residual_fp64(rad, cornell_t):
    s = _solver(rad, cornell_t, 32)
    s.step(3)
    u, _ = s.state()
    want = O.residual(u.astype(np.float64), cornell_t.emission, cornell_t.area)
    assert s.residual() == pytest.approx(want, rel=1e-9)


def test_run_converges_and_conserves(rad, cornell_t):
    say("run to 1e-2 residual")
    sc = cornell_t
    s = _solver(rad, sc, 32, K=4)
    st = s.run(20000, 1e-2, check_every=50)
    say(f"  passes {st.passes} residual {st.residual:.3e}")
    assert st.residual <= 1e-2
    u, sh = s.state()
    assert np.all(np.isfinite(u)) and np.all(u >= -1e-6) and np.all(sh >= 0)
    # emitted power ends up shot or still unshot, both weighted by area
    assert (sc.area[:, None] * (sh + u)).sum() >= (sc.area[:, None] * sc.emission).sum() * 0.999


def test_itembuf_dense_bins_bit_exact(rad):
    # a dense scene on a small hemicube with 8-pixel tiles: hundreds of
    # records per bin, many of them in several bins
    say("itembuf dense bins: 40k triangles, W=16, 8-pixel tiles")
    sc = rad.Scene.boxes(nboxes=60, room=(6.0, 3.0, 6.0), target_tris=40000, nlights=2, seed=7)
    s = rad.Solver(hemicube=16, hemicubes=3, tile_px=8)
    s.upload(sc)
    sh = [10, 20000, 39000]
    item = s.render(sh)
    bins, _ = s.debug_bins()
    say(f"  largest bin {int(bins.max())} records")
    assert bins.max() > 256
    res = s.tex_res
    for k, p in enumerate(sh):
        want = O.render_hemicube(16, sc.tri, sc.tri_patch, sc.frame[p])
        assert np.array_equal(item[k * res:(k + 1) * res], want)
    hits, ff = s.process(3)
    h_o, ff_o = O.histogram(16, 3, item, O.dff_cross(16), sc.npatch)
    assert np.array_equal(hits, h_o)
    s2 = rad.Solver(hemicube=16, hemicubes=3, tile_px=8)
    s2.upload(sc)
    u0, _ = s2.state()
    s2.shoot(sh)
    u1, _ = s2.state()
    snap = u0[sh].astype(np.float64)
    d = np.zeros_like(u0, dtype=np.float64)
    for k, i in enumerate(sh):
        d += sc.refl * snap[k] * ff_o[k][:, None] * (sc.area[i] / sc.area)[:, None]
    d[sh] -= snap
    want = u0.astype(np.float64) + d
    assert _close(u1, want, float(np.abs(want).max()))


def test_f_rows_cleared_after_other_writers(rad, cornell_t):
    """The shoot kernels leave the F rows zero and the next pass skips
    clearing them; a kernel that accumulates into F without a shoot after it
    (the standalone reduction, an autotune) must make the next pass clear
    again, eager or replayed.  Passes interrupted that way end where
    uninterrupted ones do (1e-4 relative, float atomics)."""
    ref = rad.Solver(hemicube=32, hemicubes=4)
    ref.upload(cornell_t)
    ref.step(12)
    ref.finish()
    ur, sr = ref.state()
    a = cornell_t.area[:, None].astype(np.float64)
    tot_r = (ur + sr).astype(np.float64)
    for graph in (0, 4):
        say(f"F rows dirtied between passes, graph_passes={graph}")
        s = rad.Solver(hemicube=32, hemicubes=4, graph_passes=graph)
        s.upload(cornell_t)
        s.step(4)
        s.ff_rows([0, 5, 9, 100])          # standalone reduction into F
        s.step(4)
        s.autotune(passes=1)               # render passes without a shoot
        s.step(4)
        s.finish()
        u, sh = s.state()
        tot = (u + sh).astype(np.float64)
        rel = np.abs(a * (tot - tot_r)).sum() / np.abs(a * tot_r).sum()
        assert rel < 1e-4, (graph, rel)


def test_graph_replay_matches_eager(rad, cornell_t):
    """rad_step replaying captured passes (graph_passes=4) tracks plain
    launches: every replay is checked after host work and a synchronisation
    (float atomics make the F rows order-dependent in the last bits, so the
    comparison is 1e-4 relative on the area-weighted state)."""
    e = rad.Solver(hemicube=32, hemicubes=4)
    g = rad.Solver(hemicube=32, hemicubes=4, graph_passes=4)
    e.upload(cornell_t)
    g.upload(cornell_t)
    a = cornell_t.area[:, None].astype(np.float64)
    for rnd, n in enumerate((8, 4, 5, 12)):
        say(f"graph round {rnd} passes {n}")
        e.step(n)
        g.step(n)
        e.finish()
        g.finish()
        ue, se = e.state()
        ug, sg = g.state()
        tot_e = (ue + se).astype(np.float64)
        tot_g = (ug + sg).astype(np.float64)
        rel = np.abs(a * (tot_e - tot_g)).sum() / np.abs(a * tot_e).sum()
        assert rel < 1e-4, (rnd, rel)
        assert abs(e.residual() - g.residual()) <= 1e-4 * e.residual()


def test_prepare_then_step_equals_step(rad, boxes_small):
    """rad_prepare captures the pass graph without running a pass: with F
    rows left dirty by a standalone reduction, eager passes after it, then
    replays, give the state of the same calls without it, bit for bit (the
    deterministic mode fixes the F sums)."""
    sc = boxes_small
    outs = []
    for prep in (False, True):
        say(f"prepare={prep}: 2 passes, reduction, 3 eager passes, 8 replayed")
        s = rad.Solver(hemicube=32, hemicubes=4, graph_passes=4)
        s.set_deterministic(True)
        s.upload(sc)
        s.step(2)
        s.render([0, 9, 77, 130])
        s.process(4)                      # leaves F rows [0, 4) dirty
        if prep:
            s.prepare()
        s.step(3)
        s.step(8)
        s.finish()
        u, sh = s.state()
        outs.append((u.view(np.uint32).copy(), sh.view(np.uint32).copy()))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("k", [15, 19])
def test_split_pass_item_buffers_bit_exact(rad, boxes_small, k):
    """A pass of 16 or more slots renders as two halves on two streams (here
    9 + 10 slots); 15 slots render whole.  Every item buffer bit-exact against
    the oracle, and no counter block overflows."""
    W = 64
    sc = boxes_small
    sh = [int(x) for x in np.random.default_rng(k).choice(sc.npatch, k, replace=False)]
    say(f"split pass: W={W} k={k}")
    s = rad.Solver(hemicube=W, hemicubes=24, tile_px=16, small_tri_px=32)
    s.upload(sc)
    item = s.render(sh)
    c = s.debug_counters()
    assert c["overflow"] == 0 and c["sticky"] == 0 and c["unknown"] == 0
    res = s.tex_res
    for j, p in enumerate(sh):
        it_o = O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[p])
        assert np.array_equal(item[j * res:(j + 1) * res], it_o), (j, p)


def test_autotune_picks_a_candidate_and_keeps_parity(rad, boxes_small):
    s = rad.Solver(hemicube=64, hemicubes=4)
    s.upload(boxes_small)
    say("autotune")
    t, sm, ms, cands = s.autotune(2)
    say(f"  chose tile {t} small {sm} at {ms:.3f} ms; {len(cands)} candidates")
    assert (t, sm) in cands and ms > 0 and cands[(t, sm)] == ms
    assert all(v < 0 or v >= ms for v in cands.values())
    shooters = [0, 17, boxes_small.npatch - 1]
    ids = s.render(shooters).reshape(3, -1)
    for k, p in enumerate(shooters):
        want = O.render_hemicube(64, boxes_small.tri, boxes_small.tri_patch, boxes_small.frame[p])
        assert np.array_equal(ids[k], want), p


def test_itembuf_bit_exact_max_hemicube(rad, cornell_q):
    """W = RAD_MAX_W (1024): 2048 x 1536 texels per hemicube, 3072 tiles."""
    say("itembuf cornell quads W=1024")
    light = int(np.argmax(cornell_q.emission.sum(1)))
    _render_case(rad, cornell_q, 1024, [light, cornell_q.npatch // 2])


def test_max_batch_rows_equal_oracle(rad, cornell_t):
    """K = RAD_MAX_K (256) shooters in one pass: every F row equals the
    oracle's FP64 histogram of that shooter's hemicube."""
    W, K = 16, 256
    sc = cornell_t
    s = _solver(rad, sc, W, K=K)
    rng = np.random.default_rng(7)
    sh = [int(x) for x in rng.choice(sc.npatch, K, replace=False)]
    say(f"K={K} render + process")
    item = s.render(sh)
    hits, ff = s.process(K)
    h_o, ff_o = O.histogram(W, K, item, O.dff_cross(W), sc.npatch)
    assert np.array_equal(hits, h_o)
    assert np.allclose(ff, ff_o, rtol=RTOL, atol=1e-9)
    res = s.tex_res
    for k in (0, 77, 255):
        want = O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[sh[k]])
        assert np.array_equal(item[k * res:(k + 1) * res], want), k


def test_empty_and_degenerate_scenes(rad):
    """The unsubdivided Cornell box with one triangle collapsed to a point:
    passes run and the item buffer still equals the oracle's."""
    one = rad.Scene.cornell(0.0, triangle_patches=True, physical=True, seed=1)
    tri = one.tri.reshape(-1, 3, 3).copy()
    tri[1] = tri[1][[0, 0, 0]]                      # a degenerate triangle (one point)
    sc = rad.Scene(tri.reshape(-1), one.tri_patch, one.frame, one.refl, one.area,
                   one.emission, one.shot0)
    s = _solver(rad, sc, 16, K=4)
    s.step(3)
    s.finish()
    it = s.render([1])
    want = O.render_hemicube(16, sc.tri, sc.tri_patch, sc.frame[1])
    assert np.array_equal(it, want)


def _sliver_scene(rad, rng, n_per_shooter=1500):
    """Cornell triangles plus, around four shooters, triangles whose planes
    pass close to the shooter's eye (nearly edge-on, both windings) and tiny
    ones: back-facing triangles whose snapped area can flip sign.  Each added
    triangle is its own patch."""
    base = rad.Scene.cornell(0.3127, triangle_patches=True, physical=True, seed=1)
    shooters = [int(x) for x in rng.choice(base.npatch, 4, replace=False)]
    tris = []
    for p in shooters:
        e = base.frame[p, 0:3].astype(np.float64)
        n = base.frame[p, 3:6].astype(np.float64)
        n /= np.linalg.norm(n)
        for _ in range(n_per_shooter):
            while True:
                dr = rng.normal(size=3)
                dr /= np.linalg.norm(dr)
                if dr @ n > 0.05:
                    break
            D = 10 ** rng.uniform(-1.7, 0.5)
            nrm = np.cross(dr, rng.normal(size=3))
            nrm /= np.linalg.norm(nrm)
            kind = rng.integers(4)
            if kind == 0:      # nearly edge-on, moderate size
                s = D * 10 ** rng.uniform(-3, -0.7)
                h = D * 10 ** rng.uniform(-7, -2)
            elif kind == 1:    # tiny, any orientation
                s = D * 10 ** rng.uniform(-5, -2.5)
                h = D * 10 ** rng.uniform(-3, -1)
            elif kind == 2:    # almost exactly edge-on
                s = D * 10 ** rng.uniform(-2, -0.5)
                h = D * 10 ** rng.uniform(-9, -5)
            else:              # around the snapping-error scale
                s = D * 10 ** rng.uniform(-3, -1)
                h = 10 ** rng.uniform(-1, 1) * 5.66 * (1 / 512) * D * D / (s * 32)
            h *= rng.choice([-1.0, 1.0])
            a = np.cross(nrm, dr)
            c = e + dr * D + h * nrm
            tris.append(np.concatenate([c + s * (rng.normal() * a + rng.normal() * dr)
                                        for _ in range(3)]))
    extra = np.array(tris, np.float32)
    ex = extra.reshape(-1, 3, 3).astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(ex[:, 1] - ex[:, 0], ex[:, 2] - ex[:, 0]), axis=1)
    keep = area > 1e-12
    extra, ex, area = extra[keep], ex[keep], area[keep]
    m = len(extra)
    nrm = np.cross(ex[:, 1] - ex[:, 0], ex[:, 2] - ex[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    up = ex[:, 1] - ex[:, 0]
    up /= np.linalg.norm(up, axis=1, keepdims=True)
    frame = np.concatenate([ex.mean(1), nrm, up], axis=1).astype(np.float32)
    npb = base.npatch
    sc = rad.Scene(np.concatenate([base.tri, extra]),
                   np.concatenate([base.tri_patch, np.arange(npb, npb + m, dtype=np.uint32)]),
                   np.concatenate([base.frame, frame]),
                   np.concatenate([base.refl, np.full((m, 3), 0.5, np.float32)]),
                   np.concatenate([base.area, area.astype(np.float32)]),
                   np.concatenate([base.emission, np.zeros((m, 3), np.float32)]),
                   np.concatenate([base.shot0, np.zeros((m, 3), np.float32)]))
    return sc, shooters


@pytest.mark.parametrize("W", [16, 64, 256])
def test_itembuf_bit_exact_slivers(rad, W):
    """Nearly edge-on and tiny triangles around the shooters (many of them
    back-facing with a snapped area that flips sign): the conservative culls
    must not change a texel."""
    say(f"itembuf slivers W={W}")
    sc, shooters = _sliver_scene(rad, np.random.default_rng(W))
    _render_case(rad, sc, W, shooters)


def test_itembuf_bit_exact_skewed_frame(rad, cornell_t):
    """A shooter frame whose up is not orthogonal to its normal (the side
    faces are then skewed): the culls that assume orthogonal face axes are
    switched off for it and the item buffer still equals the oracle's."""
    sc = cornell_t
    frame = sc.frame.copy()
    p = int(np.argmax(sc.emission.sum(1)))
    q = 700
    for i, t in ((p, 0.3), (q, -0.05)):
        n = frame[i, 3:6] / np.linalg.norm(frame[i, 3:6])
        frame[i, 6:9] = frame[i, 6:9] + t * n
    sk = rad.Scene(sc.tri, sc.tri_patch, frame, sc.refl, sc.area, sc.emission, sc.shot0)
    say("itembuf skewed frames W=64")
    _render_case(rad, sk, 64, [p, q, 11])


@pytest.mark.parametrize("W", [64, 128])
def test_itembuf_bit_exact_random_soup(rad, W):
    """A soup of random triangles around the shooters: large ones spanning
    several faces, ones crossing the near plane and the shooter's plane,
    ones behind the eye, both windings.  Exercises the clipped path and
    every cull next to its exact test."""
    rng = np.random.default_rng(100 + W)
    n = 3000
    centre = rng.uniform(-3.0, 3.0, (n, 3))
    big = rng.random(n) < 0.04
    scale = np.where(big, 10 ** rng.uniform(-0.5, 0.3, n), 10 ** rng.uniform(-2.5, -0.8, n))
    tri = (centre[:, None, :] + rng.normal(size=(n, 3, 3)) * scale[:, None, None]).astype(np.float32)
    t64 = tri.astype(np.float64)
    area = 0.5 * np.linalg.norm(np.cross(t64[:, 1] - t64[:, 0], t64[:, 2] - t64[:, 0]), axis=1)
    keep = area > 1e-10
    tri, t64, area = tri[keep], t64[keep], area[keep]
    m = len(tri)
    nrm = np.cross(t64[:, 1] - t64[:, 0], t64[:, 2] - t64[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    up = t64[:, 2] - t64[:, 0]
    up /= np.linalg.norm(up, axis=1, keepdims=True)
    frame = np.concatenate([t64.mean(1), nrm, up], axis=1).astype(np.float32)
    sc = rad.Scene(tri.reshape(-1), np.arange(m, dtype=np.uint32), frame,
                   np.full((m, 3), 0.5, np.float32), area.astype(np.float32),
                   np.zeros((m, 3), np.float32), np.zeros((m, 3), np.float32))
    small = np.nonzero(area < 0.01)[0]
    shooters = [int(x) for x in rng.choice(small, 4, replace=False)]
    say(f"itembuf random soup W={W}: {m} triangles, shooters {shooters}")
    _render_case(rad, sc, W, shooters)


@pytest.mark.parametrize("wx", [4, 3, 1])
def test_process_runs_equals_oracle_triplets(rad, boxes_small, wx):
    """rad_process_runs: the ProcessHemicube kernel's triplet stream for K
    hemicubes, bit for bit the oracle's restatement (Kernel_ProcessHemicube.h
    :23-69) at oclWorkitemsX 4, 3 (spans leave the last texels of a row out,
    as the reference's integer span does) and 1."""
    W, K = 64, 4
    sc = boxes_small
    s = rad.Solver(hemicube=W, hemicubes=K, workitems_x=wx)
    s.upload(sc)
    sh = [0, 17, sc.npatch // 2, sc.npatch - 1]
    item = s.render(sh)
    h, i, e = s.process_runs(K)
    ho, io, eo = O.process_triplets(W, K, item, O.dff_cross(W), workitems_x=wx)
    say(f"runs wx={wx}: {len(h)} triplets")
    assert len(h) == len(ho) and len(h) % 4 == 0
    assert np.array_equal(h, ho) and np.array_equal(i, io)
    assert np.array_equal(e.view(np.uint32), eo.view(np.uint32))
