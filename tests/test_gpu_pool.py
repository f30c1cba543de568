"""The chained POWER selection of captured passes against eager passes.

Inside a graph, a pass's selection either merges the per-workgroup lists
the previous shoot left or takes the K largest keys from the pool of every
receiver at or above a bound (when the pool holds K .. cap keys), else it
scans every patch.  Each must select exactly what the eager selection (a
scan of the whole state) selects: with the deterministic F sums, the state
after the same passes is the same bit for bit."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def boxes_mid(rad):
    return rad.Scene.boxes(nboxes=120, room=(12.0, 4.0, 12.0), target_tris=60000, nlights=8,
                           seed=5)


def _run(rad, sc, graph, passes, cap=0, K=64, mode=None):
    mode = rad.RAD_MODE_PHYSICAL if mode is None else mode
    s = rad.Solver(hemicube=32, hemicubes=K, graph_passes=graph, mode=mode)
    s.set_deterministic(True)
    s.upload(sc)
    if cap:
        s.debug_pool(cap)
    s.step(passes)
    s.finish()
    u, sh = s.state()
    return u.view(np.uint32).copy(), sh.view(np.uint32).copy(), s.debug_pool()


@pytest.mark.parametrize("graph,cap,K,ref", [(8, 0, 64, False), (8, 64, 64, False),
                                             (20, 0, 64, False), (8, 0, 24, False),
                                             (8, 0, 64, True)])
def test_chained_selection_bit_exact(rad, boxes_mid, graph, cap, K, ref):
    """40 passes replayed from graphs of 8 or 20 against 40 eager passes,
    K = 64 or 24, on 60k patches.  cap 0 keeps the default pool (most
    selections after the first two of every 8 passes of a graph come from
    it); cap 64 sends every pooled selection that finds more than 64 keys
    to the scan of all patches; ref: REFERENCE shooting (Main.cpp:1274)
    instead of PHYSICAL."""
    sc = boxes_mid
    mode = rad.RAD_MODE_REFERENCE if ref else rad.RAD_MODE_PHYSICAL
    say(f"chained selection: {sc.npatch} patches, K={K}, graph={graph} cap={cap}")
    ue, se, _ = _run(rad, sc, 0, 40, K=K, mode=mode)
    ug, sg, st = _run(rad, sc, graph, 40, cap, K, mode)
    say(f"  pool: {st}")
    assert np.array_equal(ue, ug)
    assert np.array_equal(se, sg)
    if cap:
        assert st["scans"] > 0
    else:
        assert st["hits"] > 0
    assert st["arms"] > 0


def test_pool_cap_bounds(rad, boxes_mid):
    """rad_debug_pool rejects a cap beyond 16,384 and reads the counters."""
    s = rad.Solver(hemicube=32, hemicubes=64, graph_passes=8)
    s.upload(boxes_mid)
    with pytest.raises(Exception):
        s.debug_pool(16385)
    st = s.debug_pool()
    assert st == dict(hits=0, scans=0, arms=0, armed=0)


def test_chained_passes_stop_like_eager(rad, boxes_mid):
    """The reference's per-pass stop test (the last shooter's |lastEnergy|
    below a threshold) inside captured chains, K = 64, graphs of 8 (lists,
    then the pool): a threshold that first holds at pass 4 or later stops the
    eager and the replayed runs on the same pass, with the same state bit for
    bit (deterministic F sums)."""
    sc = boxes_mid
    s = rad.Solver(hemicube=32, hemicubes=64)
    s.set_deterministic(True)
    s.upload(sc)
    e = []
    for _ in range(12):
        st = s.run_cycle(1, 0.0)
        e.append(float(np.linalg.norm(np.array(st.last_energy[:], np.float64))))
    say(f"last energies: {[round(x, 5) for x in e]}")
    # the first pass (index p) whose energy lies below all earlier ones, with
    # a gap wide enough for the device's float length
    p = next(i for i in range(3, 12) if e[i] < 0.999 * min(e[:i]))
    thr = 0.5 * (e[p] + min(e[:p]))
    outs = []
    for graph in (0, 8):
        s = rad.Solver(hemicube=32, hemicubes=64, graph_passes=graph)
        s.set_deterministic(True)
        s.upload(sc)
        st = s.run_cycle(40, thr)
        u, sh = s.state()
        say(f"graph {graph}: stopped {st.stopped} after {st.passes} passes")
        outs.append((st.stopped, st.passes, u.view(np.uint32).copy(), sh.view(np.uint32).copy()))
    assert outs[0][0] == 1 and outs[1][0] == 1
    assert outs[0][1] == outs[1][1] == p + 1
    assert np.array_equal(outs[0][2], outs[1][2])
    assert np.array_equal(outs[0][3], outs[1][3])


def test_chained_selection_headline_workload(rad):
    """The bench's workload: room200k (400 boxes, seed 3, 199,936 patches),
    W = 256, K = 64, tile 32, in-lane limit 64, graphs of 10.  30 replayed
    passes (lists, then the pool) equal 30 eager passes bit for bit in the
    deterministic mode."""
    sc = rad.Scene.boxes(nboxes=400, room=(20.0, 5.0, 20.0), target_tris=200000, nlights=16,
                         seed=3)
    outs = []
    for graph in (0, 10):
        say(f"room200k: {sc.npatch} patches, graph={graph}")
        s = rad.Solver(hemicube=256, hemicubes=64, tile_px=32, small_tri_px=64,
                       graph_passes=graph)
        s.set_deterministic(True)
        s.upload(sc)
        s.step(30)
        s.finish()
        u, sh = s.state()
        outs.append((u.view(np.uint32).copy(), sh.view(np.uint32).copy(), s.debug_pool()))
    say(f"  pool: {outs[1][2]}")
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert outs[1][2]["hits"] > 0 and outs[1][2]["scans"] == 0
