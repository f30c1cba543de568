"""The reference's stop rule and the F-row bookkeeping between passes.

Main.cpp:1137 runs the shooting loop while computeRadiosity holds, and
Main.cpp:1286-1300 clears it after the pass whose lastEnergy (the last
shooter's radiosity before its emitter update, so including what it received
in that pass) has f_Length() < 0.1.  rad_run_cycle and the drop-in
Shooter::shootCycle run that test on the device after every pass; the oracle
loop (orc_pass_k, float state) gives the pass it must stop on."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "radiosity_amd", "bin")
AREA, W, K = 0.3127, 32, 10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def flen(v):
    """Vector3f::f_Length in float."""
    v = np.asarray(v, np.float32)
    return float(np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))


@pytest.fixture(scope="module")
def scene(rad):
    return rad.Scene.cornell(AREA, triangle_patches=False, physical=False)


@pytest.fixture(scope="module")
def oracle_run(scene):
    """The oracle loop in REFERENCE mode (list rule, 0.3 * emitter colour):
    passes until lastEnergy.f_Length() < 0.1.  Returns the stopping pass
    (1-based), lastEnergy per pass and the state after every pass."""
    dff = O.dff_cross(W)
    u, s = scene.emission.copy(), scene.shot0.copy()
    energies, states, nshot = [], [], []
    for p in range(1, 2000):
        ids, le = O.pass_k(W, 0, 0, K, scene, dff, u, s)
        energies.append(np.array(le))
        states.append((u.copy(), s.copy()))
        nshot.append(len(ids))
        if flen(le) < 0.1:
            break
    stop = len(energies)
    # the test is meaningful only if the stopping pass is not a near tie
    assert flen(energies[-1]) < 0.1 * (1 - 1e-3) and flen(energies[-2]) > 0.1 * (1 + 1e-3)
    say(f"oracle stops after pass {stop} (lastEnergy {flen(energies[-1]):.5f}, "
        f"previous {flen(energies[-2]):.5f})")
    return stop, energies, states, nshot


def _close_state(got, want, tol=1e-4):
    (gu, gs), (wu, ws) = got, want
    scale = max(float(np.abs(wu).max()), float(np.abs(ws).max()))
    return (np.allclose(gu, wu, rtol=tol, atol=tol * scale)
            and np.allclose(gs, ws, rtol=tol, atol=tol * scale))


@pytest.mark.parametrize("graph,chunk", [(0, 500), (8, 500), (0, 20), (8, 20)])
def test_run_cycle_stops_on_the_oracle_pass(rad, scene, oracle_run, graph, chunk):
    """Cycles of `chunk` passes (SHOOTS_PER_CYCLE) until the stop test fires,
    eager and from captured graphs of 8 passes: the cycle that stops reports
    the oracle's stopping pass, lastEnergy and state; passes after it are
    no-ops."""
    stop, energies, states, nshot = oracle_run
    s = rad.Solver(hemicube=W, hemicubes=K, mode=rad.RAD_MODE_REFERENCE,
                   select=rad.RAD_SELECT_REFERENCE, shoots=chunk, graph_passes=graph)
    s.upload(scene)
    total, shooters, st = 0, 0, None
    for cyc in range(100):
        st = s.run_cycle(chunk, 0.1)
        total += st.passes
        shooters += st.shooters
        if st.stopped:
            break
    say(f"graph {graph} chunk {chunk}: stopped {st.stopped} after {total} passes")
    assert st.stopped == 1 and total == stop
    assert shooters == sum(nshot)
    assert flen(st.last_energy) < 0.1
    assert np.allclose(np.array(st.last_energy), energies[-1], rtol=1e-4,
                       atol=1e-4 * float(np.abs(energies[-1]).max()))
    assert _close_state(s.state(), states[-1])
    # the stop is per cycle: a new cycle runs again (and the flag is gone)
    before = s.shot_count()
    s.step(2)
    s.finish()
    after = s.shot_count()
    assert after[0] - before[0] == 2


def test_run_reports_last_energy_and_device_shooters(rad, scene, oracle_run):
    """rad_run without the stop test: lastEnergy of its last pass and the
    shooters the device shot, against the oracle after the same passes."""
    stop, energies, states, nshot = oracle_run
    n = 30
    s = rad.Solver(hemicube=W, hemicubes=K, mode=rad.RAD_MODE_REFERENCE,
                   select=rad.RAD_SELECT_REFERENCE)
    s.upload(scene)
    st = s.run(n, -1.0, check_every=n)
    assert st.passes == n and st.shooters == sum(nshot[:n])
    assert st.last_energy == pytest.approx(flen(energies[n - 1]), rel=1e-4)
    assert _close_state(s.state(), states[n - 1])


def test_dropin_shooter_stops_like_the_reference(oracle_run):
    """Shooter::shootCycle with SHOOTS_PER_CYCLE = 20: finished() after the
    cycle holding the oracle's stopping pass, no passes after it."""
    stop, energies, states, nshot = oracle_run
    p = subprocess.run([os.path.join(BIN, "dropin_check"), "shoot", str(AREA), str(W), str(K),
                        "20", "100"], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    d = json.loads(p.stdout.strip().splitlines()[-1])
    say(f"drop-in: {d['passes']} passes in {d['cycles']} cycles, finished {d['finished']}")
    assert d["finished"] == 1 and d["passes"] == stop and d["cycles"] == (stop + 19) // 20
    assert d["shooters"] == sum(nshot)
    assert d["last_energy"] == pytest.approx(flen(energies[-1]), rel=1e-4)
    got = (np.array(d["radiosity"], np.float32).reshape(-1, 3),
           np.array(d["illumination"], np.float32).reshape(-1, 3))
    assert _close_state(got, states[-1])


def test_f_rows_cleared_for_renders_of_fewer_rows(rad):
    """A writer of all K F rows (the standalone reduction) followed by a
    render or shoot of fewer rows must not leave rows behind that a later
    pass (eager or replayed) adds onto: both orders end where the same calls
    without the extra writer end (1e-4, float atomics)."""
    sc = rad.Scene.cornell(0.3127, triangle_patches=True, physical=True, seed=1)
    a = sc.area[:, None].astype(np.float64)

    def run(graph, dirty, middle):
        s = rad.Solver(hemicube=32, hemicubes=4, graph_passes=graph)
        s.upload(sc)
        s.step(2)
        if dirty:
            s.ff_rows([0, 5, 9, 100])              # K rows of F left behind
        if middle == "render":
            s.render([3])                          # one row
        else:
            s.shoot([3])                           # one row, and a shot
        s.step(8)
        s.finish()
        u, sh = s.state()
        return (u + sh).astype(np.float64)

    for middle in ("render", "shoot"):
        for graph in (0, 4):
            say(f"dirty F rows, then {middle} of one row, graph_passes={graph}")
            want = run(graph, False, middle)
            got = run(graph, True, middle)
            rel = np.abs(a * (got - want)).sum() / np.abs(a * want).sum()
            assert rel < 1e-4, (middle, graph, rel)
