"""GPU: dense form-factor matrix from the hemicube kernels and the BF16-MFMA
Jacobi gather with FP64-residual refinement, through the C ABI, against the
CPU oracle (FP64 histogram rows, FP64 Jacobi) on the same seeded inputs.

Tolerances: stored matrix (bf16 hi + lo, ~16 significant bits) within 1e-4
relative of the FP64 rows; BF16 products within 2^-7 of sum |M||x| per
element; refined products within 1e-4; radiosity within 1e-3 relative of
the FP64 fixed point."""
import sys

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cornell(rad):
    return rad.Scene.cornell(0.3127, triangle_patches=True, physical=True, seed=1)


@pytest.fixture(scope="module")
def built(rad, cornell):
    """Both layouts of the Cornell matrix at W=32 (K=16)."""
    s = rad.Solver(hemicube=32, hemicubes=16)
    s.upload(cornell)
    say(f"dense build rows N={cornell.npatch}")
    s.dense_build(rad.RAD_DENSE_ROWS)
    F = s.dense_download()
    say("dense build shoot")
    s.dense_build(rad.RAD_DENSE_SHOOT)
    G = s.dense_download()
    return s, F, G


@pytest.mark.parametrize("R", [3, 20, 48, 100, 192])
def test_dense_apply_random(rad, cornell, R):
    say(f"dense apply R={R}")
    rng = np.random.default_rng(R)
    n = cornell.npatch
    M = (rng.random((n, n)) / n).astype(np.float32)
    x = rng.uniform(-1.0, 2.0, (R, n)).astype(np.float32)
    s = rad.Solver(hemicube=16, hemicubes=1)
    s.upload(cornell)
    s.dense_upload(M)
    want = x.astype(np.float64) @ M.astype(np.float64).T          # (M x_r) per column
    mag = np.abs(x).astype(np.float64) @ np.abs(M).astype(np.float64).T
    y = s.dense_apply(x, split=False)
    assert np.max(np.abs(y - want) / mag) < 2.0 ** -7
    y2 = s.dense_apply(x, split=True)
    assert np.max(np.abs(y2 - want) / mag) < 1e-4
    # the stored matrix keeps ~16 bits: download within 2^-15 relative
    back = s.dense_download(0, 64)
    assert np.allclose(back, M[:64], rtol=2.0 ** -15, atol=0)


def test_dense_rows_equal_oracle_hemicubes(rad, cornell, built):
    s, F, G = built
    dff = O.dff_cross(32)
    for i in (0, 5, cornell.npatch // 2, cornell.npatch - 1):
        say(f"dense row {i}")
        item = O.render_hemicube(32, cornell.tri, cornell.tri_patch, cornell.frame[i])
        _, ff = O.histogram(32, 1, item, dff, cornell.npatch)
        assert np.allclose(F[i], ff[0], rtol=1e-4, atol=1e-7), i


def test_dense_shoot_layout_is_area_weighted_transpose(rad, cornell, built):
    s, F, G = built
    a = cornell.area.astype(np.float64)
    want = F.T.astype(np.float64) * a[None, :] / a[:, None]
    assert np.allclose(G, want, rtol=1e-4, atol=1e-7)


def test_dense_shoot_layout_graph_tail(rad, cornell):
    """K = 20: the shooting-layout build replays captured graphs of 8 passes
    over 77 passes (the last one partial, three passes of the last replay
    past the last patch); the matrix is still the area-weighted transpose of
    the rows layout, which builds eagerly."""
    s = rad.Solver(hemicube=32, hemicubes=20)
    s.upload(cornell)
    say(f"dense build K=20 N={cornell.npatch}")
    s.dense_build(rad.RAD_DENSE_ROWS)
    F = s.dense_download()
    s.dense_build(rad.RAD_DENSE_SHOOT)
    G = s.dense_download()
    a = cornell.area.astype(np.float64)
    want = F.T.astype(np.float64) * a[None, :] / a[:, None]
    assert np.allclose(G, want, rtol=1e-4, atol=1e-7)
    c = s.debug_counters()
    assert c["overflow"] == 0 and c["sticky"] == 0


@pytest.mark.parametrize("layout", ["rows", "shoot"])
def test_gather_matches_fp64_oracle(rad, cornell, built, layout):
    s, F, G = built
    M = F if layout == "rows" else G
    s.dense_upload(M)
    say(f"gather {layout}")
    B, st = s.gather(target=1e-6)
    assert st.residual <= 1e-6 and st.outer >= 2
    E = cornell.emission.T.astype(np.float64)
    want, it = O.gather_jacobi(M, cornell.refl, cornell.area, E, tol=1e-13)
    scale = np.maximum(np.abs(want), 1e-3 * np.abs(want).max())
    err = np.max(np.abs(B - want) / scale)
    say(f"  outer {st.outer} sweeps {st.sweeps} residual {st.residual:.2e} max rel err {err:.2e}")
    assert err < 1e-3


def test_gather_refinement_history_decreases(rad, cornell, built):
    """Each refinement step lowers the FP64-summed residual until the target
    (the floor of the ~16-bit residual product is about 5e-8 here)."""
    s, F, G = built
    s.dense_upload(G)
    B, st = s.gather(target=2e-7, inner=3)
    h = [st.history[k] for k in range(min(st.outer, 16))]
    say(f"history {h}")
    assert st.residual <= 2e-7 and st.outer >= 4
    assert all(b < a for a, b in zip(h, h[1:]))


def test_gather_matches_shooting(rad, cornell, built):
    """PHYSICAL shooting and the gather over the shooting layout reach the
    same radiosity (shot + unshot against E + diag(rho) G B)."""
    s, F, G = built
    s.dense_upload(G)
    B, st = s.gather(target=1e-7)
    sh = rad.Solver(hemicube=32, hemicubes=16)
    sh.upload(cornell)
    r = sh.run(100000, 1e-7, check_every=50)
    assert r.residual <= 1e-7
    u, shot = sh.state()
    tot = (u + shot).T.astype(np.float64)
    scale = np.maximum(np.abs(tot), 1e-3 * np.abs(tot).max())
    err = np.max(np.abs(B - tot) / scale)
    say(f"shooting vs gather max rel err {err:.2e}")
    assert err < 1e-3


def test_gather_many_rhs(rad):
    """192 columns (64 lighting conditions x RGB over 32 candidate lights)
    against the FP64 oracle on a coarse Cornell box."""
    sc = rad.Scene.cornell(1.0, triangle_patches=True, physical=True, seed=2)
    n = sc.npatch
    rng = np.random.default_rng(4)
    cand = rng.choice(n, 32, replace=False)
    E = np.zeros((192, n), np.float32)
    for c in range(64):
        E[3 * c:3 * c + 3, cand] = rng.uniform(0.0, 20.0, (3, 32)) * (rng.random(32) < 0.5)
    s = rad.Solver(hemicube=32, hemicubes=16)
    s.upload(sc)
    say(f"many rhs N={n}")
    s.dense_build(rad.RAD_DENSE_SHOOT)
    G = s.dense_download()
    B, st = s.gather(E, target=1e-6)
    want, _ = O.gather_jacobi(G, sc.refl, sc.area, E.astype(np.float64), tol=1e-12)
    scale = np.maximum(np.abs(want), 1e-3 * np.abs(want).max(axis=1, keepdims=True))
    err = np.max(np.abs(B - want) / scale)
    say(f"  outer {st.outer} residual {st.residual:.2e} max rel err {err:.2e}")
    assert st.residual <= 1e-6 and err < 1e-3


def test_gather_without_matrix_is_an_error(rad, cornell):
    s = rad.Solver(hemicube=16, hemicubes=1)
    s.upload(cornell)
    with pytest.raises(rad.RadError):
        s.gather()


def test_gather_check_fp64_matches_host(rad, cornell):
    """rad_gather_check_fp64 (rows rendered again and streamed on the
    device, FP64 accumulation) against the same residual and contraction
    computed on the host from FP32 rows, for a solution of the gather and
    for a perturbed one."""
    sc = cornell
    s = rad.Solver(hemicube=32, hemicubes=16)
    s.upload(sc)
    s.dense_build(rad.RAD_DENSE_SHOOT)
    B, st = s.gather(target=1e-6)
    n = sc.npatch
    F = np.concatenate([s.ff_rows(np.arange(b, min(n, b + 16), dtype=np.uint32))
                        for b in range(0, n, 16)]).astype(np.float64)
    A = sc.area.astype(np.float64)
    E = sc.emission.T.astype(np.float64)
    rho = sc.refl.T.astype(np.float64)
    q_host = float(np.max(F @ sc.refl.astype(np.float64)))
    for scale in (0.0, 1e-3):
        Bp = (B * (1.0 + scale * np.sin(np.arange(n)))[None, :]).astype(np.float32)
        Bd = Bp.astype(np.float64)
        r = E + rho * ((F.T * A[None, :]) @ Bd.T).T / A[None, :] - Bd
        want = float((A[None, :] * np.abs(r)).sum() / (A[None, :] * np.abs(Bd)).sum())
        res, q = s.gather_check_fp64(Bp)
        say(f"perturbation {scale}: residual {res:.3e} (host {want:.3e}), q {q:.4f}")
        # the rows here and on the device come from separate renders whose
        # float atomics round differently (~1e-7 of an entry)
        assert res == pytest.approx(want, rel=1e-3, abs=3e-7)
        assert q == pytest.approx(q_host, rel=1e-6)
    assert q < 1.0
