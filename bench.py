"""Benchmark of the hemicube shooting path on one MI355X.

Workload (default): the 20 x 5 x 20 room with 400 seeded boxes, ~200k
triangle patches, 16 emitters, hemicube 256.  A step is one pass of K=64
shooters (select, render, reduce, shoot).  Reported:
  value          batched shooters/s (K=64, FP32)
  plain          the single-shooter drop-in (K=1): shooters/s
  converge       wall ms until the FP64 residual unshot energy <= 1e-3,
                 batched and plain (plain is capped; see its fields)
  roofline       the dominant kernel, timed with HIP events on the library's
                 stream, against the MI355X HBM peak
  cpu_baseline   the oracle's single-thread CPU pass on the same scene
  cpu_baseline_multicore  the same pass on all host cores (OpenMP)
  rasterizer     (triangle, hemicube) pairs set up per second and covered
                 fragments tested per second (fragments per hemicube counted
                 by the oracle on sample shooters)

    python bench.py [--steps 50] [--warmup 5] [--no-converge] [--cpu-sample-s 10]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def host_threads():
    """Host cores this process may use (OMP_NUM_THREADS caps it)."""
    return max(1, min(len(os.sched_getaffinity(0)),
                      int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _profile_line(name, key):
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return json.loads(line.split(":", 1)[1])
    except OSError:
        pass
    return {}


def measured_traffic(W, K):
    """HBM bytes per launch from the committed counter summary of the default
    workload (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size)."""
    if (W, K) != (256, 64):
        return {}
    return _profile_line("room200k_W256_K64_hbm_counters.md", "traffic_bytes_per_launch")


def measured_valu(W, K):
    """Wavefront VALU instructions per launch (SQ_INSTS_VALU) from the
    committed counter summary of the default workload."""
    if (W, K) != (256, 64):
        return {}
    return _profile_line("room200k_W256_K64_valu_counters.md", "valu_insts_per_launch")


# MI355X: 256 CUs x 4 SIMDs x 2.4 GHz, a wavefront VALU instruction every 4
# cycles per SIMD (64 lanes over 16-wide SIMDs)
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4
MFMA_BF16_PEAK_TFLOPS = 2500.0


def stop_rule_leg(s, sc, chunk, cap):
    """The reference's own stop rule (Main.cpp:1137, 1288-1300): OnIdle
    cycles of `chunk` passes with the stop test after every pass (on the
    device: lastEnergy.f_Length() < 0.1, the last shooter's radiosity before
    its emitter update), from the initial state."""
    import numpy as np
    s.upload(sc)
    passes = shooters = 0
    st = None
    t0 = time.perf_counter()
    while passes < cap:
        st = s.run_cycle(chunk, 0.1)
        passes += st.passes
        shooters += st.shooters
        if st.stopped:
            break
    ms = (time.perf_counter() - t0) * 1e3
    le = np.array(st.last_energy, np.float32)
    out = {"ms": ms, "passes": passes, "shooters": shooters,
           "last_energy": float(np.sqrt(np.float32((le * le).sum()))),
           "residual": s.residual(), "stopped": bool(st.stopped), "cycle_passes": chunk}
    log(f"reference stop rule (lastEnergy < 0.1): {'after' if out['stopped'] else 'not within'} "
        f"{passes} passes, {ms:.0f} ms, residual {out['residual']:.3e}")
    return out


def gather_leg(R, s, sc, args):
    """Dense matrix from the K=64 hemicube passes (shooting layout, bf16 hi +
    lo) and the BF16-MFMA Jacobi gather with FP64-residual refinement to the
    same residual target.  The context `s` has just converged by batched
    shooting; its radiosity (shot + unshot) is compared with the gather's."""
    import numpy as np
    u, shot = s.state()
    shoot_B = (u + shot).T.astype(np.float64)
    t0 = time.perf_counter()
    s.dense_build(R.RAD_DENSE_SHOOT)
    t1 = time.perf_counter()
    B, st = s.gather(target=args.target, inner=args.gather_inner)
    t2 = time.perf_counter()
    # the same solve run to the floor of the FP64 residual: the error of the
    # 1e-3 answer against the converged fixed point of the stored matrix
    B7, st7 = s.gather(target=1e-7, inner=args.gather_inner, max_outer=20)
    s.dense_release()
    # against the exact FP64 solution of the FP32 form factors: every
    # hemicube rendered again and streamed into an FP64 residual on the device
    t3 = time.perf_counter()
    res64, q = s.gather_check_fp64(B)
    chk = {"fp64_residual_rel_l1": res64, "contraction_q": q,
           "error_bound_rel_l1": res64 / (1.0 - q) if q < 1 else None,
           "seconds": time.perf_counter() - t3}
    a = sc.area.astype(np.float64)[None, :]
    diff = float((a * np.abs(B - shoot_B)).sum() / (a * np.abs(shoot_B)).sum())
    B7 = B7.astype(np.float64)
    err7 = float((a * np.abs(B - B7)).sum() / (a * np.abs(B7)).sum())
    errmax = float(np.max(np.abs(B - B7)) / np.max(np.abs(B7)))
    npad = (sc.npatch + 511) // 512 * 512
    sweep_gbs = 2.0 * npad * npad / (st.sweep_ms * 1e-3) / 1e9 if st.sweep_ms else None
    resid_gbs = 4.0 * npad * npad / (st.resid_ms * 1e-3) / 1e9 if st.resid_ms else None
    out = {"ms": (t2 - t0) * 1e3, "build_ms": (t1 - t0) * 1e3, "solve_ms": (t2 - t1) * 1e3,
           "outer": st.outer, "sweeps": st.sweeps, "residual": st.residual,
           "history": [st.history[k] for k in range(min(st.outer, 16))],
           "sweep_kernel_ms": st.sweep_ms, "sweep_hbm_gbs": sweep_gbs,
           "resid_kernel_ms": st.resid_ms, "resid_hbm_gbs": resid_gbs,
           "matrix_bytes": 4 * npad * npad,
           "rel_l1_vs_batched_shooting": diff,
           "rel_l1_vs_converged": err7, "max_rel_vs_converged": errmax,
           "converged_residual": st7.residual, "fp64_check": chk}
    log(f"gather bf16 to {args.target}: {out['ms'] / 1e3:.2f} s (build {out['build_ms'] / 1e3:.2f}"
        f" s, solve {out['solve_ms']:.0f} ms, {st.outer} steps, {st.sweeps} sweeps), residual "
        f"{st.residual:.2e}; sweep {st.sweep_ms:.2f} ms ({sweep_gbs:.0f} GB/s), residual product "
        f"{st.resid_ms:.2f} ms ({resid_gbs:.0f} GB/s); vs shooting {diff:.2e}; vs the solve "
        f"converged to {st7.residual:.1e}: {err7:.2e} rel L1, {errmax:.2e} max; against the "
        f"FP64 problem of the FP32 rows: residual {res64:.2e}, q {q:.3f}, error bound "
        f"{chk['error_bound_rel_l1']}")
    return out


def cornell_leg(R, O, area, W, shots, seed):
    """Plain single-shooter drop-in on the GPU against the oracle's CPU pass
    (the reference's CPU path restated), same scene, same shots, same rule."""
    import numpy as np
    sc = R.Scene.cornell(area, triangle_patches=True, physical=True, seed=seed)
    s = R.Solver(hemicube=W, hemicubes=1)
    s.upload(sc)
    s.step(3)
    s.finish()
    s.upload(sc)
    t0 = time.perf_counter()
    s.step(shots)
    s.finish()
    gpu = time.perf_counter() - t0
    ug, sg = s.state()
    dff = O.dff_cross(W)
    u = sc.emission.copy()
    sh = sc.shot0.copy()
    item = np.zeros(3 * W * W, np.uint32)
    ff = np.zeros(sc.npatch, np.float32)
    t0 = time.perf_counter()
    for _ in range(shots):
        O.plain_pass(W, 1, 1, sc, dff, u, sh, item, ff)
    cpu = time.perf_counter() - t0
    threads = host_threads()
    u2, sh2 = sc.emission.copy(), sc.shot0.copy()
    t0 = time.perf_counter()
    for _ in range(shots):
        O.plain_pass_mt(W, 1, 1, sc, dff, u2, sh2, threads, item, ff)
    cpu_mt = time.perf_counter() - t0
    a = sc.area.astype(np.float64)[:, None]
    tg, tc = (ug + sg).astype(np.float64), (u + sh).astype(np.float64)
    diff = float((a * np.abs(tg - tc)).sum() / (a * np.abs(tc)).sum())
    out = {"patches": sc.npatch, "hemicube": W, "shots": shots,
           "gpu_ms": gpu * 1e3, "cpu_ms": cpu * 1e3, "cpu_cores": 1,
           "cpu_multicore_ms": cpu_mt * 1e3, "cpu_multicore_cores": threads,
           "speedup_vs_multicore": cpu_mt / gpu,
           "gpu_shooters_per_s": shots / gpu, "cpu_shooters_per_s": shots / cpu,
           "speedup": cpu / gpu, "rel_l1_gpu_vs_cpu_state": diff,
           "residual_gpu": s.residual(), "residual_cpu": O.residual(u.astype(np.float64),
                                                                  sc.emission, sc.area)}
    log(f"cornell {sc.npatch} W={W}: {shots} shots GPU {gpu * 1e3:.1f} ms, CPU "
        f"{cpu * 1e3:.0f} ms ({cpu / gpu:.0f}x), CPU {threads} threads {cpu_mt * 1e3:.0f} ms "
        f"({cpu_mt / gpu:.0f}x); state rel L1 {diff:.1e}")
    return out


def fp64_gather_check(s, sc, E, B, K):
    """Bound on the error of a gather solution B (R x n) against the exact
    solution of B = E + diag(rho) G B, G_ji = F_ij A_i/A_j, built in FP64 from
    the FP32 form-factor rows the hemicube kernels render (before any BF16
    rounding).  One FP64 product on the host gives the residual
    r = E + rho G B - B.  In the area-weighted norm |x|_A = sum_j A_j |x_j|
    the map x -> rho G x contracts by q = max_c max_i sum_j rho_jc F_ij (< 1:
    a hemicube row sums to about 1), so |B - B*|_A <= |r|_A / (1 - q)."""
    import numpy as np
    n = sc.npatch
    F = np.empty((n, n), np.float32)
    for b in range(0, n, K):
        F[b:b + K] = s.ff_rows(np.arange(b, min(n, b + K), dtype=np.uint32))
    G = F.T.astype(np.float64)                                  # G[j, i] = F[i, j]
    del F
    A = sc.area.astype(np.float64)
    G *= A[None, :]
    G /= A[:, None]
    R_ = E.shape[0]
    rho = sc.refl.astype(np.float64)[:, np.arange(R_) % 3]       # n x R
    Bt = B.T.astype(np.float64)
    r = E.T.astype(np.float64) + rho * (G @ Bt) - Bt
    # sum_j rho_jc F_ij = (sum_j G_ji rho_jc A_j) / A_i
    q = float(np.max((G.T @ (sc.refl.astype(np.float64) * A[:, None])) / A[:, None]))
    del G
    rl1 = float((A[:, None] * np.abs(r)).sum() / (A[:, None] * np.abs(Bt)).sum())
    out = {"fp64_residual_rel_l1": rl1, "contraction_q": q,
           "error_bound_rel_l1": rl1 / (1.0 - q) if q < 1 else None}
    return out


def dense_leg(R, args):
    """Dense 32k x 32k matrix (hemicube 128) and the gather of 64 lighting
    conditions x RGB (192 columns) over 32 candidate lights, BF16 MFMA with
    FP64-residual refinement."""
    import numpy as np
    sc = R.Scene.boxes(nboxes=60, room=(12.0, 4.0, 12.0), target_tris=32768, nlights=16, seed=4)
    n = sc.npatch
    s = R.Solver(hemicube=128, hemicubes=64)
    s.upload(sc)
    t0 = time.perf_counter()
    s.dense_build(R.RAD_DENSE_SHOOT)
    t1 = time.perf_counter()
    rng = np.random.default_rng(4)
    cand = rng.choice(n, 32, replace=False)
    E = np.zeros((192, n), np.float32)
    for c in range(64):
        E[3 * c:3 * c + 3, cand] = rng.uniform(0.0, 20.0, (3, 32)) * (rng.random(32) < 0.5)
    s.gather(E, target=1e-2, max_outer=1, inner=1)            # warm the kernels
    t2 = time.perf_counter()
    B, st = s.gather(E, target=1e-4, inner=args.gather_inner)
    t3 = time.perf_counter()
    npad = (n + 511) // 512 * 512
    flop = 2.0 * npad * npad * 192
    out = {"patches": n, "hemicube": 128, "columns": 192, "build_ms": (t1 - t0) * 1e3,
           "solve_ms": (t3 - t2) * 1e3, "outer": st.outer, "sweeps": st.sweeps,
           "residual": st.residual, "sweep_kernel_ms": st.sweep_ms,
           "sweep_hbm_gbs": 2.0 * npad * npad / (st.sweep_ms * 1e-3) / 1e9,
           "sweep_mfma_tflops": flop / (st.sweep_ms * 1e-3) / 1e12,
           "sweep_mfma_frac_of_2500": flop / (st.sweep_ms * 1e-3) / 2.5e15,
           "resid_kernel_ms": st.resid_ms,
           "resid_mfma_tflops": 3 * flop / (st.resid_ms * 1e-3) / 1e12}
    s.dense_release()
    log(f"dense 32k x 192: build {out['build_ms']:.0f} ms, solve {out['solve_ms']:.1f} ms "
        f"({st.outer} steps, {st.sweeps} sweeps, residual {st.residual:.1e}); sweep "
        f"{st.sweep_ms:.3f} ms = {out['sweep_hbm_gbs']:.0f} GB/s, "
        f"{out['sweep_mfma_tflops']:.0f} TFLOP/s")
    # error bound against the exact FP64 solution for the FP32 form factors
    t4 = time.perf_counter()
    chk = fp64_gather_check(s, sc, E, B, 64)
    chk["seconds"] = time.perf_counter() - t4
    out["fp64_check"] = chk
    log(f"dense 32k x 192 against the FP64 problem of the FP32 form factors: residual "
        f"{chk['fp64_residual_rel_l1']:.2e} rel L1 (area-weighted), error bound "
        f"{chk['error_bound_rel_l1']} (contraction {chk['contraction_q']:.3f})")
    return out


def ingest_leg(R):
    """The step before the path at 200k patches: the Cornell faces split into
    about 200,000 triangle patches, built on the host and uploaded
    (rad_scene_faces + rad_scene_upload + rad_set_state) against built on the
    device (rad_scene_ingest); both give the same arrays."""
    import numpy as np
    q, c, e, sh = R.cornell_faces()
    area = 0.00226
    s = R.Solver(hemicube=256, hemicubes=64)
    s.ingest(q, area, R.RAD_FACES_TRIANGLES, c, e, sh)            # warm
    t0 = time.perf_counter()
    host = R.Scene.faces(q, area, R.RAD_FACES_TRIANGLES, c, e, sh)
    t1 = time.perf_counter()
    s.upload(host)
    t2 = time.perf_counter()
    nt, npch, nq = s.ingest(q, area, R.RAD_FACES_TRIANGLES, c, e, sh)
    t3 = time.perf_counter()
    same = bool(np.array_equal(s.download().tri, host.tri))
    out = {"faces": int(q.shape[0]), "patches": npch, "triangles": nt, "quads": nq,
           "host_build_ms": (t1 - t0) * 1e3, "host_upload_ms": (t2 - t1) * 1e3,
           "device_ingest_ms": (t3 - t2) * 1e3, "same_triangles": same,
           "speedup": (t2 - t0) / (t3 - t2)}
    log(f"ingest {npch} patches: host build {out['host_build_ms']:.1f} ms + upload "
        f"{out['host_upload_ms']:.1f} ms vs device {out['device_ingest_ms']:.1f} ms "
        f"({out['speedup']:.1f}x), same triangles {same}")
    return out


def profile_leg(R, name):
    """One path run briefly, for rocprofv3 (kernel trace or counters):
      dense32k    the 32,760-patch dense build and one refinement step of
                  the 192-column gather (4 BF16 sweeps + one two-half
                  residual product), three times
      reduce200k  64 item buffers of the 200k room (hemicube 256) reduced by
                  the standalone ProcessHemicube kernel, five times"""
    import numpy as np
    if name == "dense32k":
        sc = R.Scene.boxes(nboxes=60, room=(12.0, 4.0, 12.0), target_tris=32768, nlights=16,
                           seed=4)
        s = R.Solver(hemicube=128, hemicubes=64)
        s.upload(sc)
        s.dense_build(R.RAD_DENSE_SHOOT)
        rng = np.random.default_rng(4)
        cand = rng.choice(sc.npatch, 32, replace=False)
        E = np.zeros((192, sc.npatch), np.float32)
        for c in range(64):
            E[3 * c:3 * c + 3, cand] = rng.uniform(0.0, 20.0, (3, 32)) * (rng.random(32) < 0.5)
        for _ in range(3):
            _, st = s.gather(E, target=1e-30, max_outer=1, inner=4)
        s.dense_release()
        log(f"profile dense32k: sweep {st.sweep_ms:.3f} ms, residual product {st.resid_ms:.3f} ms")
    elif name == "reduce200k":
        sc = R.Scene.boxes()
        s = R.Solver(hemicube=256, hemicubes=64, tile_px=32, small_tri_px=64)
        s.upload(sc)
        shooters = np.random.default_rng(12).choice(sc.npatch, 64, replace=False)
        s.render(shooters)
        for _ in range(5):
            s.process(64)
        log("profile reduce200k: 64 item buffers reduced 5 times")
    else:
        raise SystemExit(f"unknown profile leg {name}")


def _pack_scenes(R, args):
    """256 independent box rooms, 500-20,000 triangle patches (log-uniform,
    seed 5), hemicube 64 / 128 / 256 by size, shooters wanted per pass
    ceil(N / 2000) (1..16)."""
    import math
    import numpy as np
    rng = np.random.default_rng(5)
    sizes = np.exp(rng.uniform(math.log(500), math.log(20000), args.pack_scenes)).astype(int)
    scenes, Ws, ks = [], [], []
    for i, n in enumerate(sizes):
        side = 5.0 + 7.0 * n / 20000.0
        sc = R.Scene.boxes(nboxes=max(2, int(n) // 400), room=(side, 3.0 + n / 10000.0, side),
                           target_tris=int(n), nlights=4, seed=1000 + i)
        scenes.append(sc)
        Ws.append(64 if sc.npatch < 1500 else (128 if sc.npatch < 6000 else 256))
        ks.append(int(min(16, max(1, math.ceil(sc.npatch / 2000)))))
    return scenes, Ws, ks


def _run_packs(packs, nscenes, target, cap_s):
    import numpy as np
    for s, _ in packs:
        s.pack_begin(target)
    t0 = time.perf_counter()
    done_t = np.full(nscenes, np.nan)
    shooters = np.zeros(nscenes, np.int64)
    live = list(packs)
    npass = 0
    while live and time.perf_counter() - t0 < cap_s:
        for s, _ in live:
            s.step(16)
        now_live = []
        for s, idx in live:
            done, res, passes, rem = s.pack_status()
            npass = max(npass, passes)
            now = time.perf_counter() - t0
            for k, i in enumerate(idx):
                if done[k] != 0xFFFFFFFF and np.isnan(done_t[i]):
                    done_t[i] = now
                shooters[i] = s.pack_shooters[k]
            if rem:
                now_live.append((s, idx))
        live = now_live
    wall = time.perf_counter() - t0
    finished = ~np.isnan(done_t)
    return {"wall_s": wall, "finished": int(finished.sum()), "passes": int(npass),
            "shooters": int(shooters.sum()), "shooters_per_s": float(shooters.sum() / wall),
            "completion_s_p50": float(np.percentile(done_t[finished], 50)) if finished.any() else None,
            "completion_s_p99": float(np.percentile(done_t[finished], 99)) if finished.any() else None}


def pack_leg(R, args):
    """The 256 scenes solved to the residual target in ONE pack: every
    kernel of a pass is one launch for all scenes, each scene rendering at
    its own hemicube side (rad_pack_hemicubes), K = 256 slots per pass shared
    out by water-filling over the scenes' quotas.  For comparison the earlier
    arrangement is timed too (--pack-multi): one pack per hemicube class and
    256-slot budget, side by side on their own streams."""
    scenes, Ws, ks = _pack_scenes(R, args)
    s = R.Solver(hemicube=256, hemicubes=256, graph_passes=8)
    s.pack_upload(scenes, ks)
    s.pack_hemicubes(Ws)
    out = {"scenes": len(scenes), "patches_total": int(sum(sc.npatch for sc in scenes)),
           "packs": 1, "launch": "one launch per kernel per pass for all scenes",
           "slots_per_pass": 256, "quota_total": int(sum(ks)),
           "hemicube_classes": {str(W): Ws.count(W) for W in (64, 128, 256)}}
    out.update(_run_packs([(s, list(range(len(scenes))))], len(scenes), args.target,
                          args.pack_cap_s))
    del s
    log(f"pack: {out['scenes']} scenes ({out['patches_total']} patches) in one pack: "
        f"{out['finished']} done in {out['wall_s']:.2f} s ({out['passes']} passes), "
        f"{out['shooters_per_s']:.0f} shooters/s, completion p50 {out['completion_s_p50']:.2f} s "
        f"p99 {out['completion_s_p99']:.2f} s")
    if args.pack_multi:
        packs = []   # (Solver, scene indices)
        for W in (256, 128, 64):
            idx = [i for i in range(len(scenes)) if Ws[i] == W]
            cur, tot = [], 0
            for i in idx + [None]:
                if i is None or tot + ks[i] > 256:
                    if cur:
                        q = R.Solver(hemicube=W, hemicubes=256, graph_passes=8)
                        q.pack_upload([scenes[j] for j in cur], [ks[j] for j in cur])
                        packs.append((q, cur))
                    cur, tot = [], 0
                if i is not None:
                    cur.append(i)
                    tot += ks[i]
        m = _run_packs(packs, len(scenes), args.target, args.pack_cap_s)
        m["packs"] = len(packs)
        out["multi_pack"] = m
        log(f"pack, {len(packs)} packs by class: {m['finished']} done in {m['wall_s']:.2f} s, "
            f"{m['shooters_per_s']:.0f} shooters/s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--K", type=int, default=64)
    ap.add_argument("--W", type=int, default=256)
    ap.add_argument("--plain-passes", type=int, default=2000)
    ap.add_argument("--target", type=float, default=1e-3)
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--plain-cap-s", type=float, default=90.0)
    ap.add_argument("--cpu-sample-s", type=float, default=10.0)
    ap.add_argument("--small-tri-px", type=int, default=0)
    ap.add_argument("--no-plain", action="store_true", help="skip the K=1 leg (profiling)")
    ap.add_argument("--no-gather", action="store_true", help="skip the dense BF16 gather leg")
    ap.add_argument("--no-reduce", action="store_true",
                    help="skip the standalone ProcessHemicube timing")
    ap.add_argument("--gather-inner", type=int, default=4)
    ap.add_argument("--graph-passes", type=int, default=10,
                    help="passes per captured HIP graph in the batched leg (0: plain launches); "
                         "the default divides the default step count, so every timed pass is "
                         "replayed")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--pack-scenes", type=int, default=256)
    ap.add_argument("--pack-cap-s", type=float, default=120.0)
    ap.add_argument("--pack-multi", action="store_true",
                    help="also time the scenes as one pack per hemicube class")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the Cornell (GPU vs CPU) and dense 32k x 192 legs")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB, so that a "
                         "counter run can scale FETCH_SIZE / WRITE_SIZE to bytes")
    ap.add_argument("--profile", default="",
                    help="run one path briefly for rocprofv3: dense32k or reduce200k")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); sharding shooters "
              "across GPUs is not built", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import radiosity_amd as R
    if args.profile:
        profile_leg(R, args.profile)
        return

    if args.copy_calibration_mib:
        import torch
        n = args.copy_calibration_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {n} bytes read, {n} bytes written")
        del src, dst

    log("scene: room 20x5x20, 400 boxes, seed 3")
    sc = R.Scene.boxes(nboxes=400, room=(20.0, 5.0, 20.0), target_tris=200000, nlights=16, seed=3)
    W, K = args.W, args.K
    names = ["select", "prepare_clear", "raster_setup", "bin", "tile_raster_reduce", "shoot"]

    # ---------------------------------------------------------- batched K
    s = R.Solver(hemicube=W, hemicubes=K, small_tri_px=args.small_tri_px,
                 graph_passes=args.graph_passes)
    s.upload(sc)
    tune = None
    if not args.no_autotune:
        t, sm, tms, cands = s.autotune(3)
        tune = {"tile_px": t, "small_tri_px": sm, "render_ms": tms,
                "candidates": {f"{a}x{b}": round(v, 5) for (a, b), v in sorted(cands.items())}}
        log(f"autotune K={K}: tile {t}, in-lane limit {sm} px ({tms:.3f} ms per render pass)")
    s.step(args.warmup)
    s.prepare()             # capture the pass graph before the clock starts
    s.finish()
    c0 = s.shot_count()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.finish()
    dt = time.perf_counter() - t0
    c1 = s.shot_count()
    shot_timed = c1[1] - c0[1]                 # shooters the device shot
    batched_sps = shot_timed / dt
    ms_per_step = dt * 1e3 / args.steps
    log(f"batched K={K}: {ms_per_step:.3f} ms/pass, {batched_sps:.0f} shooters/s "
        f"({shot_timed} shooters in {c1[0] - c0[0]} passes, counted on the device)")
    # per-kernel device times over the same number of passes (HIP events on
    # the library stream)
    s.kernel_times(1)
    s.step(args.steps)
    s.finish()
    kt, npass = s.kernel_times(0)
    kt = dict(zip(names, kt))
    log("kernel ms/pass: " + json.dumps({k: round(v, 4) for k, v in kt.items()}))

    # --------------------------------------- standalone ProcessHemicube
    # the reference-shaped reduction (rad_process_hemicube) over 64 item
    # buffers of this scene, timed with HIP events: per hemicube 4 B of ids
    # per texel + 4 B F and 4 B hit count per patch, the ΔFF table once
    reduce_leg = None
    if not args.no_reduce:
        import numpy as np
        r = R.Solver(hemicube=W, hemicubes=K, tile_px=32, small_tri_px=64)
        r.upload(sc)
        r.render(np.random.default_rng(12).choice(sc.npatch, K, replace=False))
        r.process(K)
        r.kernel_times(1)
        ms = []
        for _ in range(5):
            r.process(K)
            ms.append(r.reduce_kernel_ms())
        r.kernel_times(0)
        rms = float(np.mean(ms))
        rb = K * (4.0 * 3 * W * W + 8.0 * sc.npatch) + 4.0 * 3 * W * W
        reduce_leg = {"bound": "hbm", "avg_launch_ms": round(rms, 5), "alg_bytes_per_launch": rb,
                      "achieved_gbs": round(rb / (rms * 1e-3) / 1e9, 1),
                      "frac": round(rb / (rms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"standalone ProcessHemicube, {K} item buffers: {rms * 1e3:.1f} us, "
            f"{reduce_leg['achieved_gbs']:.0f} GB/s")
        del r

    # ----------------------------------------------------------- plain K=1
    # the single-shooter drop-in, launched eagerly and replayed from captured
    # graphs of 8 passes; the faster of the two is the plain baseline
    p = plain_sps = pdt = pkt = None
    plain_modes = {}
    if not args.no_plain:
        for gp in (0, 8):
            q = R.Solver(hemicube=W, hemicubes=1, small_tri_px=args.small_tri_px, graph_passes=gp)
            q.upload(sc)
            q.step(24)
            q.finish()
            c0 = q.shot_count()
            t0 = time.perf_counter()
            q.step(args.plain_passes)
            q.finish()
            qdt = time.perf_counter() - t0
            nshot = q.shot_count()[1] - c0[1]
            plain_modes["graph8" if gp else "eager"] = {
                "shooters_per_s": round(nshot / qdt, 2),
                "ms_per_pass": round(qdt * 1e3 / args.plain_passes, 5), "graph_passes": gp}
            log(f"plain K=1 {'graph' if gp else 'eager'}: {qdt * 1e3 / args.plain_passes:.4f} "
                f"ms/pass, {nshot / qdt:.0f} shooters/s")
            if plain_sps is None or nshot / qdt > plain_sps:
                p, plain_sps, pdt = q, nshot / qdt, qdt
        p.kernel_times(1)
        p.step(200)
        p.finish()
        pkt, _ = p.kernel_times(0)
        pkt = dict(zip(names, pkt))

    # --------------------------------------------------------- convergence
    conv = None
    if not args.no_converge:
        s.upload(sc)
        st = s.run(10 ** 9, args.target, check_every=100)
        conv = {"batched": {"ms": st.seconds * 1e3, "passes": st.passes, "shooters": st.shooters,
                            "residual": st.residual}}
        log(f"batched to {args.target}: {st.seconds:.2f} s, {st.shooters} shooters, "
            f"residual {st.residual:.3e}")
        if not args.no_gather:
            conv["gather_bf16"] = gather_leg(R, s, sc, args)
        conv["batched_reference_stop"] = stop_rule_leg(s, sc, 1, 20000)
        if p is not None:
            p.upload(sc)
            cap = int(max(1000, plain_sps * args.plain_cap_s))
            st = p.run(cap, args.target, check_every=2000)
            reached = st.residual <= args.target
            conv["plain"] = {"ms": st.seconds * 1e3, "passes": st.passes, "shooters": st.shooters,
                             "residual": st.residual, "reached": bool(reached),
                             "cap_passes": cap}
            log(f"plain to {args.target}: {st.seconds:.2f} s, {st.shooters} shooters, residual "
                f"{st.residual:.3e} ({'reached' if reached else 'capped'})")
            if reached:
                conv["speedup_batched_vs_plain"] = conv["plain"]["ms"] / conv["batched"]["ms"]
                if "gather_bf16" in conv:
                    conv["speedup_gather_vs_plain"] = (conv["plain"]["ms"]
                                                       / conv["gather_bf16"]["ms"])

    # -------------------------------------- Cornell configs and dense 32k
    configs = None
    if not args.no_configs:
        from oracle import oracle as O
        configs = {"cornell1500_W128_100": cornell_leg(R, O, 0.3127, 128, 100, 1),
                   "cornell20k_W256_1000": cornell_leg(R, O, 0.02314, 256, 1000, 1),
                   "dense32k_R192": dense_leg(R, args),
                   "ingest200k": ingest_leg(R),
                   "pack256": pack_leg(R, args)}

    # ------------------------------------------------------------ roofline
    # Dominant kernel of the batched pass and its algorithmic bytes per launch
    # (K shooters):
    #  tile_raster_reduce  ProcessHemicube: a 4 B item-buffer id per texel and
    #                      shooter (resolved in LDS here, never written), a
    #                      4 B F-row entry per patch and shooter, and the 4 B
    #                      per texel ΔFF table once per launch
    #  raster_setup        36 B of vertices + 4 B patch id per triangle and
    #                      shooter
    #  shoot               4 B of F per shooter and patch + 40 B of patch state
    #  bin, select, prepare_clear: list entries, 16 B / 4 B per patch
    RES = 3 * W * W
    per_launch = {
        "raster_setup": K * 40.0 * sc.ntri,
        "tile_raster_reduce": K * (4.0 * RES + 4.0 * sc.npatch) + 4.0 * RES,
        "shoot": sc.npatch * (4.0 * K + 40.0),
        "select": sc.npatch * 16.0,
        "bin": K * 3.0 * RES,
        "prepare_clear": K * 4.0 * sc.npatch,
    }
    dom = max(per_launch, key=lambda k: kt[k])
    alg = per_launch[dom]
    achieved = alg / (kt[dom] * 1e-3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "alg_bytes_per_launch": alg, "avg_launch_ms": round(kt[dom], 5),
            "traffic": measured_traffic(W, K).get(dom)}
    # the rasterizing kernels are bound by instruction issue, not HBM: their
    # wavefront VALU instructions per launch (counter run of this workload)
    # over the launch time, against 614 G/s
    valu = measured_valu(W, K)
    roof["issue"] = {k: {"valu_insts_per_launch": valu[k], "avg_launch_ms": round(kt[k], 5),
                         "achieved_ginst_s": round(valu[k] / (kt[k] * 1e-3) / 1e9, 1),
                         "peak_ginst_s": VALU_PEAK_GINST,
                         "frac": round(valu[k] / (kt[k] * 1e-3) / 1e9 / VALU_PEAK_GINST, 4)}
                     for k in ("tile_raster_reduce", "raster_setup") if k in valu}
    sh_gbs = per_launch["shoot"] / (kt["shoot"] * 1e-3) / 1e9
    roof["others"] = {"shoot": {"bound": "hbm", "achieved_gbs": round(sh_gbs, 1),
                                "frac": round(sh_gbs / HBM_PEAK_GBS, 4),
                                "alg_bytes_per_launch": per_launch["shoot"],
                                "avg_launch_ms": round(kt["shoot"], 5)}}
    if reduce_leg:
        roof["others"]["process_hemicube"] = reduce_leg
    if configs and "dense32k_R192" in configs:
        d = configs["dense32k_R192"]
        roof["others"]["gather_gemm_32k_R192"] = {
            "bound": "mfma", "achieved_tflops": round(d["sweep_mfma_tflops"], 1),
            "peak_tflops": MFMA_BF16_PEAK_TFLOPS,
            "frac": round(d["sweep_mfma_tflops"] / MFMA_BF16_PEAK_TFLOPS, 4),
            "hbm_gbs": round(d["sweep_hbm_gbs"], 1),
            "hbm_frac": round(d["sweep_hbm_gbs"] / HBM_PEAK_GBS, 4),
            "sweep_ms": d["sweep_kernel_ms"]}

    # --------------------------------------------------------- CPU baseline
    cpu = cpu_mt = raster = None
    if args.cpu_sample_s > 0:
        from oracle import oracle as O
        dff = O.dff_cross(W)
        u = sc.emission.copy()
        sh = sc.shot0.copy()
        item = np.zeros(RES, np.uint32)
        ff = np.zeros(sc.npatch, np.float32)
        n = 0
        t0 = time.perf_counter()
        while True:
            O.plain_pass(W, 1, 1, sc, dff, u, sh, item, ff)
            n += 1
            el = time.perf_counter() - t0
            if el >= args.cpu_sample_s and n >= 3:
                break
        model = cpu_model()
        cpu = {"value": n / el, "unit": "shooters/s", "cores": 1, "kind": "port",
               "sample": f"{n} plain passes (K=1, W={W}) of the oracle on the same 200k-patch "
                         f"scene from the initial state, {el:.1f} s",
               "cpu_model": model, "nproc": os.cpu_count()}
        log(f"cpu oracle: {n / el:.2f} shooters/s")
        # the same passes on all the host cores this process may use (OpenMP)
        threads = host_threads()
        u = sc.emission.copy()
        sh = sc.shot0.copy()
        n = 0
        t0 = time.perf_counter()
        while True:
            O.plain_pass_mt(W, 1, 1, sc, dff, u, sh, threads, item, ff)
            n += 1
            el = time.perf_counter() - t0
            if el >= args.cpu_sample_s and n >= 3:
                break
        cpu_mt = {"value": n / el, "unit": "shooters/s", "cores": threads, "kind": "port",
                  "sample": f"{n} plain passes (K=1, W={W}) of the oracle's OpenMP pass on the "
                            f"same scene from the initial state, {el:.1f} s",
                  "cpu_model": model, "nproc": os.cpu_count()}
        log(f"cpu oracle, {threads} threads: {n / el:.2f} shooters/s")
        # covered fragments per hemicube on sample shooters, for fragments/s
        frags = [O.render_hemicube_mt(W, sc.tri, sc.tri_patch, sc.frame[p], threads)[1]
                 for p in np.random.default_rng(11).choice(sc.npatch, 4, replace=False)]
        fph = float(np.mean(frags))
        raster = {"triangle_hemicubes_per_s": sc.ntri * batched_sps,
                  "fragments_per_hemicube": fph,
                  "fragments_per_s": fph * batched_sps,
                  "sample": "fragments counted by the oracle on 4 seeded shooters"}
        log(f"rasterizer: {sc.ntri * batched_sps / 1e9:.1f} G triangle-hemicubes/s, "
            f"{fph * batched_sps / 1e9:.1f} G fragments/s ({fph:.0f} per hemicube)")

    out = {
        "metric": "shooters/sec + ms to 1e-3 unshot energy at 200k patches, plain FP32 vs "
                  "batched/BF16",
        "value": round(batched_sps, 2),
        "unit": "shooters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random-box room)",
        "config": {"workload": f"room200k: 20x5x20 room, 400 boxes (seed 3), {sc.npatch} "
                               f"triangle patches, 16 emitters, hemicube {W}, K={K} shooters "
                               f"per pass", "patches": sc.npatch, "hemicube": W,
                   "hemicubes_per_pass": K, "parallelism": "single GPU"},
        "plain": None if p is None else {
            "shooters_per_s": round(plain_sps, 2),
            "ms_per_pass": round(pdt * 1e3 / args.plain_passes, 5),
            "modes": plain_modes, "kernel_ms": {k: round(v, 5) for k, v in pkt.items()}},
        "batched": {"shooters_per_s": round(batched_sps, 2), "autotune": tune,
                    "graph_passes": args.graph_passes,
                    "kernel_ms": {k: round(v, 5) for k, v in kt.items()}},
        "converge": conv,
        "configs": configs,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_baseline_multicore": cpu_mt,
        "rasterizer": raster,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
