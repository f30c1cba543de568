"""The C++ drop-in layer (include/radhip_dropin.hpp) through its test driver
(tests/dropin_check.cpp) and the headless CLI.  CPU tests cover Config
arithmetic, scene assembly and argument handling; GPU tests compare the
device-backed calls with the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "radiosity_amd", "bin")


def run(*args, ok=True):
    p = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]],
                       capture_output=True, text=True, timeout=300)
    if ok:
        assert p.returncode == 0, p.stderr
        return json.loads(p.stdout.strip().splitlines()[-1])
    return p


@pytest.mark.parametrize("W,K,wx", [(16, 10, 4), (256, 64, 4), (2, 3, 16), (100, 1, 1)])
def test_config_freeze_arithmetic(W, K, wx):
    """Config.cpp:30-48: TEX 2W x floor(1.5W), workitems min(x, TEX_W) x TEX_H*K;
    setters after freeze are refused."""
    c = run("dropin_check", "config", W, K, wx, 500, 0.25)
    tw, th = 2 * W, int(W * 1.5)
    assert (c["W"], c["H"], c["TW"], c["TH"], c["RES"]) == (W, W, tw, th, tw * th)
    assert c["WX"] == min(wx, tw) and c["WY"] == th * K and c["K"] == K
    assert c["frozen_err"] == 1 and c["W"] == W


@pytest.mark.parametrize("area", [0.5, 0.3127])
def test_model_container_matches_scene_source(rad, area):
    """load + updateData: patch order, vertices (4 x 3 per patch), indices
    (0,1,2,0,2,3 per patch), frames, colours and initial fields equal the
    reference-rule scene of rad_scene_cornell."""
    d = run("dropin_check", "patches", area)
    sc = rad.Scene.cornell(area, triangle_patches=False, physical=False)
    n = sc.npatch
    assert d["n"] == n and d["nv"] == 12 * n and d["ni"] == 6 * n
    tri = sc.tri.reshape(n, 2, 3, 3)
    quads = np.concatenate([tri[:, 0], tri[:, 1, 2:3]], axis=1)      # v0, v1, v2, v3
    assert np.array_equal(np.array(d["vertices"], np.float32), quads.reshape(-1))
    ix = np.array(d["indices"]).reshape(n, 6) - 4 * np.arange(n)[:, None]
    assert np.all(ix == [0, 1, 2, 0, 2, 3])
    fr = np.array(d["frames"], np.float32).reshape(4, 9)
    for k, i in enumerate((0, 5, n // 2, n - 1)):
        assert np.array_equal(fr[k], sc.frame[i]), i
    assert np.array_equal(np.array(d["colors"], np.float32).reshape(n, 3), sc.refl)
    assert np.array_equal(np.array(d["radiosity_r"], np.float32), sc.emission[:, 0])
    assert np.array_equal(np.array(d["illumination_r"], np.float32), sc.shot0[:, 0])


def test_cli_rejects_bad_arguments():
    p = run("radhip_cli", "area", ok=False)
    assert p.returncode == 2 and "wrong number of parameters" in p.stderr
    p = run("radhip_cli", "colour", "red", ok=False)
    assert p.returncode == 2 and "unknown parameter" in p.stderr


SMALL_OBJ = (
    "v 0 0 0\nv 2000 0 0\nv 2000 0 1500\nv 0 0 1500\nv 1000 800 700\n"
    "f 1 2 3 4\nf 1/1 2/2 5/5\nf 4 3 5\n"
)


def test_dropin_wavefront_model(rad, tmp_path):
    """WaveFrontModel: the OBJ faces subdivided by Patch::divide, black,
    with the neighbour links of Patch.cpp:127-222, as the face-scene builder
    produces them."""
    path = tmp_path / "small.obj"
    path.write_text(SMALL_OBJ)
    d = run("dropin_check", "obj", str(path), 0.2)
    faces, ok = rad.obj_parse(str(path))
    assert ok and d["error"] == 0 and d["color0"] == [0, 0, 0]
    sc = rad.Scene.faces(faces, 0.2, rad.RAD_FACES_QUADS)
    assert d["n"] == sc.nquad
    tri = sc.tri.reshape(-1, 2, 3, 3)
    quads = np.concatenate([tri[:, 0], tri[:, 1, 2:3]], axis=1)
    assert np.array_equal(np.array(d["vertices"], np.float32), quads.reshape(-1))
    assert np.array_equal(np.array(d["neighbours"]).reshape(-1, 8), sc.quad_nbr)
    d = run("dropin_check", "obj", str(tmp_path / "missing.obj"), 0.2)
    assert d["n"] == 0 and d["error"] == 1


def test_cli_rejects_bad_view():
    p = run("radhip_cli", "view", "1,2,3", ok=False)
    assert p.returncode == 2 and "view takes 9" in p.stderr


# ------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("W,K", [(16, 3), (64, 10)])
def test_dropin_form_factors(W, K):
    d = run("dropin_check", "dff", W, K)
    assert d["copies_equal"] == 1
    assert np.array_equal(np.array(d["table"], np.float32).view(np.uint32),
                          O.dff_cross(W).view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("K,seed", [(1, 1), (10, 2), (37, 3)])
def test_dropin_selection_list_rule(K, seed):
    d = run("dropin_check", "select", 0.3127, K, seed)
    rr = np.array(d["radiosity"], np.float32).reshape(-1, 3)
    want, n = O.select_reference(rr, K)
    assert np.array_equal(np.array(d["ids"], np.uint32), want)
    ok = np.array(d["ptr_ok"])
    assert np.all(ok[:n] == 1) and np.all(ok[n:] == -1)
    l2 = (rr.astype(np.float32) ** 2).sum(1)
    assert d["first"] == int(np.argmax(l2))                        # first index of the max


@pytest.mark.gpu
def test_dropin_shooter_matches_oracle(rad):
    """Two cycles of 15 single-shooter passes in REFERENCE mode (the
    reference's rules) against the oracle's CPU pass, same fields."""
    area, W = 0.5, 32
    d = run("dropin_check", "shoot", area, W, 1, 15, 2)
    sc = rad.Scene.cornell(area, triangle_patches=False, physical=False)
    dff = O.dff_cross(W)
    u = sc.emission.copy()
    s = sc.shot0.copy()
    for _ in range(30):
        O.plain_pass(W, 0, 0, sc, dff, u, s)
    assert d["shooters"] == 30
    got_u = np.array(d["radiosity"], np.float32).reshape(-1, 3)
    got_s = np.array(d["illumination"], np.float32).reshape(-1, 3)
    assert np.allclose(got_u, u, rtol=1e-4, atol=1e-4 * np.abs(u).max())
    assert np.allclose(got_s, s, rtol=1e-4, atol=1e-4 * np.abs(s).max())


@pytest.mark.gpu
def test_dropin_process_hemicube_triplets(rad):
    area, W = 0.3127, 64
    d = run("dropin_check", "process", area, W)
    sc = rad.Scene.cornell(area, triangle_patches=False, physical=False)
    light = d["light"]
    item = np.array(d["item"], np.uint32)
    assert np.array_equal(item, O.render_hemicube(W, sc.tri, sc.tri_patch, sc.frame[light]))
    # the kernel's own stream: per-run triplets in blocks of 4 with zero
    # padding, bit for bit the oracle's (work-item order)
    h, i, e = O.process_triplets(W, 1, item, O.dff_cross(W))
    assert d["count"] == len(h) and d["count"] % 4 == 0
    assert np.array_equal(np.array(d["hemicubes"], np.uint32), h)
    assert np.array_equal(np.array(d["ids"], np.uint32), i)
    assert np.array_equal(np.array(d["energies"], np.float32).view(np.uint32), e.view(np.uint32))
    # the host accumulation of Main.cpp:1257-1266 over it gives the F row
    hits, ff = O.histogram(W, 1, item, O.dff_cross(W), sc.npatch)
    acc = np.zeros(sc.npatch, np.float64)
    np.add.at(acc, i, e.astype(np.float64))
    assert np.allclose(acc, ff[0], rtol=1e-5, atol=1e-9)


@pytest.mark.gpu
def test_cli_runs_to_the_reference_stop_rule():
    p = subprocess.run([os.path.join(BIN, "radhip_cli"), "area", "0.5", "hemicube", "32",
                        "shoots", "20", "hemicubes", "4", "cycles", "200", "json", "1"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    d = json.loads(p.stdout.strip().splitlines()[-1])
    assert d["last_energy"] < 0.1 and d["cycles"] < 200
    assert d["shooters"] > 0 and 0 <= d["residual"] < 1


@pytest.mark.gpu
def test_dropin_smooth_shade_device_equals_host():
    """ModelContainer::smoothShadeAll (device) equals Colors::smoothShadePatch
    over the patch fields (host) for every patch, bit for bit."""
    d = run("dropin_check", "smooth", 0.3127, 64, 3)
    assert d["ok"] == 1 and d["values"] == 12 * d["n"]
    assert d["same"] == d["values"] and d["max"] > 0


@pytest.mark.gpu
def test_cli_exports_ply_and_ppm(tmp_path):
    ply, ppm = tmp_path / "out.ply", tmp_path / "out.ppm"
    p = subprocess.run([os.path.join(BIN, "radhip_cli"), "area", "0.5", "hemicube", "64",
                        "shoots", "20", "hemicubes", "4", "cycles", "30", "json", "1",
                        "ply", str(ply), "ppm", str(ppm)],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    n = json.loads(p.stdout.strip().splitlines()[-1])["patches"]
    data = ply.read_bytes()
    head, body = data.split(b"end_header\n", 1)
    assert f"element vertex {4 * n}".encode() in head and f"element face {n}".encode() in head
    assert len(body) == 4 * n * 15 + n * 17
    img = ppm.read_bytes()
    assert img.startswith(b"P6\n64 64\n255\n")
    pix = np.frombuffer(img[len(b"P6\n64 64\n255\n"):], np.uint8)
    assert pix.size == 64 * 64 * 3 and pix.max() > 0
    # the open side is closed by ROOMCLOSURE: inside the box no pixel is background
    assert np.count_nonzero(pix.reshape(-1, 3).max(1)) > 0.9 * 64 * 64


@pytest.mark.gpu
def test_cli_loads_obj(tmp_path):
    path = tmp_path / "small.obj"
    path.write_text(SMALL_OBJ)
    p = subprocess.run([os.path.join(BIN, "radhip_cli"), "obj", str(path), "area", "0.2",
                        "hemicube", "32", "json", "1"],
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    d = json.loads(p.stdout.strip().splitlines()[-1])
    assert d["patches"] > 3 and d["sum_rgb"] == [0, 0, 0]   # black, nothing emits

