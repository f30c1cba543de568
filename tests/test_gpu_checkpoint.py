"""GPU: checkpoint / resume of the solver state (portable successor of the
reference's .rr save and load) through the C ABI and the drop-in layer."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def scene(rad):
    return rad.Scene.cornell(0.5, triangle_patches=True, physical=True, seed=1)


def test_resume_continues_the_solve(rad, scene, tmp_path):
    path = str(tmp_path / "state.rck")
    a = rad.Solver(hemicube=32, hemicubes=4)
    a.upload(scene)
    a.step(10)
    a.checkpoint_save(path)
    ua, sa = a.state()
    b = rad.Solver(hemicube=32, hemicubes=4)
    b.upload(scene)
    b.checkpoint_load(path)
    ub, sb = b.state()
    assert np.array_equal(ua, ub) and np.array_equal(sa, sb)
    assert b.residual() == pytest.approx(a.residual(), rel=1e-12)
    a.step(10)
    b.step(10)
    ua, sa = a.state()
    ub, sb = b.state()
    w = scene.area[:, None].astype(np.float64)
    tot_a, tot_b = (ua + sa).astype(np.float64), (ub + sb).astype(np.float64)
    assert np.abs(w * (tot_a - tot_b)).sum() <= 1e-5 * np.abs(w * tot_a).sum()


def test_checkpoint_rejects_other_scenes_and_corruption(rad, scene, tmp_path):
    path = str(tmp_path / "state.rck")
    a = rad.Solver(hemicube=16, hemicubes=1)
    a.upload(scene)
    a.step(3)
    a.checkpoint_save(path)
    other = rad.Scene.cornell(0.5, triangle_patches=True, physical=True, seed=2)  # other rho
    b = rad.Solver(hemicube=16, hemicubes=1)
    b.upload(other)
    with pytest.raises(rad.RadError, match="another scene"):
        b.checkpoint_load(path)
    raw = bytearray(open(path, "rb").read())
    raw[100] ^= 0x40
    bad = str(tmp_path / "bad.rck")
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(rad.RadError, match="checksum"):
        a.checkpoint_load(bad)
    open(bad, "wb").write(bytes(raw[:50]))
    with pytest.raises(rad.RadError):
        a.checkpoint_load(bad)
    m = rad.Solver(hemicube=16, hemicubes=1, mode=rad.RAD_MODE_REFERENCE)
    m.upload(scene)
    with pytest.raises(rad.RadError, match="mode"):
        m.checkpoint_load(path)


def test_dropin_save_and_load(tmp_path):
    path = str(tmp_path / "dropin.rck")
    p = subprocess.run([os.path.join(ROOT, "radiosity_amd", "bin", "dropin_check"), "resume",
                        "0.5", "16", path], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    d = json.loads(p.stdout.strip().splitlines()[-1])
    assert d["saved"] == 1 and d["loaded"] == 1 and d["same"] == d["n"] > 0


def test_oracle_written_checkpoint_loads(rad, scene, tmp_path):
    """A file written by the independent writer of oracle/scene_io.py from
    the documented layout loads, and the device state equals its planes bit
    for bit."""
    from oracle import scene_io as S
    rng = np.random.default_rng(9)
    n = scene.npatch
    u = rng.uniform(0, 3, (n, 3)).astype(np.float32)
    sh = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    path = str(tmp_path / "oracle.rck")
    S.ckpt_write(path, scene, 1, u, sh, scene.emission)
    a = rad.Solver(hemicube=16, hemicubes=1)
    a.upload(scene)
    a.checkpoint_load(path)
    gu, gs = a.state()
    assert np.array_equal(gu.view(np.uint32), u.view(np.uint32))
    assert np.array_equal(gs.view(np.uint32), sh.view(np.uint32))
    want = (scene.area[:, None] * u.astype(np.float64)).sum() / \
        (scene.area[:, None] * scene.emission.astype(np.float64)).sum()
    assert a.residual() == pytest.approx(want, rel=1e-12)
    # the same file with a wrong mode word is refused
    S.ckpt_write(path, scene, 0, u, sh, scene.emission)
    with pytest.raises(rad.RadError, match="mode"):
        a.checkpoint_load(path)


def test_device_checkpoint_parsed_by_oracle(rad, scene, tmp_path):
    """A file the library writes, read by the independent reader: header
    fields, the scene hash recomputed from the scene arrays, the checksum and
    the three planes."""
    from oracle import scene_io as S
    path = str(tmp_path / "device.rck")
    a = rad.Solver(hemicube=32, hemicubes=4)
    a.upload(scene)
    a.step(6)
    a.checkpoint_save(path)
    gu, gs = a.state()
    d = S.ckpt_read(path)
    assert d["checksum_ok"] and d["version"] == 1 and d["mode"] == 1
    assert d["npatch"] == scene.npatch and d["ntri"] == scene.ntri
    assert d["size"] == 8 + 16 + 8 + 36 * scene.npatch + 8
    assert d["scene_hash"] == S.scene_hash(scene.tri, scene.tri_patch, scene.frame, scene.refl,
                                           scene.area)
    assert np.array_equal(d["unshot"].view(np.uint32), gu.view(np.uint32))
    assert np.array_equal(d["shot"].view(np.uint32), gs.view(np.uint32))
    assert np.array_equal(d["emission"], scene.emission)
