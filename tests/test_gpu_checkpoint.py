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
