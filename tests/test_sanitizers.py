"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

  - oracle.c with oracle/asan_check.c (every oracle function on a box soup
    with near-plane and guard-band clipping), built by `make -C oracle asan`;
  - the scene builders (rad_scene.cpp), the OBJ reader (rad_obj.cpp) and the
    drop-in layer (rad_dropin.cpp) linked into the drop-in test driver
    (radiosity_amd.build.build_asan), run on its CPU modes.

Leak detection is on; any sanitizer report fails the test."""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24", OMP_NUM_THREADS="3")


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=ENV)
    report = "AddressSanitizer" in p.stderr or "runtime error" in p.stderr or \
        "LeakSanitizer" in p.stderr
    assert p.returncode == 0 and not report, (cmd, p.returncode, p.stderr[-3000:])
    return p.stdout


def _linked_with_asan(exe):
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    return "libasan" in out


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True)
    exe = os.path.join(ROOT, "oracle", "_asan", "oracle_asan")
    assert _linked_with_asan(exe)
    assert '"ok": 1' in _run([exe])


def test_host_library_code_under_sanitizers(rad, tmp_path):
    from radiosity_amd import build as B
    exe = B.build_asan()
    assert _linked_with_asan(exe)
    obj = tmp_path / "t.obj"
    obj.write_text("v 0 0 0\nv 1000 0 0\nv 1000 1000 0\nv 0 1000 0\nv 0 0 1000\n"
                   "f 1 2 3 4\nf 1 2 5\nf 2/1 3/1/1 5//2\nf 1 2 99\nf -1 0 3\n\n"
                   "f 1 2 3 4 5\nv 1 2\n")
    _run([exe, "config", "16", "10", "4", "500", "0.3"])
    _run([exe, "patches", "0.3127"])
    # the file ends with a bad vertex line: three faces kept, parse incomplete
    out = json.loads(_run([exe, "scenes", str(obj)]).strip().splitlines()[-1])
    assert out["obj_faces"] == 3 and out["obj_rc"] != 0 and out["sum"] > 0
    out = json.loads(_run([exe, "scenes", str(tmp_path / "missing.obj")]).strip().splitlines()[-1])
    assert out["obj_faces"] == 0 and out["obj_rc"] != 0
    _run([exe, "obj", str(obj), "0.1"])
