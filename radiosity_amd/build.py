"""Builds libradhip.so in-tree with hipcc for gfx950.

    python -m radiosity_amd.build        # or radiosity_amd.build.build()

Every HIP translation unit gets -ffp-contract=off (the item-buffer arithmetic
is specified operation by operation) and the packed-FP32 feature disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libradhip.so")
BUILD = os.path.join(HERE, "_build")
BIN = os.path.join(HERE, "bin")
ROOT = os.path.dirname(HERE)

# host programs linked against libradhip.so (rpath: the package directory)
PROGRAMS = [
    ("radhip_cli", os.path.join(CSRC, "rad_cli.cpp")),
    ("dropin_check", os.path.join(ROOT, "tests", "dropin_check.cpp")),
]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RADHIP_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall"]
HIPFLAGS = [f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    # (source, compiled as HIP)
    ("rad_kernels.hip", True),
    ("rad_gather.hip", True),
    ("rad_pack.hip", True),
    ("rad_ingest.hip", True),
    ("rad_static.hip", True),
    ("rad_api.cpp", True),
    ("rad_scene.cpp", False),
    ("rad_obj.cpp", False),
    ("rad_dropin.cpp", False),
]


def _compile(src, as_hip):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src + ".o")
    cmd = [HIPCC] + COMMON
    if as_hip:
        cmd += ["-x", "hip"] + HIPFLAGS
    cmd += ["-I", os.path.join(HERE, "..", "include"), "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [(s, h) for s, h in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda a: _compile(*a), srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    os.makedirs(BIN, exist_ok=True)
    for name, src in PROGRAMS:
        if not os.path.exists(src):
            continue
        exe = os.path.join(BIN, name)
        cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), src,
               "-o", exe + ".tmp", "-L", HERE, "-lradhip", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
        os.replace(exe + ".tmp", exe)
    if verbose:
        print(f"built {OUT}")
    return OUT


# Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only,
# g++, no device code): the scene builders, the OBJ reader and the drop-in
# layer linked into the drop-in test driver; the device entry points come from
# libradhip.so.  Built by the sanitizer test, not by build().
ASAN_SOURCES = [os.path.join(CSRC, f) for f in ("rad_scene.cpp", "rad_obj.cpp", "rad_dropin.cpp")]
ASAN_EXE = os.path.join(BUILD, "asan", "dropin_check_asan")


def build_asan(verbose=False):
    os.makedirs(os.path.dirname(ASAN_EXE), exist_ok=True)
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-Wall", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I", os.path.join(ROOT, "include")]
    cmd += ASAN_SOURCES + [os.path.join(ROOT, "tests", "dropin_check.cpp"),
                           "-o", ASAN_EXE, "-L", HERE, "-lradhip", f"-Wl,-rpath,{HERE}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"sanitizer build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {ASAN_EXE}")
    return ASAN_EXE


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
