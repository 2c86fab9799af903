"""Builds libumhip.so (all HIP sources, gfx950) in-tree.

    python -m umagnus_amd.build            # build
    python -m umagnus_amd.build --clean    # remove objects and the library
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libumhip.so")
SOURCES = ["api.hip", "demag.hip", "newell.hip", "kcache.hip", "field.hip", "vec.hip", "sim.hip", "adjoint.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("UMHIP_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 (v_pk_*_f32) stays off: scalar FP32 only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-Wno-unknown-pragmas",
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4, 8)) as ex:
        results = list(ex.map(_compile, SOURCES))
    failed = False
    for src, _obj, rc, out in results:
        msgs = [l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if rc != 0 or (verbose and msgs):
            print(f"--- {src} (rc={rc})")
            print("\n".join(msgs))
        failed |= rc != 0
    if failed:
        raise RuntimeError("libumhip build failed")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + [o for _, o, _, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        print(r.stdout + r.stderr)
        raise RuntimeError("libumhip link failed")
    os.replace(tmp, LIB)
    return LIB


def clean():
    shutil.rmtree(OBJ, ignore_errors=True)
    if os.path.exists(LIB):
        os.remove(LIB)


if __name__ == "__main__":
    if "--clean" in sys.argv:
        clean()
    else:
        print(build(verbose="-v" in sys.argv))
