"""Build recipe for libshrinkers_hip.so (the gfx950 HIP library).

Compiles every ``csrc/*.hip`` translation unit with hipcc into one shared
library placed next to this file.  Flags:

* ``--offload-arch=gfx950``: CDNA4 (MI355X) code objects only.
* ``-ffp-contract=off``: no implicit a*b+c fusion; the kernels spell out
  every fused multiply-add they want, so the arithmetic is the one written.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32
  (v_pk_*_f32) instructions are generated.
"""
from __future__ import annotations

import glob
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_NAME = "libshrinkers_hip.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def build(verbose: bool = False, jobs: int = 0) -> str:
    """Compile the library (always recompiles) and return its path.

    Each translation unit is compiled to an object file in a temporary
    directory, ``jobs`` at a time (default: up to 8), and the objects are
    linked into the shared library.  The units share no device symbols (the
    device helpers they have in common are header-inline), so separate
    compilation needs no relocatable device code."""
    import concurrent.futures
    import tempfile
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]
    compile_flags = [f for f in FLAGS if f != "-shared"]
    jobs = jobs or min(8, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory(prefix="shk_build_") as tmpdir:
        def compile_one(src: str) -> str:
            obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
            cmd = [HIPCC, *compile_flags, *inc, "-c", src, "-o", obj]
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
            return obj
        with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
            objs = list(ex.map(compile_one, sources()))
        tmp = LIB_PATH + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose=True))
