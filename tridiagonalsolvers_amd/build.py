"""Build recipe for libtrid.so (gfx950 only) -- invoked by ``__graft_entry__.build()``.

Compiles every HIP/C++ source under ``csrc/`` with hipcc for
``--offload-arch=gfx950`` into ``tridiagonalsolvers_amd/libtrid.so`` (in
tree, so the library travels with the repository snapshot). Objects are
cached under ``tridiagonalsolvers_amd/_build/`` and rebuilt when a source or
header is newer. Packed FP32 instructions are disabled
(``-target-feature -packed-fp32-ops``) on every device compile. The
``lds_poison`` variant (``libtrid_lds_poison.so``, objects under
``_build_lds_poison/``) is the same library compiled with
``-DTRID_LDS_POISON``: every LDS staging and exchange area is filled with
NaN before each use, so a read of an element no thread wrote shows up as
NaN in the parity tests.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtrid.so")
# debug variant: every LDS staging / exchange area filled with NaN before use
LIB_LDS_POISON = os.path.join(HERE, "libtrid_lds_poison.so")
VARIANTS = {
    "release": (LIB, BUILD, []),
    "lds_poison": (LIB_LDS_POISON, os.path.join(HERE, "_build_lds_poison"), ["-DTRID_LDS_POISON"]),
}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    "-I", CSRC, "-I", os.path.join(ROOT, "include"),
]
DEVICE = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]


def _sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _headers():
    return sorted(glob.glob(os.path.join(CSRC, "*.hpp")) +
                  glob.glob(os.path.join(ROOT, "include", "*.h")))


def _compile(src: str, force: bool, build_dir: str = BUILD, defines=()) -> str:
    obj = os.path.join(build_dir, os.path.basename(src) + ".o")
    newest = max([os.path.getmtime(src)] + [os.path.getmtime(h) for h in _headers()])
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= newest:
        return obj
    cmd = [HIPCC, *COMMON, *DEVICE, *defines, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(force: bool = True, jobs: int = 8, variant: str = "release") -> str:
    """Compile and link libtrid.so (variant "release") or the LDS-poisoning
    debug library libtrid_lds_poison.so ("lds_poison", -DTRID_LDS_POISON).
    ``force`` recompiles every object."""
    lib_path, build_dir, defines = VARIANTS[variant]
    os.makedirs(build_dir, exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force, build_dir, defines), srcs))
    tmp = lib_path + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib_path)
    return lib_path


if __name__ == "__main__":
    print(build(force="--force" in sys.argv,
                variant="lds_poison" if "--lds-poison" in sys.argv else "release"))
