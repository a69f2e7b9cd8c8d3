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
NaN in the parity tests. The ``asan`` and ``tsan`` variants
(``_san/libtrid_asan.so``, ``_san/libtrid_tsan.so``) sanitize the host code
only (AddressSanitizer + UndefinedBehaviorSanitizer, or ThreadSanitizer; the
device code is the release code), for the CPU tests run with the runtime
preloaded, e.g.::

    ASAN_OPTIONS=detect_leaks=0 \
    LD_PRELOAD=/opt/rocm/llvm/lib/clang/22/lib/linux/libclang_rt.asan-x86_64.so \
    TRID_LIBRARY=tridiagonalsolvers_amd/_san/libtrid_asan.so \
    python -m pytest tests -m "not gpu"
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
_ASAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]
_TSAN = ["-Xarch_host", "-fsanitize=thread"]
_NOFP = ["-Xarch_host", "-fno-omit-frame-pointer"]
# variant: (library, object directory, compile flags, link flags)
VARIANTS = {
    "release": (LIB, BUILD, [], []),
    "lds_poison": (LIB_LDS_POISON, os.path.join(HERE, "_build_lds_poison"), ["-DTRID_LDS_POISON"],
                   []),
    "asan": (os.path.join(HERE, "_san", "libtrid_asan.so"), os.path.join(HERE, "_build_asan"),
             _ASAN + _NOFP + ["-Xarch_host", "-fno-sanitize-recover=undefined"], _ASAN),
    "tsan": (os.path.join(HERE, "_san", "libtrid_tsan.so"), os.path.join(HERE, "_build_tsan"),
             _TSAN + _NOFP, _TSAN),
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
    """Compile and link libtrid.so (variant "release"), the LDS-poisoning
    debug library libtrid_lds_poison.so ("lds_poison", -DTRID_LDS_POISON) or
    a host-sanitized library ("asan", "tsan"). ``force`` recompiles every
    object."""
    lib_path, build_dir, defines, link_flags = VARIANTS[variant]
    os.makedirs(build_dir, exist_ok=True)
    os.makedirs(os.path.dirname(lib_path), exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force, build_dir, defines), srcs))
    tmp = lib_path + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *link_flags, "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib_path)
    return lib_path


if __name__ == "__main__":
    variant = "release"
    for v in VARIANTS:
        if "--" + v.replace("_", "-") in sys.argv:
            variant = v
    print(build(force="--force" in sys.argv, variant=variant))
