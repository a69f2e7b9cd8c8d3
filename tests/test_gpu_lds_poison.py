"""The LDS-poisoning debug build on the GPU. libtrid_lds_poison.so is the
library compiled with -DTRID_LDS_POISON (by __graft_entry__.build()): every
LDS staging and exchange area -- the LDS-resident CR / CR-PCR systems, the
pcrGauss LDS form and global-level records, the staged strided elimination,
the interleaved gather tiles, the register kernels' coalescing stage, the
multi-wave exchange, the ragged stage and exchange areas, the separator
records, the ADI maps and tiles -- is filled with NaN before each use. The
parity tests of those kernels run once against it in a child process
(TRID_LIBRARY selects the library) and must pass: no kernel reads an LDS
element that it did not write."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "tridiagonalsolvers_amd", "libtrid_lds_poison.so")
FILES = ["test_gpu_parity.py", "test_gpu_methods.py", "test_gpu_interleaved.py",
         "test_gpu_part.py", "test_gpu_ragged.py", "test_gpu_refine.py", "test_gpu_strided.py",
         "test_gpu_adi.py", "test_gpu_pivoting.py"]
# the largest batches repeat what smaller ones cover and take the most time
SKIP = "not full_size and not 100k and not beyond_2_31 and not million"


def test_lds_kernels_pass_with_poisoned_lds():
    assert os.path.exists(LIB), "built by __graft_entry__.build()"
    env = dict(os.environ, TRID_LIBRARY=LIB)
    probe = subprocess.run([sys.executable, "-c", "import tridiagonalsolvers_amd._lib as L; "
                            "L.lib(); print(L.LIB_PATH)"],
                           env=env, cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert probe.returncode == 0 and probe.stdout.strip().endswith("libtrid_lds_poison.so"), \
        probe.stdout + probe.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "gpu",
                        "-p", "no:cacheprovider", "-k", SKIP,
                        *(os.path.join(ROOT, "tests", f) for f in FILES)],
                       env=env, cwd=ROOT, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-3000:]
    print(tail)
    assert r.returncode == 0, tail
