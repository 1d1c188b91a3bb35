"""Loading the reference build (oracle/_ref) or the oracle's own C must not change the
floating-point mode of the process.  A shared object linked with -ffast-math carries
crtfastmath.o, whose constructor sets flush-to-zero / denormals-are-zero for every thread
of the loading process; the reference links its extension without it
(create_module.py:14-15), and so does oracle/Makefile.  Checked in a fresh interpreter so
that nothing loaded earlier in the test session hides the effect."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROBE = r"""
import ctypes, sys
import numpy as np
for p in sys.argv[1:]:
    ctypes.CDLL(p)
    f32 = np.float32(1e-40) * np.float32(1.0)
    f64 = np.float64(5e-324) * np.float64(1.0)
    if f32 == 0 or f64 == 0:
        print("FTZ after loading", p)
        sys.exit(1)
print("ok", len(sys.argv) - 1)
"""


def test_loading_the_c_checkers_keeps_subnormals():
    libs = sorted(glob.glob(os.path.join(ROOT, "oracle", "_ref", "*.so")))
    libs += sorted(glob.glob(os.path.join(ROOT, "oracle", "_build", "*.so")))
    if not libs:
        pytest.skip("oracle/_ref and oracle/_build are not built")
    r = subprocess.run([sys.executable, "-c", PROBE] + libs, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok %d" % len(libs)
