"""Build libitomo.so (HIP, gfx950) in-tree with hipcc.

Every HIP translation unit under ``csrc/`` is compiled and linked into
``libitomo.so`` next to this file.  Flags:

* ``--offload-arch=gfx950``: MI355X only.
* ``-ffp-contract=off``: the projectors reproduce the reference's float arithmetic
  operation for operation; fused multiply-adds are written explicitly where the
  reference build has them, and nowhere else.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
* ``-fno-honor-nans`` for the distance-driven translation units only: their min/max
  of boundary keys then compile to a single v_min/v_max with no sNaN-quieting
  canonicalize.  Results on finite data are unchanged; the reference build itself
  assumes finite data (-ffast-math implies -ffinite-math-only).  The ray-driven
  walk tests for NaN and keeps full NaN semantics.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libitomo.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
    "-ffp-contract=off", "-Wall", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# extra flags per translation unit
PER_FILE = {
    "dd.hip": ["-fno-honor-nans", "-falign-loops=64"],
    "batch.hip": ["-fno-honor-nans", "-falign-loops=64"],
}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)


def _finish(proc):
    out, _ = proc.communicate()
    msg = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + out)
    return msg


def build(verbose=False):
    """Compile libitomo.so (always recompiles): one object per source, in parallel,
    then one link."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC]
    compile_flags = [f for f in FLAGS if f != "-shared"]
    objs, procs = [], []
    for src in sources():
        obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
        extra = PER_FILE.get(os.path.basename(src), [])
        procs.append(_run([HIPCC] + compile_flags + extra + inc + ["-c", src, "-o", obj],
                          verbose))
        objs.append(obj)
    msgs = [_finish(p) for p in procs]
    link = _run([HIPCC, "--offload-arch=gfx950", "-fPIC", "-shared"] + objs
                + ["-o", LIB + ".tmp"], verbose)
    msgs.append(_finish(link))
    if verbose:
        for m in msgs:
            if m:
                print(m)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
