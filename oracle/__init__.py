"""CPU oracle for the radon projector path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``, always as the reference the GPU results are compared
with or timed against.  The product package never imports it and never falls back to it.

* :mod:`oracle.radon`   -- ctypes bindings of ``radon_oracle.c`` (float32 IEEE, float32
  gcc-fast-math form, float64) plus the Cython-wrapper semantics around them.
* :mod:`oracle.ref`     -- ctypes bindings of the reference's own radon C, compiled from
  the reference sources into ``oracle/_ref`` by ``oracle/Makefile`` (optional).
* :mod:`oracle.solvers` -- numpy restatements of the SIRT / MLEM / SPS / CG / ADMM-TV
  iterations that sit on top of the projector pair.
"""
