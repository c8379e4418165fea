"""CPU oracle for the PIC hot path -- TEST INFRASTRUCTURE ONLY.

Two ctypes-loaded libraries live here:

* ``orc`` -- ``oracle/liboracle.so``, the C FP64 restatement of the reference
  kernels (``oracle/pxr_oracle.c``, built by ``oracle/Makefile``).
* ``ref`` -- ``oracle/_ref/libpxrref.so``, the reference's own Fortran leaf
  kernels compiled from the PICSAR sources by ``oracle/build_ref.sh``.  It is
  optional: it exists wherever it was built (it is carried along with the
  tree but never committed).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``picsar_amd``) never imports it.
"""
from .pxr_oracle import Oracle, Reference, load_oracle, load_reference, build_oracle  # noqa: F401
