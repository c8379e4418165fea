"""CPU oracle for the PIC hot path -- TEST INFRASTRUCTURE ONLY.

Two ctypes-loaded libraries live here:

* ``orc`` -- ``oracle/liboracle.so``, the C FP64 restatement of the reference
  kernels (``oracle/pxr_oracle.c``, built by ``oracle/Makefile``).
* ``ref`` -- ``oracle/_ref/libpxrref.so``, the reference's own Fortran leaf
  kernels compiled from the PICSAR sources by ``oracle/build_ref.sh``.  It is
  optional: it exists wherever it was built (it is carried along with the
  tree but never committed).  Beside it, ``libpxrref_o3.so`` (the same
  kernels with the reference's production flags) and
  ``libpxrref_driver.so`` (``ref_tile_driver.c``, an OpenMP loop over tiles
  that calls them) give bench.py's CPU baseline.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
(``picsar_amd``) never imports it.
"""
from .pxr_oracle import (Oracle, Reference, RefTileDriver, build_oracle,  # noqa: F401
                         load_oracle, load_ref_driver, load_reference)
