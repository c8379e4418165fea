"""Drop-in for the reference's ``picsar_python`` package
(python_module/picsar_python/__init__.py): ``from picsar_python import
picsarpy`` gives ``picsarpy.picsar``, the module state and routines of
PICSAR for one rank, run on the GPU by picsar_amd."""
from . import picsarpy  # noqa: F401
