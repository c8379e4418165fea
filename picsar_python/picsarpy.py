"""``picsarpy.picsar``: PICSAR's module variables, set-up routines, managers
and leaf routines (picsar_amd.pxrmodule.PicsarModule) under the name the
reference's Forthon module has."""
from picsar_amd.pxrmodule import PicsarModule, picsar  # noqa: F401
