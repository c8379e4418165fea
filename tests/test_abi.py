"""CPU-side checks of the C ABI: the library loads and exports every entry
point include/pxrhip.h declares; host-side argument checks raise before any
device work.  No kernel is launched here."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pxrhip.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pxrhip_\w+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib_path():
    from picsar_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from picsar_amd.build import build
        build(verbose=False)
    return _lib.LIB_PATH


def test_library_exports_every_declared_symbol(lib_path):
    from picsar_amd import _lib
    names = declared()
    assert len(names) >= 20
    exported = set(_lib.exported_symbols(lib_path))
    missing = [n for n in names if n not in exported]
    assert not missing, f"declared in pxrhip.h but not exported: {missing}"


def test_python_binding_covers_header(lib_path):
    from picsar_amd import _lib
    bound = set(_lib.SIGNATURES) | {"pxrhip_last_error", "pxrhip_get_stream"}
    assert set(declared()) <= bound


def declared_arity():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(pxrhip_\w+)\s*\(([^)]*)\)\s*;", txt):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_python_argument_counts_match_header(lib_path):
    from picsar_amd import _lib
    for name, n in declared_arity().items():
        if name in _lib.SIGNATURES:
            assert len(_lib.SIGNATURES[name]) == n, (name, len(_lib.SIGNATURES[name]), n)


def test_library_is_gfx950(lib_path):
    data = open(lib_path, "rb").read()
    assert b"gfx950" in data


def test_host_argument_checks_raise_before_device_work(lib_path):
    from picsar_amd.picsarpy import picsar
    x = np.zeros(4, dtype=np.float32)
    with pytest.raises(TypeError):
        picsar.pxr_pushxyz(4, x, x, x, x, x, x, x, 0.1)
    y = np.zeros(2)
    with pytest.raises(ValueError):
        picsar.pxr_set_gamma(4, y, y, y, y)
    with pytest.raises(ValueError):
        picsar.pxrpush_em3d_evec_norder(*[np.zeros((5, 5, 5), order="F")] * 9, 0.1, [1.0],
                                        [1.0], [1.0], 2, 2, 2, 4, 4, 4, 1, 1, 1, 0, 0, 0)


def test_struct_layouts_match_the_header(tmp_path):
    """The ctypes mirrors of pxrhip_grid and pxrhip_species have the C
    layout: a C program compiled against include/pxrhip.h prints the sizes
    and field offsets."""
    import ctypes as C
    import shutil
    import subprocess
    from picsar_amd.sim import PxrGrid, PxrSpecies
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "pxrhip.h"\n'
        "int main(void) {\n"
        '  printf("%zu %zu %zu %zu\\n", sizeof(pxrhip_grid), offsetof(pxrhip_grid, pbound),\n'
        "         offsetof(pxrhip_grid, partcom), sizeof(pxrhip_species));\n"
        '  printf("%zu %zu %zu %zu\\n", offsetof(pxrhip_species, capacity),\n'
        "         offsetof(pxrhip_species, precision), offsetof(pxrhip_species, npid_extra),\n"
        "         offsetof(pxrhip_species, pid_extra));\n"
        "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    got = [int(v) for v in out]
    want = [C.sizeof(PxrGrid), PxrGrid.pbound.offset, PxrGrid.partcom.offset, C.sizeof(PxrSpecies),
            PxrSpecies.capacity.offset, PxrSpecies.precision.offset,
            PxrSpecies.npid_extra.offset, PxrSpecies.pid_extra.offset]
    assert got == want
