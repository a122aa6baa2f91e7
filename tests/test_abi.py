"""The C ABI without a GPU: the library loads, exports every function the
header declares, agrees with the header's constants, rejects bad handles and
shapes before touching the device, and the Fortran ISO_C_BINDING module
compiles against it."""
import ctypes as ct
import os
import re
import shutil
import subprocess

import pytest

from mohid_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mohid_gpu.h")
FORTRAN = os.path.join(ROOT, "mohid_amd", "fortran", "ModuleAdvectionDiffusionGPU.F90")


def _header():
    with open(HEADER) as f:
        return f.read()


def _declared_functions():
    text = re.sub(r"/\*.*?\*/", "", _header(), flags=re.S)
    return sorted(set(re.findall(r"\b(mgpu_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    declared = _declared_functions()
    assert declared, "no prototypes parsed"
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == declared


def test_header_constants_match_binding():
    defs = dict(re.findall(r"#define\s+(MGPU_[A-Z0-9_]+)\s+\(?(-?\d+)\)?", _header()))
    assert defs, "no constants parsed"
    for name, value in defs.items():
        assert getattr(_lib, name) == int(value), name


def test_struct_layouts_match_header_order():
    """Field names of the ctypes mirrors appear in the header's struct
    definitions in the same order."""
    text = _header()
    for struct, cls in (("mgpu_size3d", _lib.Size3D), ("mgpu_geometry", _lib.Geometry),
                        ("mgpu_ad_args", _lib.ADArgs), ("mgpu_transport_opts",
                                                        _lib.TransportOpts)):
        m = re.search(r"typedef struct\s*(?:\w+\s*)?\{([^{}]*)\}\s*" + struct + r"\s*;", text)
        assert m, struct
        body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
        fields = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*[;,]", body)
        assert fields == [f[0] for f in cls._fields_], struct


def test_rejects_bad_handles_and_shapes_without_a_device():
    L = _lib.lib()
    assert L.mgpu_version().startswith(b"mohid")
    assert L.mgpu_ad_destroy(None) == _lib.MGPU_OFF_ERR
    assert L.mgpu_ad_step(None, None, None) == _lib.MGPU_OFF_ERR
    assert L.mgpu_transport_run(None, 1, 0) == _lib.MGPU_OFF_ERR
    assert L.mgpu_transport_destroy(None) == _lib.MGPU_OFF_ERR
    size = _lib.Size3D(0, 5, 0, 6, 0, 4)
    ws = _lib.Size3D(1, 4, 1, 5, 1, 3)
    h = ct.c_void_p()
    assert L.mgpu_ad_create(ct.byref(size), ct.byref(ws), 1, 0, 0, None, None) == \
        _lib.MGPU_ERR_BADARG
    # WorkSize must start at 1 and Size must be WorkSize +/- 1
    for bad in (_lib.Size3D(0, 4, 1, 5, 1, 3), _lib.Size3D(1, 5, 1, 5, 1, 3)):
        assert L.mgpu_ad_create(ct.byref(size), ct.byref(bad), 1, 0, 0, None,
                                ct.byref(h)) == _lib.MGPU_ERR_BADARG
    assert L.mgpu_thomasz(ct.byref(size), ct.byref(ws), None, None, None, None, None, None,
                          0, None) == _lib.MGPU_ERR_BADARG


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/amdflang"), reason="amdflang missing")
def test_fortran_binding_compiles(tmp_path):
    out = subprocess.run(
        ["/opt/rocm/bin/amdflang", "-c", "-cpp", "-fdefault-real-8", "-fdefault-double-8",
         "-O0", FORTRAN, "-o",
         str(tmp_path / "adgpu.o"), "-J", str(tmp_path)],
        capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert (tmp_path / "adgpu.o").exists()


def test_autotune_candidates_are_valid_shapes():
    from mohid_amd.autotune import FUSED, LANES, candidates
    for T in range(1, 13):
        c = candidates(T)
        assert c, T
        for kern, pl, e, col in c:
            assert kern in (LANES, FUSED)
            chunk = min(pl or 8, T)
            assert chunk % col == 0 and col <= chunk
            if kern == LANES:
                assert chunk % e == 0 and e <= chunk
            else:
                assert e == 0 and col <= 4      # the fused walk: no explicit shape
            assert pl == 0 or pl < T
        assert len(set(c)) == len(c)
    assert (LANES, 0, 4, 2) in candidates(8)
    assert (FUSED, 0, 0, 2) in candidates(8)
