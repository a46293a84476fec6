"""CPU checks of the C ABI: the library loads and exports every symbol that
include/*.h declares, and the ctypes table covers exactly those symbols.
No compute call is made (there is no GPU here)."""
import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(nb_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "nb_lbvh_build_points" in names and "nb_traverse_count" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from neighbor_amd import _capi
    L = ctypes.CDLL(_capi.library_path())
    missing = [n for n in sorted(declared_symbols()) if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    from neighbor_amd import _capi
    table = {name for name, _, _ in _capi.SIGNATURES}
    assert table == declared_symbols()
    _capi.lib()  # binds every signature


def test_version_and_error_string_without_gpu():
    from neighbor_amd import _capi
    L = _capi.lib()
    assert L.nb_version() >= 100
    # invalid-argument paths return codes without touching the device
    assert L.nb_lbvh_info(None, None, None, None, None) == _capi.NB_ERR_INVALID
    assert b"null" in L.nb_last_error_string()
    assert L.nb_traverse_count(None, None, None, 0, None, 0, None, None, None, 256, None) == _capi.NB_ERR_INVALID


def test_batch_argument_errors_without_gpu():
    """Batch entry points reject bad arguments before any device work."""
    from neighbor_amd import _capi
    L = _capi.lib()
    h = ctypes.c_void_p()
    assert L.nb_batch_create(ctypes.byref(h)) == _capi.NB_OK
    try:
        assert L.nb_batch_build(h, None, 0, None, None, None, 0, None) == _capi.NB_ERR_INVALID      # no boxes
        assert b"box" in L.nb_last_error_string()
        dummy = ctypes.c_void_p(0x1000)
        assert L.nb_batch_build(h, None, 5, dummy, dummy, dummy, 1, None) == _capi.NB_ERR_INVALID   # null points
        assert L.nb_batch_set_params(h, 100, 16, 0, 0) == _capi.NB_ERR_TUNING                     # block not /32
        assert L.nb_batch_set_params(h, 256, 65, 0, 0) == _capi.NB_ERR_INVALID                    # bucket > 64
        assert L.nb_batch_set_params(h, 256, 16, 7, 0) == _capi.NB_ERR_INVALID                    # variant
        assert L.nb_batch_set_params(h, 512, 32, _capi.NB_BATCH_LANE, 64) == _capi.NB_OK
        b, k, g = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
        v = ctypes.c_int()
        assert L.nb_batch_get_params(h, ctypes.byref(b), ctypes.byref(k), ctypes.byref(v), ctypes.byref(g)) == 0
        assert (b.value, k.value, v.value, g.value) == (512, 32, _capi.NB_BATCH_LANE, 64)
        assert L.nb_batch_plan_launch(None, None) == _capi.NB_ERR_INVALID
    finally:
        assert L.nb_batch_destroy(h) == _capi.NB_OK


def test_header_compiles_as_plain_c_and_cpp():
    """include/neighbor_amd.h is usable from C (cgo, ctypes header generators, plain C hosts) and
    from C++ without HIP: no C++ or HIP types in the signatures."""
    import shutil
    import subprocess
    hdr = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "neighbor_amd.h")
    for compiler, args in (("gcc", ["-x", "c", "-std=c99"]), ("g++", ["-x", "c++", "-std=c++11"])):
        if shutil.which(compiler) is None:
            continue
        p = subprocess.run([compiler, "-fsyntax-only", "-Wall", "-Wextra", "-Werror", *args, hdr], capture_output=True,
                           text=True)
        assert p.returncode == 0, p.stderr
