"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls here)."""
import ctypes
import glob
import os
import re

from conftest import ROOT


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(mpcf_\w+)\s*\(", text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("mpcf_create", "mpcf_destroy", "mpcf_upload", "mpcf_download", "mpcf_load_ghosts",
                 "mpcf_max_sos", "mpcf_convection", "mpcf_update", "mpcf_process_all", "mpcf_step"):
        assert must in names, must


def test_library_exports_every_declared_symbol(mpcf):
    lib = mpcf.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # the binding's own list stays in sync with the header
    assert sorted(set(mpcf.EXPORTS)) == sorted(set(mpcf.EXPORTS) & set(declared_functions()))


def test_error_path_without_device_work(mpcf):
    lib = mpcf.lib()
    rc = lib.mpcf_create(None, 8, 8, 8, 0.1, 0)
    assert rc == mpcf.EINVAL
    assert b"NULL" in lib.mpcf_last_error()
    h = ctypes.c_void_p()
    assert lib.mpcf_create(ctypes.byref(h), 4, 8, 8, 0.1, 0) == mpcf.EINVAL   # extents >= 5
    assert lib.mpcf_create(ctypes.byref(h), 8, 8, 8, -1.0, 0) == mpcf.EINVAL  # h > 0
    assert lib.mpcf_create(ctypes.byref(h), 8, 8, 8, 0.1, 64) == mpcf.EINVAL  # unknown flag
    assert lib.mpcf_create(ctypes.byref(h), 8, 8, 8, 0.1, 256 | 32) == mpcf.EINVAL  # E_PAIR with FP64
    assert lib.mpcf_create(ctypes.byref(h), 8, 8, 8, 0.1, 256 | 4) == mpcf.EINVAL   # E_PAIR with SPLIT_XY
    assert lib.mpcf_set_bc(None, None) == mpcf.EINVAL
    assert lib.mpcf_step(None, 0.3, 1.0, None) == mpcf.EINVAL


def test_noalloc_refuses_bad_arguments_before_device_work(mpcf):
    """mpcf_max_sos_noalloc checks its arguments before it touches a device:
    a null state, a state beyond 2^31 - 1 cells, a field stride below the
    cell count and unknown flags are refused with MPCF_EINVAL (the pointer
    below is never dereferenced: every call fails its checks first)."""
    lib = mpcf.lib()
    fake = 4096
    assert lib.mpcf_max_sos_noalloc(None, 8, 2, 2, 2, 0.1, 0.3, 1.0, 0, None, None, None, None) == mpcf.EINVAL
    assert b"NULL" in lib.mpcf_last_error()
    rc = lib.mpcf_max_sos_noalloc(fake, 1 << 33, 2048, 2048, 1024, 0.1, 0.3, 1.0, 0, None, None, None, None)
    assert rc == mpcf.EINVAL
    assert b"2^31" in lib.mpcf_last_error()
    assert lib.mpcf_max_sos_noalloc(fake, 7, 2, 2, 2, 0.1, 0.3, 1.0, 0, None, None, None, None) == mpcf.EINVAL
    assert lib.mpcf_max_sos_noalloc(fake, 8, 2, 2, 2, 0.1, 0.3, 1.0, 4, None, None, None, None) == mpcf.EINVAL
    assert lib.mpcf_max_sos_noalloc(fake, 8, 2, 2, 2, 0.0, 0.3, 1.0, 0, None, None, None, None) == mpcf.EINVAL
