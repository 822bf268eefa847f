"""The C-ABI library loads and exports every entry point include/transfer.h
declares (no device calls)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "transfer.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tr_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    import transfer_amd
    lib = ctypes.CDLL(transfer_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_layouts_match_reference():
    import transfer_amd as ta
    # tetrahedron.h:20 and face.h:18 assert 128 and 64 bytes
    assert ta.TET_DTYPE.itemsize == 128
    assert ta.FACE_DTYPE.itemsize == 64


def test_errors_are_reported_not_thrown():
    import transfer_amd as ta
    import pytest
    with pytest.raises(ta.TransferError, match="angorder"):
        ta.Angular(9)
    with pytest.raises(ta.TransferError):
        ta.Mesh.box((0, 1, 1), (0, 0, 0), (1, 1, 1))


def test_batch_create_rejects_unknown_flags():
    """tr_batch_create takes 0 or TR_CTX_GENERAL_NORMALS; any other flag is
    an argument error, raised before any device call."""
    import ctypes as C
    import numpy as np
    import pytest
    import transfer_amd as ta
    from transfer_amd.api import lib, check, MeshView
    m = ta.Mesh.box((2, 2, 2), (0, 0, 0), (1, 1, 1))
    ang = ta.Angular(1)
    views = (MeshView * 1)(m.view)
    h = C.c_void_p()
    alb = np.zeros(1)
    with pytest.raises(ta.TransferError, match="flags"):
        check(lib.tr_batch_create(0, 1, views, C.byref(ang.view), alb.ctypes.data_as(C.POINTER(C.c_double)),
                                  ta.TR_CTX_GENERAL_NORMALS << 1, C.byref(h)))
