"""CPU checks of the drop-in boundary: libssg_hip.so loads and exports every entry
point include/ssg.h declares, with the signatures the Python host binds."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ssg.h")
LIB = os.path.join(ROOT, "similaritysearch.jl_amd", "lib", "libssg_hip.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ssg_[a-z_]+)\s*\(", src)))


def test_header_declares_the_api():
    names = declared()
    for must in ["ssg_create", "ssg_destroy", "ssg_upload_db", "ssg_upload_graph",
                 "ssg_exhaustive_knn", "ssg_beamsearch", "ssg_allknn_graph", "ssg_rerank",
                 "ssg_build_graph", "ssg_optimize_minrecall", "ssg_append_db", "ssg_neardup",
                 "ssg_hsp_queries", "ssg_hsp_queries_dev"]:
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libssg_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="libssg_hip.so not built")
def test_python_bindings_cover_the_header():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "similaritysearch.jl_amd"))
    from similaritysearch import _lib
    assert sorted(_lib.SIGNATURES) == declared()
    lib = _lib.load()
    assert lib.ssg_version().decode().startswith("ssg_hip")


def test_null_handle_is_an_argument_error():
    if not os.path.exists(LIB):
        pytest.skip("libssg_hip.so not built")
    import sys
    sys.path.insert(0, os.path.join(ROOT, "similaritysearch.jl_amd"))
    from similaritysearch import _lib
    lib = _lib.load()
    assert lib.ssg_destroy(None) == 0
    assert lib.ssg_upload_db(None, None, 4, 10, 0) == -1
    assert b"null handle" in lib.ssg_last_error(None)
