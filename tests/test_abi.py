"""The drop-in boundary: the C-ABI library loads without a GPU and exports
every function that include/*.h declares (no compute calls)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "acd-alm_amd", "lib", "libacdalm.so")
KEYWORDS = {"if", "while", "for", "return", "sizeof", "void", "int", "double",
            "float", "char", "size_t", "uint32_t", "uint64_t", "defined"}


def declared_functions():
    out = {}
    for h in sorted(glob.glob(os.path.join(ROOT, "include", "**", "*.h"),
                              recursive=True)):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
        text = re.sub(r"//[^\n]*", " ", text)
        text = re.sub(r"^\s*#[^\n]*", " ", text, flags=re.M)
        text = re.sub(r"typedef[^;]*;", " ", text)
        names = re.findall(r"[\s*]([A-Za-z_]\w*)\s*\([^;{}]*\)\s*;", text)
        out[os.path.relpath(h, ROOT)] = [n for n in names if n not in KEYWORDS]
    return out


def test_headers_declare_the_reference_entry_points():
    decl = {k: set(v) for k, v in declared_functions().items()}
    assert {"alm_make", "alm_free", "alm_solve", "alm_read",
            "read_doubles"} <= decl["include/alm.h"]
    assert {"approx_make", "approx_update", "approx_solve", "mpg_solve",
            "cd_solve"} <= decl["include/approx/approx.h"]
    assert {"sparse_matrix_make", "sparse_matrix_multiply",
            "sparse_matrix_multiply_2", "sparse_matrix_row_permute",
            "sparse_matrix_col_permute"} <= decl["include/spmv/spmv.h"]
    assert {"thread_pool_init", "thread_pool_free"} <= \
        decl["include/thread_pool/thread_pool.h"]


def test_library_exports_every_declared_function():
    if not os.path.exists(LIB):
        pytest.fail("libacdalm.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [(h, n) for h, names in declared_functions().items()
               for n in names if not hasattr(lib, n)]
    assert not missing, missing
