"""The C-ABI library builds, loads on a host without a GPU and exports every
entry point declared in include/slab_hip.h (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "slab_hip.h")).read()
    return sorted(set(re.findall(r"^(?:int|size_t)\s+(slab_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("slab_gemm", "slab_getrf", "slab_btri_factor", "slab_btri_solve", "slab_btri_inv_factor",
                 "slab_btri_inv_solve", "slab_leaf_assemble", "slab_leaf_assemble_nodes", "slab_leaf_residual", "slab_btri_factor_twisted", "slab_btri_solve_twisted",
                 "slab_btri_fused", "slab_getri", "slab_getri_ws"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from slablu_amd import hip
    if not os.path.exists(hip.LIB_PATH):
        from slablu_amd import _build
        _build.build()
    lib = ctypes.CDLL(hip.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.slab_version() >= 100


def test_bad_arguments_are_rejected_without_a_gpu():
    """Argument validation happens before any launch: negative sizes and
    inconsistent LU shapes return -1."""
    from slablu_amd import hip
    if not os.path.exists(hip.LIB_PATH):
        from slablu_amd import _build
        _build.build()
    lib = ctypes.CDLL(hip.LIB_PATH)
    lib.slab_getrf.restype = ctypes.c_int
    args = [ctypes.c_int(0), ctypes.c_int(1), ctypes.c_int(10), ctypes.c_int(10), ctypes.c_int(11),
            ctypes.c_int(10), ctypes.c_void_p(0), ctypes.c_int64(10), ctypes.c_int64(100), ctypes.c_void_p(0),
            ctypes.c_int64(10), ctypes.c_void_p(0), ctypes.c_void_p(0)]
    assert lib.slab_getrf(*args) == -1  # nelim > npiv
    assert lib.slab_btri_inv_fused(0, 100) == 1 and lib.slab_btri_inv_fused(0, 1000) == 0
    assert lib.slab_btri_inv_fused(7, 10) == -1
    assert lib.slab_btri_fused(1, 100) == 1 and lib.slab_btri_fused(0, 100) == 1
    assert lib.slab_btri_fused(0, 1000) == 0 and lib.slab_btri_fused(3, 10) == -1
    # twisted factorization needs blocks that fit the fused kernels
    lib.slab_btri_factor_twisted.restype = ctypes.c_int
    targs = [ctypes.c_int(0), ctypes.c_int(1), ctypes.c_int(3), ctypes.c_int(1000)] + \
        [ctypes.c_void_p(0), ctypes.c_int64(0), ctypes.c_int64(0)] * 3 + \
        [ctypes.c_void_p(0), ctypes.c_int64(0), ctypes.c_int64(0), ctypes.c_void_p(0), ctypes.c_void_p(0)]
    assert lib.slab_btri_factor_twisted(*targs) == -1
    # getri workspace: W and X (2 batch padded matrices each), pair products,
    # pivots; a missing workspace is rejected
    lib.slab_getri_ws.restype = ctypes.c_size_t
    np_, b = 1024, 2
    assert lib.slab_getri_ws(0, b, 1000) == 8 * (4 * b * np_ * np_ + 2 * b * np_ * np_ // 4) + 4 * b * 1000 + 256
    assert lib.slab_getri_ws(5, b, 1000) == 0
    lib.slab_getri.restype = ctypes.c_int
    gargs = [ctypes.c_int(0), ctypes.c_int(1), ctypes.c_int(8), ctypes.c_void_p(0), ctypes.c_int64(8),
             ctypes.c_int64(64), ctypes.c_void_p(0), ctypes.c_int64(8), ctypes.c_void_p(0), ctypes.c_int64(8),
             ctypes.c_int64(64), ctypes.c_void_p(0), ctypes.c_void_p(0)]
    assert lib.slab_getri(*gargs) == -1
