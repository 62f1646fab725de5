"""CPU checks of the C ABI: the library loads and exports every function
declared in include/*.h; argument validation returns errors (no GPU work is
launched by any of these calls)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(qphix_amd_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "qphix_amd_dslash" in names and "qphix_amd_tm_clov_dslash_achimbdpsi" in names
    assert len(names) >= 19


def test_library_exports_every_declared_symbol(qamd):
    lib = ctypes.CDLL(qamd.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header(qamd):
    from importlib import import_module
    sig = import_module("qphix_codegen_amd._lib").SIGNATURES
    assert set(declared_functions()) <= set(sig), set(declared_functions()) - set(sig)


def test_ctx_validation(qamd):
    lib = qamd.lib()
    ctx = ctypes.c_void_p()
    assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), 3, 4, 4, 4, -1.0) == -2
    assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), 4, 4, 4, 5, -1.0) == -2
    assert lib.qphix_amd_ctx_create(None, 4, 4, 4, 4, -1.0) == -1
    # every field index must fit 32 bits: 72 reals x sites per checkerboard < 2^31
    assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), 128, 128, 128, 128, -1.0) == -2
    assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), 96, 96, 96, 96, -1.0) == -2
    # extents whose product overflows 32 or 64 bits are refused, not wrapped
    # (65536^4 / 2 = 2^63 wraps a signed 64-bit product to a negative number)
    for dims in ((65536, 65536, 65536, 65536), (65536, 65536, 2, 2), (2, 2, 65536, 65536),
                 (1 << 30, 2, 2, 2), (2, 2, 2, 1 << 30), (2, 65534, 65536, 2)):
        assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), *dims, -1.0) == -2, dims
    big = ctypes.c_void_p()  # 64^4: 72 x 8,388,608 < 2^31
    assert lib.qphix_amd_ctx_create(ctypes.byref(big), 64, 64, 64, 64, -1.0) == 0
    lib.qphix_amd_ctx_destroy(big)
    small = ctypes.c_void_p()
    assert lib.qphix_amd_ctx_create(ctypes.byref(small), 2, 2, 2, 2, -1.0) == 0  # smallest lattice
    assert lib.qphix_amd_sites_cb(small) == 8
    lib.qphix_amd_ctx_destroy(small)
    assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), 8, 8, 8, 8, -1.0) == 0
    try:
        assert lib.qphix_amd_sites_cb(ctx) == 2048
        assert lib.qphix_amd_field_bytes(ctx, 0, 1, 0) == 2048 * 24 * 4
        assert lib.qphix_amd_field_bytes(ctx, 1, 2, 12) == 2048 * 96 * 2
        assert lib.qphix_amd_field_bytes(ctx, 1, 0, 13) == 0
        assert lib.qphix_amd_field_bytes(ctx, 3, 3, 0) == 2048 * 144 * 2
        # invalid isign / cb / precision / null pointers are rejected up front
        buf = ctypes.create_string_buffer(64)
        p = ctypes.cast(buf, ctypes.c_void_p)
        assert lib.qphix_amd_dslash(ctx, 1, 12, 2, 0, p, p, p, 1.0, 1.0, 1.0, None) == -1
        assert lib.qphix_amd_dslash(ctx, 1, 12, 1, 2, p, None, p, 1.0, 1.0, 1.0, None) == -1
        assert lib.qphix_amd_dslash(ctx, 7, 12, 1, 0, p, p, p, 1.0, 1.0, 1.0, None) == -1
        assert lib.qphix_amd_dslash(ctx, 1, 11, 1, 0, p, p, p, 1.0, 1.0, 1.0, None) == -1
        assert lib.qphix_amd_clov_dslash(ctx, 1, 12, 1, 0, p, p, p, None, 1.0, 1.0, 1.0, None) == -1
        assert lib.qphix_amd_error_string(-2).startswith(b"unsupported lattice")
    finally:
        lib.qphix_amd_ctx_destroy(ctx)


def test_face_entry_points_validate(qamd):
    lib = qamd.lib()
    ctx = ctypes.c_void_p()
    assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), 8, 4, 6, 10, -1.0) == 0
    try:
        # half-spinors per face: one checkerboard of the slice
        assert [lib.qphix_amd_face_sites(ctx, d) for d in range(4)] == [10 * 6 * 2, 10 * 6 * 4,
                                                                       10 * 4 * 4, 6 * 4 * 4]
        assert lib.qphix_amd_face_sites(ctx, 4) == 0
        assert lib.qphix_amd_set_comm(ctx, None) == -1
        buf = ctypes.create_string_buffer(256)
        p = ctypes.c_void_p((ctypes.addressof(buf) + 15) & ~15)
        q = ctypes.c_void_p(p.value + 64)
        assert lib.qphix_amd_face_pack(ctx, 1, 1, 0, 4, 0, p, q, None) == -1   # dim
        assert lib.qphix_amd_face_pack(ctx, 1, 1, 0, 0, 3, p, q, None) == -1   # fb (2: both faces)
        assert lib.qphix_amd_face_pack(ctx, 1, 1, 0, 0, -1, p, q, None) == -1
        assert lib.qphix_amd_face_finish(ctx, 1, 12, 1, 0, 0, 3, p, q, q, 1.0, None) == -1
        assert lib.qphix_amd_face_pack(ctx, 1, 0, 0, 0, 0, p, q, None) == -1   # isign
        assert lib.qphix_amd_face_pack(ctx, 1, 1, 0, 0, 0, p, p, None) == -1   # buf == in
        assert lib.qphix_amd_face_pack(ctx, 1, 1, 0, 0, 0, ctypes.c_void_p(p.value + 4), q, None) == -6
        assert lib.qphix_amd_face_finish(ctx, 1, 13, 1, 0, 0, 0, p, q, q, 1.0, None) == -1
        assert lib.qphix_amd_clov_face_finish(ctx, 1, 12, 1, 0, 0, 0, p, q, q, None, 1.0, None) == -1
        assert lib.qphix_amd_tm_clov_face_finish(ctx, 1, 12, 1, 0, 3, 1, p, p, q, q, 1.0, None) == -1
        assert lib.qphix_amd_tm_face_finish(ctx, 7, 12, 1, 0, 3, 1, p, q, q, 1.0, 0.1, 0.2, None) == -1
        # operators that need the whole lattice refuse a split context
        comm = (ctypes.c_int * 4)(0, 0, 0, 1)
        assert lib.qphix_amd_set_comm(ctx, comm) == 0
        op = ctypes.c_void_p()
        assert lib.qphix_amd_eo_op_create(ctypes.byref(op), ctx, p, p, 0.0, 1, 12, None) == -1
        assert lib.qphix_amd_dslash_force(ctx, 0, 1, 0, p, q, q, p, 1.0, 1.0, 1.0, 1.0, 1, None) == -1
    finally:
        lib.qphix_amd_ctx_destroy(ctx)


def test_traversal_choice(qamd):
    """Automatic work-group traversal: order 1 (XCDs sweep t together) when a
    t-slice holds more than 32768 sites per checkerboard or Lt <= Lz (and
    Lz % 8 == 0), else 3 (t-slabs swept one z-part at a time; Lt % 8 == 0);
    requested orders the geometry does not admit run as 0; bad arguments
    rejected."""
    lib = qamd.lib()
    cases = {  # dims: (automatic, effective order for requests 0..3)
        (32, 32, 32, 64): (3, [0, 1, 2, 3]),
        (48, 48, 48, 96): (1, [0, 1, 2, 3]),
        (64, 32, 32, 64): (3, [0, 1, 2, 3]),
        (32, 32, 32, 128): (3, [0, 1, 2, 3]),
        (32, 32, 32, 32): (1, [0, 1, 2, 3]),   # Lt <= Lz
        (32, 32, 32, 8): (1, [0, 1, 2, 3]),
        (24, 24, 24, 48): (0, [0, 0, 0, 0]),   # 12 x 24 = 288 sites per plane
        (48, 48, 44, 96): (0, [0, 0, 2, 3]),   # Lz % 8 != 0
        (40, 40, 40, 80): (0, [0, 0, 0, 0]),   # (Lx/2) Ly % 128 != 0
        (8, 8, 8, 8): (0, [0, 0, 0, 0]),
    }
    for dims, (auto, eff) in cases.items():
        ctx = ctypes.c_void_p()
        assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), *dims, -1.0) == 0
        try:
            assert lib.qphix_amd_traversal(ctx) == auto, dims
            for req, want in enumerate(eff):
                assert lib.qphix_amd_set_traversal(ctx, req) == 0
                assert lib.qphix_amd_traversal(ctx) == want, (dims, req)
            assert lib.qphix_amd_set_traversal(ctx, -1) == 0
            assert lib.qphix_amd_traversal(ctx) == auto
            assert lib.qphix_amd_set_traversal(ctx, 4) == -1
            assert lib.qphix_amd_set_traversal(ctx, -2) == -1
        finally:
            lib.qphix_amd_ctx_destroy(ctx)
    assert lib.qphix_amd_set_traversal(None, 0) == -1
    assert lib.qphix_amd_traversal(None) == -1
