"""ctypes binding of libqphix_amd.so (the C ABI in include/qphix_amd.h).

The library is built in-tree by ``build.build()``; importing this module
loads it and raises if it is missing -- there is no fallback path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqphix_amd.so")

F64, F32, F16, BF16 = 0, 1, 2, 3
FIELD_SPINOR, FIELD_GAUGE, FIELD_CLOVER, FIELD_FULLCLOVER = 0, 1, 2, 3

_c_int, _c_double, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_void_p


class CGParams(ctypes.Structure):
    """qphix_amd_cg_params"""
    _fields_ = [("rsd_target", ctypes.c_double), ("max_iters", ctypes.c_int),
                ("delta", ctypes.c_double), ("mixed", ctypes.c_int)]


class CGResult(ctypes.Structure):
    """qphix_amd_cg_result"""
    _fields_ = [("iters", ctypes.c_int), ("reliable_updates", ctypes.c_int),
                ("rsd_final", ctypes.c_double), ("seconds", ctypes.c_double),
                ("mv_apps", ctypes.c_long)]

class QPhiXLayout(ctypes.Structure):
    """qphix_amd_qphix_layout: storage and block geometry of QPhiX arrays."""
    _fields_ = [("ft", ctypes.c_int), ("veclen", ctypes.c_int), ("soalen", ctypes.c_int),
                ("compress12", ctypes.c_int), ("pad_xy", ctypes.c_int), ("pad_xyz", ctypes.c_int)]


_lay = ctypes.POINTER(QPhiXLayout)

# name -> (restype, argtypes); the order follows include/qphix_amd.h
SIGNATURES = {
    "qphix_amd_ctx_create": (_c_int, [ctypes.POINTER(_vp), _c_int, _c_int, _c_int, _c_int, _c_double]),
    "qphix_amd_ctx_destroy": (_c_int, [_vp]),
    "qphix_amd_set_tbc": (_c_int, [_vp, ctypes.POINTER(_c_int), ctypes.POINTER(_c_double)]),
    "qphix_amd_field_bytes": (ctypes.c_size_t, [_vp, _c_int, _c_int, _c_int]),
    "qphix_amd_sites_cb": (ctypes.c_long, [_vp]),
    "qphix_amd_error_string": (ctypes.c_char_p, [_c_int]),
    "qphix_amd_pack_spinor": (_c_int, [_vp, _vp, _c_int, _vp, _c_int, _vp]),
    "qphix_amd_unpack_spinor": (_c_int, [_vp, _vp, _vp, _c_int, _c_int, _vp]),
    "qphix_amd_pack_gauge": (_c_int, [_vp, _vp, _c_int, _c_int, _vp, _c_int, _vp]),
    "qphix_amd_pack_clover": (_c_int, [_vp, _vp, _c_int, _vp, _c_int, _vp]),
    "qphix_amd_pack_full_clover": (_c_int, [_vp, _vp, _c_int, _vp, _c_int, _vp]),
    "qphix_amd_qphix_field_bytes": (ctypes.c_size_t, [_vp, _lay, _c_int]),
    "qphix_amd_spinor_from_qphix": (_c_int, [_vp, _lay, _vp, _c_int, _vp, _c_int, _vp]),
    "qphix_amd_spinor_to_qphix": (_c_int, [_vp, _lay, _vp, _vp, _c_int, _c_int, _vp]),
    "qphix_amd_gauge_from_qphix": (_c_int, [_vp, _lay, _vp, _c_int, _c_int, _vp, _c_int, _vp]),
    "qphix_amd_clover_from_qphix": (_c_int, [_vp, _lay, _vp, _c_int, _vp, _c_int, _vp]),
    "qphix_amd_full_clover_from_qphix": (_c_int, [_vp, _lay, _vp, _c_int, _vp, _c_int, _vp]),
    "qphix_amd_qphix_dslash": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                        _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_qphix_dslash_achimbdpsi": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp,
                                                   _c_double, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_qphix_clov_dslash": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp,
                                             _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_qphix_clov_dslash_achimbdpsi": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                                        _vp, _vp, _c_double, _c_double, _c_double, _c_double,
                                                        _vp]),
    "qphix_amd_qphix_tm_dslash": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_double,
                                           _c_double, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_qphix_tm_dslash_achimbdpsi": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                                      _vp, _c_double, _c_double, _c_double, _c_double,
                                                      _c_double, _vp]),
    "qphix_amd_qphix_tm_clov_dslash": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp,
                                                _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_qphix_tm_clov_dslash_achimbdpsi": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                                           _vp, _vp, _vp, _c_double, _c_double, _c_double,
                                                           _c_double, _vp]),
    "qphix_amd_dslash": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                  _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_dslash_achimbdpsi": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                             _c_double, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_clov_dslash": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                       _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_clov_dslash_achimbdpsi": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                                  _vp, _c_double, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_tm_dslash": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                     _c_double, _c_double, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_tm_dslash_achimbdpsi": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                                _c_double, _c_double, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_tm_clov_dslash": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                          _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_tm_clov_dslash_achimbdpsi": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                                     _vp, _vp, _c_double, _c_double, _c_double, _c_double,
                                                     _vp]),
    "qphix_amd_ndtm_dslash": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                       _c_double, _c_double, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_ndtm_dslash_achimbdpsi": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                                  _c_double, _c_double, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_ndtm_clov_dslash": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp,
                                            _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_ndtm_clov_dslash_achimbdpsi": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                                       _vp, _vp, _c_double, _c_double, _c_double, _c_double,
                                                       _vp]),
    "qphix_amd_set_comm": (_c_int, [_vp, ctypes.POINTER(_c_int)]),
    "qphix_amd_set_traversal": (_c_int, [_vp, _c_int]),
    "qphix_amd_traversal": (_c_int, [_vp]),
    "qphix_amd_face_sites": (ctypes.c_long, [_vp, _c_int]),
    "qphix_amd_face_pack": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "qphix_amd_face_finish": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                       _c_double, _vp]),
    "qphix_amd_clov_face_finish": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                                            _vp, _vp, _c_double, _vp]),
    "qphix_amd_tm_face_finish": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                          _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_tm_clov_face_finish": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                                               _vp, _vp, _c_double, _vp]),
    "qphix_amd_qphix_face_pack": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp]),
    "qphix_amd_qphix_face_finish": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                             _c_double, _vp]),
    "qphix_amd_qphix_clov_face_finish": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                                                  _vp, _vp, _c_double, _vp]),
    "qphix_amd_qphix_tm_face_finish": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _vp,
                                                _vp, _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_qphix_tm_clov_face_finish": (_c_int, [_vp, _lay, _c_int, _c_int, _c_int, _c_int, _c_int, _vp,
                                                     _vp, _vp, _vp, _c_double, _vp]),
    "qphix_amd_dslash_adjoint": (_c_int, [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _vp, _vp,
                                          _c_double, _c_double, _c_double, _vp]),
    "qphix_amd_dslash_force": (_c_int, [_vp, _c_int, _c_int, _c_int, _vp, _vp, _vp, _vp, _c_double,
                                        _c_double, _c_double, _c_double, _c_int, _vp]),
    "qphix_amd_reunitarize": (_c_int, [_vp, _vp, _vp]),
    "qphix_amd_clover_from_links": (_c_int, [_vp, _vp, _vp, _c_double, _c_double, _vp]),
    "qphix_amd_clover_invert": (_c_int, [_vp, _vp, _vp, _vp]),
    "qphix_amd_tmclover_blocks": (_c_int, [_vp, _vp, _vp, _c_double, _c_int, _c_int, _vp]),
    "qphix_amd_eo_op_create": (_c_int, [ctypes.POINTER(_vp), _vp, _vp, _vp, _c_double, _c_int, _c_int, _vp]),
    "qphix_amd_eo_op_destroy": (_c_int, [_vp]),
    "qphix_amd_eo_op_apply": (_c_int, [_vp, _c_int, _c_int, _vp, _vp, _vp]),
    "qphix_amd_eo_op_force": (_c_int, [_vp, _vp, _vp, _vp, _c_double, _c_int, _vp]),
    "qphix_amd_eo_op_force_clover": (_c_int, [_vp, _vp, _vp, _vp, _c_double, _c_double, _c_int, _vp]),
    "qphix_amd_clover_force": (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_double, _c_double, _c_double, _c_int, _vp]),
    "qphix_amd_clover_force_noalloc": (_c_int, [_vp, _vp, _vp, _vp, _vp, _c_double, _c_double, _c_double, _c_int,
                                                _vp]),
    "qphix_amd_cg_solve": (_c_int, [_vp, _vp, _vp, ctypes.POINTER(CGParams), ctypes.POINTER(CGResult), _vp]),
}


class QPhiXAMDError(RuntimeError):
    """A non-zero status returned by the C ABI."""


_lib = None


def lib():
    """Load the in-tree shared library (once); raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise QPhiXAMDError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(there is no CPU fallback)")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def check(status, what=""):
    if status != 0:
        msg = lib().qphix_amd_error_string(status).decode()
        raise QPhiXAMDError(f"{what}: {msg} (status {status})")
    return status
