"""ctypes wrapper of oracle/_ref/libref_<config>.so: the reference's own
generated kernel bodies (qphix-codegen) under a QPhiX-style block driver
(oracle/ref_driver.cpp).  Built from /root/reference by oracle/ref.mk.

TEST INFRASTRUCTURE ONLY: the tests use it to pin the FP64 oracle and write
golden fixtures, bench.py times it as the CPU baseline.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
CONFIGS = ("scalar_f64", "avx512_f64", "avx512_f32", "avx2_f64", "avx2_f32", "avx512_f16",
           "avx512_f64_s4", "avx512_f32_s8")
_libs = {}
_dp = ctypes.POINTER(ctypes.c_double)


def path(config):
    return os.path.join(_HERE, "_ref", f"libref_{config}.so")


def available(config):
    return os.path.exists(path(config))


def lib(config):
    if config not in _libs:
        l = ctypes.CDLL(path(config))
        i, d = ctypes.c_int, ctypes.c_double
        l.ref_apply.argtypes = [i, i, i, i, i, i, i, i, i, _dp, _dp, _dp, _dp, _dp, d, d, d, d, d,
                                d, d, d, i, i, _dp, _dp, ctypes.POINTER(ctypes.c_ubyte)]
        l.ref_apply.restype = i
        l.ref_veclen.restype = i
        l.ref_precision.restype = i
        l.ref_lowprec.restype = i
        l.ref_pack_field.argtypes = [i, i, i, i, i, i, i, _dp, ctypes.c_void_p]
        l.ref_pack_field.restype = ctypes.c_long
        l.ref_soalen.restype = i
        l.ref_face_pack.argtypes = [i, i, i, i, i, i, i, _dp, _dp]
        l.ref_face_pack.restype = i
        l.ref_face_finish.argtypes = [i, i, i, i, i, i, i, i, i, _dp, _dp, _dp, _dp, d, d, d, _dp]
        l.ref_face_finish.restype = i
        l.ref_stream_triad.argtypes = [ctypes.c_long, i, i]
        l.ref_stream_triad.restype = d
        _libs[config] = l
    return _libs[config]


def soalen(config):
    return lib(config).ref_soalen()


def veclen(config):
    return lib(config).ref_veclen()


def stream_triad(config, nthreads, n=1 << 25, reps=5):
    """Host STREAM triad (GB/s) on nthreads OpenMP threads of the same
    runtime the reference's kernels run under (three FP64 arrays of n)."""
    return lib(config).ref_stream_triad(int(n), int(nthreads), int(reps))


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def apply(config, dims, psi, U, isign, cb, kind=0, achi=False, chi=None, clov=None, recon=18,
          coeff_s=1.0, coeff_t_f=1.0, coeff_t_b=1.0, t_boundary=-1.0, alpha=0.0, mu=0.0,
          muinv=0.0, rho=0.0, nthreads=1, reps=0, tbc_phases=None, mask=None):
    """Run one reference kernel over checkerboard cb.  Returns (out, times):
    out is (V,4,3,2) with the cb sites set; with reps > 0 the block loop
    (not the packing) runs reps more times and times holds each call's
    seconds (a numpy array), else None.
    tbc_phases: four complex forward-hop phases (x, y, z, t) -> the bodies
    generated with twisted boundary conditions in every dimension.
    mask: optional uint8 array of one byte per lexicographic site, read at
    the output sites: bit d = 2 mu + fb enables direction d there (the
    driver turns it into the bodies' per-lane accumulate[d] masks)."""
    Lx, Ly, Lz, Lt = dims
    V = Lx * Ly * Lz * Lt
    c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    psi, U, chi, clov = c(psi), c(U), c(chi), c(clov)
    out = np.zeros((V, 4, 3, 2))
    sec = np.zeros(max(int(reps), 1))
    ph = None
    if tbc_phases is not None:
        ph = np.array([v for z in tbc_phases for v in (complex(z).real, complex(z).imag)], dtype=np.float64)
    pm = None
    if mask is not None:
        mask = np.ascontiguousarray(mask, dtype=np.uint8)
        if mask.size != V:
            raise ValueError("mask needs one byte per lattice site")
        pm = mask.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))
    rc = lib(config).ref_apply(kind, int(achi), isign, recon, cb, Lx, Ly, Lz, Lt, _p(psi), _p(chi),
                               _p(U), _p(clov), _p(out), coeff_s, coeff_t_f, coeff_t_b, t_boundary,
                               alpha, mu, muinv, rho, nthreads, reps, _p(sec), _p(ph), pm)
    if rc != 0:
        raise ValueError(f"reference kernel rejected the call (config {config}, dims {dims})")
    return out, (sec if reps > 0 else None)


def pack(config, field, dims, src, cb, recon=18):
    """The block array (flat, FT = float32 or float64) that ref_apply hands
    to the reference's bodies for a lexicographic FP64 field: field 0 spinor,
    1 gauge (recon 12 | 18), 2 clover, 3 full clover."""
    l = lib(config)
    src = np.ascontiguousarray(src, dtype=np.float64)
    n = l.ref_pack_field(field, recon, cb, *dims, _p(src), None)
    if n < 0:
        raise ValueError(f"unsupported geometry {dims} for {config}")
    if l.ref_lowprec():
        dt = np.float16  # IEEE binary16 bits, the reference's `half`
    else:
        dt = np.float32 if l.ref_precision() == 1 else np.float64
    out = np.empty(n, dtype=dt)
    l.ref_pack_field(field, recon, cb, *dims, _p(src), out.ctypes.data_as(ctypes.c_void_p))
    return out


def _phases(tbc_phases):
    if tbc_phases is None:
        return None
    return np.array([v for z in tbc_phases for v in (complex(z).real, complex(z).imag)], dtype=np.float64)


def face_pack(config, dims, psi, isign, dir, pcb):
    """The reference's face pack body (face_proj_dir_{plus,minus}) over face
    dir = 2 mu + fb of checkerboard pcb: returns hs (V, 12), the projected
    half-spinor ([spin 0|1][colour][re|im]) at each face site, zero
    elsewhere."""
    V = int(np.prod(dims))
    psi = np.ascontiguousarray(psi, dtype=np.float64)
    hs = np.zeros((V, 12))
    rc = lib(config).ref_face_pack(isign, dir, pcb, *dims, _p(psi), _p(hs))
    if rc != 0:
        raise ValueError(f"reference face pack rejected the call (config {config}, dims {dims})")
    return hs


def face_finish(config, dims, kind, isign, recon, dir, cb, hs, U, out, beta, clov=None, mu=0.0, muinv=0.0,
                tbc_phases=None):
    """The reference's face finish body of family kind (0 face_finish_dir,
    1 face_clov_finish_dir, 2 tm_face_finish_dir, 3 tm_clov_face_finish_dir)
    on face dir of checkerboard cb: the face sites receive hs (V, 12, read
    at those sites) and update a copy of out (V, 4, 3, 2), which is
    returned."""
    c = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    hs, U, clov = c(hs), c(U), c(clov)
    res = np.array(out, dtype=np.float64, copy=True, order="C")
    rc = lib(config).ref_face_finish(kind, isign, recon, dir, cb, *dims, _p(hs), _p(U), _p(clov), _p(res),
                                     beta, mu, muinv, _p(_phases(tbc_phases)))
    if rc != 0:
        raise ValueError(f"reference face finish rejected the call (config {config}, dims {dims})")
    return res
