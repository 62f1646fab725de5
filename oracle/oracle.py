"""ctypes wrapper of the FP64 C oracle (oracle/qphix_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  Product code never imports it.

All arrays are float64 numpy arrays in the lexicographic layouts described
in qphix_oracle.c (spinor (V,4,3,2), gauge (V,4,3,3,2), clover (V,2,36),
full clover (V,2,6,6,2)).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# QAMD_ORACLE_SO: another build of the same source (the sanitizer build,
# oracle/Makefile target liboracle_san.so)
_SO = os.environ.get("QAMD_ORACLE_SO") or os.path.join(_HERE, "liboracle.so")

WILSON, CLOVER, TM, TMCLOVER = 0, 1, 2, 3

_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    r = subprocess.run(["make", "-B", "-C", _HERE, "liboracle.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        l = ctypes.CDLL(_SO)
        i, d = ctypes.c_int, ctypes.c_double
        l.oracle_hop.argtypes = [_dp, _dp, _dp, i, i, i, i, i, i, d, d, d, d, _ip, _dp]
        l.oracle_dslash.argtypes = [_dp, _dp, _dp, _dp, _dp, i, i, i, i, i, i, i, i, d, d, d, d, _ip,
                                    _dp, d, d, d, d]
        l.oracle_dslash_sites.argtypes = [_dp, ctypes.POINTER(ctypes.c_long), ctypes.c_long, _dp, _dp,
                                          _dp, _dp, i, i, i, i, i, i, i, d, d, d, d, _ip, _dp, d, d,
                                          d, d]
        l.oracle_force.argtypes = [_dp, _dp, _dp, _dp, i, i, i, i, i, i, d, d, d, d, _ip, _dp, i]
        l.oracle_force_links.argtypes = [_dp, ctypes.POINTER(ctypes.c_long), ctypes.c_long, _dp, _dp, _dp,
                                         i, i, i, i, i, i, d, d, d, d, _ip, _dp, i]
        l.oracle_clover_from_links_sites.argtypes = [_dp, ctypes.POINTER(ctypes.c_long), ctypes.c_long,
                                                     _dp, i, i, i, i, d, d]
        l.oracle_clover_apply.argtypes = [_dp, _dp, _dp, i, i, i, i, i, i]
        l.oracle_clover_from_links.argtypes = [_dp, _dp, i, i, i, i, d, d]
        l.oracle_clover_invert.argtypes = [_dp, _dp, ctypes.c_long]
        l.oracle_clover_invert.restype = i
        l.oracle_tmclover_blocks.argtypes = [_dp, _dp, ctypes.c_long, d, i, i]
        l.oracle_tmclover_blocks.restype = i
        _lib = l
    return _lib


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _bc(tbc, phases):
    if tbc is None:
        return None, None
    t = (ctypes.c_int * 4)(*[int(bool(x)) for x in tbc])
    p = (ctypes.c_double * 8)()
    for m, z in enumerate(phases):
        p[2 * m], p[2 * m + 1] = complex(z).real, complex(z).imag
    return t, p


def dslash(dims, psi, U, isign, cb, kind=WILSON, achi=False, chi=None, clov=None,
           coeff_s=1.0, coeff_t_f=1.0, coeff_t_b=1.0, t_boundary=-1.0, tbc=None, tbc_phases=None,
           alpha=0.0, mu=0.0, muinv=0.0, rho=0.0, out=None):
    """Every operator family; returns a (V,4,3,2) array with the cb sites set
    (the other checkerboard is zero unless `out` is given)."""
    Lx, Ly, Lz, Lt = dims
    V = Lx * Ly * Lz * Lt
    psi, pp = _c(psi)
    U, pu = _c(U)
    chi_a, pc = _c(chi) if chi is not None else (None, None)
    clov_a, pcl = _c(clov) if clov is not None else (None, None)
    if out is None:
        out = np.zeros((V, 4, 3, 2))
    out, po = _c(out)
    t, p = _bc(tbc, tbc_phases)
    lib().oracle_dslash(po, pp, pc, pu, pcl, kind, int(achi), Lx, Ly, Lz, Lt, isign, cb, coeff_s,
                        coeff_t_f, coeff_t_b, t_boundary, t, p, alpha, mu, muinv, rho)
    return out


def dslash_sites(dims, sites, psi, U, isign, kind=WILSON, achi=False, chi=None, clov=None,
                 coeff_s=1.0, coeff_t_f=1.0, coeff_t_b=1.0, t_boundary=-1.0, tbc=None,
                 tbc_phases=None, alpha=0.0, mu=0.0, muinv=0.0, rho=0.0):
    """The operator at the listed lexicographic sites only -> (n,4,3,2).
    Inputs are full-lattice arrays (any checkerboard content)."""
    Lx, Ly, Lz, Lt = dims
    sites = np.ascontiguousarray(sites, dtype=np.int64)
    psi, pp = _c(psi)
    U, pu = _c(U)
    chi_a, pc = _c(chi) if chi is not None else (None, None)
    clov_a, pcl = _c(clov) if clov is not None else (None, None)
    out = np.zeros((len(sites), 4, 3, 2))
    out, po = _c(out)
    t, p = _bc(tbc, tbc_phases)
    lib().oracle_dslash_sites(po, sites.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), len(sites),
                              pp, pc, pu, pcl, kind, int(achi), Lx, Ly, Lz, Lt, isign, coeff_s,
                              coeff_t_f, coeff_t_b, t_boundary, t, p, alpha, mu, muinv, rho)
    return out


def force(dims, chi, psi, U, isign, cb, coeff_s=1.0, coeff_t_f=1.0, coeff_t_b=1.0, t_boundary=-1.0,
          tbc=None, tbc_phases=None, project=True):
    """Link gradient of L = Re<chi, hop psi> (chi read on cb, psi on 1-cb);
    (V,4,3,3,2): TA(U G) if project else G^dagger."""
    Lx, Ly, Lz, Lt = dims
    V = Lx * Ly * Lz * Lt
    chi, pc = _c(chi)
    psi, pp = _c(psi)
    U, pu = _c(U)
    out = np.zeros((V, 4, 3, 3, 2))
    out, po = _c(out)
    t, p = _bc(tbc, tbc_phases)
    lib().oracle_force(po, pc, pp, pu, Lx, Ly, Lz, Lt, isign, cb, coeff_s, coeff_t_f, coeff_t_b,
                       t_boundary, t, p, int(bool(project)))
    return out


def force_links(dims, links, chi, psi, U, isign, cb, coeff_s=1.0, coeff_t_f=1.0, coeff_t_b=1.0,
                t_boundary=-1.0, tbc=None, tbc_phases=None, project=True):
    """force() at the listed links only (link = 4 * lexicographic site + mu)
    -> (n,3,3,2).  Inputs are full-lattice arrays; only the sites the listed
    links' hops touch are read."""
    Lx, Ly, Lz, Lt = dims
    links = np.ascontiguousarray(links, dtype=np.int64)
    chi, pc = _c(chi)
    psi, pp = _c(psi)
    U, pu = _c(U)
    out = np.zeros((len(links), 3, 3, 2))
    out, po = _c(out)
    t, p = _bc(tbc, tbc_phases)
    lib().oracle_force_links(po, links.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), len(links), pc, pp,
                             pu, Lx, Ly, Lz, Lt, isign, cb, coeff_s, coeff_t_f, coeff_t_b, t_boundary, t,
                             p, int(bool(project)))
    return out


def clover_from_links_sites(dims, sites, U, alpha, csw):
    """clover_from_links() at the listed lexicographic sites only -> (n,2,36)."""
    Lx, Ly, Lz, Lt = dims
    sites = np.ascontiguousarray(sites, dtype=np.int64)
    U, pu = _c(U)
    out = np.zeros((len(sites), 2, 36))
    out, po = _c(out)
    lib().oracle_clover_from_links_sites(po, sites.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                         len(sites), pu, Lx, Ly, Lz, Lt, alpha, csw)
    return out


def clover_apply(dims, clov, psi, cb, full=False):
    Lx, Ly, Lz, Lt = dims
    V = Lx * Ly * Lz * Lt
    clov, pc = _c(clov)
    psi, pp = _c(psi)
    out = np.zeros((V, 4, 3, 2))
    out, po = _c(out)
    lib().oracle_clover_apply(po, pp, pc, int(full), Lx, Ly, Lz, Lt, cb)
    return out


def clover_from_links(dims, U, alpha, csw):
    Lx, Ly, Lz, Lt = dims
    V = Lx * Ly * Lz * Lt
    U, pu = _c(U)
    out = np.zeros((V, 2, 36))
    out, po = _c(out)
    lib().oracle_clover_from_links(po, pu, Lx, Ly, Lz, Lt, alpha, csw)
    return out


def clover_invert(clov):
    clov, pc = _c(clov)
    n = clov.size // 72
    out = np.zeros_like(clov)
    out, po = _c(out)
    if lib().oracle_clover_invert(po, pc, n):
        raise ZeroDivisionError("singular clover block")
    return out


def tmclover_blocks(clov, mu, sign, invert):
    clov, pc = _c(clov)
    n = clov.size // 72
    out = np.zeros((n, 2, 6, 6, 2))
    out, po = _c(out)
    if lib().oracle_tmclover_blocks(po, pc, n, mu, sign, int(invert)):
        raise ZeroDivisionError("singular clover block")
    return out
