"""Host-side operator classes over the C ABI, mirroring QPhiX's caller-side
interface to the reference's generated kernels.

QPhiX's Dslash / ClovDslash / TMDslash / TMClovDslash classes drive the
generated bodies (codegen/jinja/*_decl.h.j2) block by block; here each
method is one whole-checkerboard call into libqphix_amd.so.  Method names,
argument order and meaning follow those classes:

  Dslash.dslash(res, psi, u, isign, cb)                      res = D psi
  Dslash.dslashAChiMinusBDPsi(res, psi, chi, u, alpha, beta, isign, cb)
                                                   res = alpha chi - beta D psi
  ClovDslash.dslash(res, psi, u, invclov, isign, cb)         res = A^-1 D psi
  ClovDslash.dslashAChiMinusBDPsi(res, psi, chi, u, clov, beta, isign, cb)
                                        res = A chi (+ i rho g5 chi) - beta D psi
  TMDslash / TMClovDslash likewise, with the twisted mass.

The hopping term D carries the anisotropy coefficients (coeff_s for x, y, z,
coeff_t for t) and the geometry's t_boundary on hops across the time
boundary (codegen/dslash.cc:79-86).  Fields are device tensors in the
library's native layout (``Geometry.pack_*``).  Errors raise QPhiXAMDError.
"""
import ctypes

import torch

from ._lib import (BF16, F16, F32, F64, FIELD_CLOVER, FIELD_FULLCLOVER, FIELD_GAUGE,
                   FIELD_SPINOR, QPhiXAMDError, QPhiXLayout, check, lib)

TORCH_DTYPE = {F64: torch.float64, F32: torch.float32, F16: torch.float16, BF16: torch.bfloat16}
PREC_NAME = {F64: "f64", F32: "f32", F16: "f16", BF16: "bf16"}


def _ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Field:
    """A native-layout field on one checkerboard: a flat device tensor plus
    its storage precision (and link reconstruction for gauge fields)."""

    def __init__(self, data, prec, kind, recon=0, cb=None):
        self.data = data
        self.prec = prec
        self.kind = kind
        self.recon = recon
        self.cb = cb

    def ptr(self):
        return ctypes.c_void_p(self.data.data_ptr())


class Geometry:
    """A lattice Lx x Ly x Lz x Lt (all even) with its boundary conditions."""

    def __init__(self, dims, t_boundary=-1.0):
        self.dims = tuple(int(d) for d in dims)
        self.t_boundary = float(t_boundary)
        ctx = ctypes.c_void_p()
        check(lib().qphix_amd_ctx_create(ctypes.byref(ctx), *self.dims, self.t_boundary),
              "qphix_amd_ctx_create")
        self._ctx = ctx
        self.vcb = int(lib().qphix_amd_sites_cb(ctx))

    @property
    def ctx(self):
        return self._ctx

    def close(self):
        if getattr(self, "_ctx", None):
            lib().qphix_amd_ctx_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tbc(self, tbc, phases):
        """tbc: 4 flags; phases: 4 complex numbers (forward-hop phase per dim)."""
        t = (ctypes.c_int * 4)(*[int(bool(b)) for b in tbc])
        p = (ctypes.c_double * 8)()
        for m, z in enumerate(phases):
            p[2 * m], p[2 * m + 1] = complex(z).real, complex(z).imag
        check(lib().qphix_amd_set_tbc(self._ctx, t, p), "qphix_amd_set_tbc")

    def set_comm(self, comm):
        """comm: 4 flags, split dimensions (their wrap-around hops belong to
        the faces, see face_pack / face_finish)."""
        c = (ctypes.c_int * 4)(*[int(bool(b)) for b in comm])
        check(lib().qphix_amd_set_comm(self._ctx, c), "qphix_amd_set_comm")

    def face_sites(self, dim):
        return int(lib().qphix_amd_face_sites(self._ctx, dim))

    def set_traversal(self, order):
        """Work-group traversal of the operator kernels: -1 automatic (1 for
        t-slices above 32768 sites per checkerboard or Lt <= Lz, else 3), 0 a
        t-slab per XCD, 1 XCDs sweep t together owning z-ranges, 2 t-slabs
        z-major, 3 t-slabs in two z-halves (bit-identical results)."""
        check(lib().qphix_amd_set_traversal(self._ctx, int(order)), "qphix_amd_set_traversal")

    @property
    def traversal(self):
        """The traversal order the operator calls on this geometry use."""
        return int(lib().qphix_amd_traversal(self._ctx))

    def new_face(self, prec, dim, device="cuda", both=False):
        """A face buffer of dim (both: the two faces, for fb = 2)."""
        n = self.face_sites(dim) * 12 * (2 if both else 1)
        return torch.empty(n, dtype=TORCH_DTYPE[prec], device=device)

    # ---- allocation ---------------------------------------------------
    def field_bytes(self, kind, prec, recon=18):
        return int(lib().qphix_amd_field_bytes(self._ctx, kind, prec, recon))

    def new_field(self, kind, prec, recon=18, device="cuda", cb=None):
        nbytes = self.field_bytes(kind, prec, recon)
        if nbytes == 0:
            raise QPhiXAMDError("invalid field request")
        dt = TORCH_DTYPE[prec]
        n = nbytes // torch.tensor([], dtype=dt).element_size()
        return Field(torch.empty(n, dtype=dt, device=device), prec, kind, recon, cb)

    def new_spinor(self, prec, device="cuda", cb=None):
        return self.new_field(FIELD_SPINOR, prec, device=device, cb=cb)

    # ---- converters (device lex FP64 <-> native) ------------------------
    @staticmethod
    def _lex(t):
        if t.dtype != torch.float64 or not t.is_cuda:
            raise QPhiXAMDError("lexicographic fields must be float64 device tensors")
        return t.contiguous()

    def pack_spinor(self, lex, prec, cb):
        lex = self._lex(lex)
        f = self.new_spinor(prec, device=lex.device, cb=cb)
        check(lib().qphix_amd_pack_spinor(self._ctx, f.ptr(), prec, _ptr(lex), cb, _stream()),
              "qphix_amd_pack_spinor")
        return f

    def unpack_spinor(self, f, out=None):
        """Native spinor -> lex FP64 (only the field's checkerboard is written)."""
        V = self.vcb * 2
        if out is None:
            out = torch.zeros(V, 4, 3, 2, dtype=torch.float64, device=f.data.device)
        check(lib().qphix_amd_unpack_spinor(self._ctx, _ptr(out), f.ptr(), f.prec, f.cb, _stream()),
              "qphix_amd_unpack_spinor")
        return out

    def pack_gauge(self, u_lex, prec, recon, cb):
        u_lex = self._lex(u_lex)
        f = self.new_field(FIELD_GAUGE, prec, recon, device=u_lex.device, cb=cb)
        check(lib().qphix_amd_pack_gauge(self._ctx, f.ptr(), prec, recon, _ptr(u_lex), cb, _stream()),
              "qphix_amd_pack_gauge")
        return f

    def pack_clover(self, clov_lex, prec, cb):
        clov_lex = self._lex(clov_lex)
        f = self.new_field(FIELD_CLOVER, prec, device=clov_lex.device, cb=cb)
        check(lib().qphix_amd_pack_clover(self._ctx, f.ptr(), prec, _ptr(clov_lex), cb, _stream()),
              "qphix_amd_pack_clover")
        return f

    def pack_full_clover(self, fclov_lex, prec, cb):
        fclov_lex = self._lex(fclov_lex)
        f = self.new_field(FIELD_FULLCLOVER, prec, device=fclov_lex.device, cb=cb)
        check(lib().qphix_amd_pack_full_clover(self._ctx, f.ptr(), prec, _ptr(fclov_lex), cb, _stream()),
              "qphix_amd_pack_full_clover")
        return f

    # ---- converters (device QPhiX block arrays <-> native) ---------------
    # A layout is a QPhiXLayout (ft, veclen, soalen, compress12, pad_xy,
    # pad_xyz); QPhiX arrays are flat device tensors of the layout's FT
    # (float64, float32, or float16 holding the reference's IEEE `half`).
    def qphix_field_bytes(self, layout, kind):
        n = int(lib().qphix_amd_qphix_field_bytes(self._ctx, ctypes.byref(layout), kind))
        if n == 0:
            raise QPhiXAMDError(f"layout {_layout_str(layout)} does not fit lattice {self.dims}")
        return n

    def new_qphix_field(self, layout, kind, device="cuda"):
        dt = TORCH_DTYPE[layout.ft]
        n = self.qphix_field_bytes(layout, kind) // torch.tensor([], dtype=dt).element_size()
        return torch.zeros(n, dtype=dt, device=device)

    def _qphix_in(self, arr, layout, kind):
        if arr.dtype != TORCH_DTYPE[layout.ft] or not arr.is_cuda:
            raise QPhiXAMDError("QPhiX arrays are device tensors of the layout's FT")
        if arr.numel() * arr.element_size() < self.qphix_field_bytes(layout, kind):
            raise QPhiXAMDError("QPhiX array smaller than the layout requires")
        return arr.contiguous()

    def spinor_from_qphix(self, arr, layout, prec, cb):
        arr = self._qphix_in(arr, layout, FIELD_SPINOR)
        f = self.new_spinor(prec, device=arr.device, cb=cb)
        check(lib().qphix_amd_spinor_from_qphix(self._ctx, ctypes.byref(layout), f.ptr(), prec, _ptr(arr),
                                                cb, _stream()), "qphix_amd_spinor_from_qphix")
        return f

    def spinor_to_qphix(self, f, layout, out=None):
        """Native spinor -> QPhiX FourSpinorBlocks (padding blocks untouched)."""
        if out is None:
            out = self.new_qphix_field(layout, FIELD_SPINOR, device=f.data.device)
        out = self._qphix_in(out, layout, FIELD_SPINOR)
        check(lib().qphix_amd_spinor_to_qphix(self._ctx, ctypes.byref(layout), _ptr(out), f.ptr(), f.prec,
                                              f.cb, _stream()), "qphix_amd_spinor_to_qphix")
        return out

    def gauge_from_qphix(self, arr, layout, prec, recon, cb):
        arr = self._qphix_in(arr, layout, FIELD_GAUGE)
        f = self.new_field(FIELD_GAUGE, prec, recon, device=arr.device, cb=cb)
        check(lib().qphix_amd_gauge_from_qphix(self._ctx, ctypes.byref(layout), f.ptr(), prec, recon,
                                               _ptr(arr), cb, _stream()), "qphix_amd_gauge_from_qphix")
        return f

    def clover_from_qphix(self, arr, layout, prec, cb):
        arr = self._qphix_in(arr, layout, FIELD_CLOVER)
        f = self.new_field(FIELD_CLOVER, prec, device=arr.device, cb=cb)
        check(lib().qphix_amd_clover_from_qphix(self._ctx, ctypes.byref(layout), f.ptr(), prec, _ptr(arr),
                                                cb, _stream()), "qphix_amd_clover_from_qphix")
        return f

    def full_clover_from_qphix(self, arr, layout, prec, cb):
        arr = self._qphix_in(arr, layout, FIELD_FULLCLOVER)
        f = self.new_field(FIELD_FULLCLOVER, prec, device=arr.device, cb=cb)
        check(lib().qphix_amd_full_clover_from_qphix(self._ctx, ctypes.byref(layout), f.ptr(), prec,
                                                     _ptr(arr), cb, _stream()),
              "qphix_amd_full_clover_from_qphix")
        return f


def _qphix_spinor_ok(geom, f):
    n = geom.qphix_field_bytes(f.layout, FIELD_SPINOR) // f.data.element_size()
    if f.data.numel() < n:
        raise QPhiXAMDError("QPhiX array smaller than the layout requires")


def face_pack(geom, buf, psi, isign, cb, dim, fb):
    """buf <- projected half-spinors of psi (checkerboard 1-cb) on the slice
    x_dim = 0 (fb 0, sent backwards) or L_dim - 1 (fb 1, sent forwards), or
    both in one launch (fb 2, buf holding face 0 then face 1); psi a native
    spinor or a QPhiXSpinor (read in place)."""
    if buf.dtype != TORCH_DTYPE[psi.prec] or buf.numel() < geom.face_sites(dim) * 12 * (2 if fb == 2 else 1):
        raise QPhiXAMDError("face buffer of the wrong type or size")
    if isinstance(psi, QPhiXSpinor):
        _qphix_spinor_ok(geom, psi)
        st = lib().qphix_amd_qphix_face_pack(geom.ctx, ctypes.byref(psi.layout), isign, cb, dim, fb, _ptr(buf),
                                             psi.ptr(), _stream())
    else:
        st = lib().qphix_amd_face_pack(geom.ctx, psi.prec, isign, cb, dim, fb, _ptr(buf), psi.ptr(), _stream())
    check(st, "face pack")
    return buf


def face_finish(geom, kind, out, buf, u, isign, cb, dim, fb, beta, clov=None, mu=0.0, muinv=0.0):
    """Complete the boundary sites of out (checkerboard cb) from a received
    face (fb 0 or 1; fb 2: both in one launch, buf holding face 0 then face
    1); kind 0 Wilson (out -= r), 1 clover (out += clov r), 2 twisted mass
    (out -= muinv (1 - isign i mu g5) r), 3 twisted-mass clover (out += clov r)
    with r the reconstructed beta * U h."""
    if buf.dtype != TORCH_DTYPE[out.prec] or buf.numel() < geom.face_sites(dim) * 12 * (2 if fb == 2 else 1):
        raise QPhiXAMDError("face buffer of the wrong type or size")
    if u.prec != out.prec or (clov is not None and clov.prec != out.prec):
        raise QPhiXAMDError("out, u and clov must share a storage precision")
    if kind in (1, 3) and clov is None:
        raise QPhiXAMDError("the clover finishes need clov")
    l = lib()
    if isinstance(out, QPhiXSpinor):  # the output array updated in place
        _qphix_spinor_ok(geom, out)
        a, pre = (geom.ctx, ctypes.byref(out.layout), u.recon, isign, cb, dim, fb, out.ptr(), _ptr(buf),
                  u.ptr()), "qphix_amd_qphix_"
    else:
        a, pre = (geom.ctx, out.prec, u.recon, isign, cb, dim, fb, out.ptr(), _ptr(buf), u.ptr()), "qphix_amd_"
    if kind == 0:
        st = getattr(l, pre + "face_finish")(*a, float(beta), _stream())
    elif kind == 1:
        st = getattr(l, pre + "clov_face_finish")(*a, clov.ptr(), float(beta), _stream())
    elif kind == 2:
        st = getattr(l, pre + "tm_face_finish")(*a, float(beta), float(mu), float(muinv), _stream())
    else:
        st = getattr(l, pre + "tm_clov_face_finish")(*a, clov.ptr(), float(beta), _stream())
    check(st, "face finish")
    return out


def qphix_layout(ft, veclen, soalen, compress12=True, pad_xy=0, pad_xyz=0):
    """QPhiXLayout for arrays of storage ft (F64 / F32 / F16)."""
    return QPhiXLayout(int(ft), int(veclen), int(soalen), int(bool(compress12)), int(pad_xy), int(pad_xyz))


def _layout_str(l):
    return (f"(ft={PREC_NAME.get(l.ft, l.ft)}, veclen={l.veclen}, soalen={l.soalen}, "
            f"compress12={l.compress12}, pad_xy={l.pad_xy}, pad_xyz={l.pad_xyz})")


class QPhiXSpinor:
    """A QPhiX FourSpinorBlock array used in place: a flat device tensor of
    the layout's FT (float64 / float32 / float16) in the block order of the
    layout (qphix_layout).  The operator methods below accept it wherever
    they take a native spinor (res, psi and chi all QPhiX arrays of one
    layout, the links and clover native fields of the same precision), and
    then call the qphix_amd_qphix_* entry points, which read and write the
    arrays where they are.  accumulate: optional uint8 device tensor, one
    byte per site slot of the array (block * soalen + lane), bit d = 2 mu +
    fb enabling direction d at that output site -- the reference's
    accumulate[8] lane masks (codegen/dslash.cc:248,285-290); None enables
    every direction."""

    def __init__(self, data, layout, accumulate=None, cb=None):
        if data.dtype != TORCH_DTYPE.get(layout.ft) or not data.is_cuda or not data.is_contiguous():
            raise QPhiXAMDError("a QPhiX array is a contiguous device tensor of the layout's FT")
        if accumulate is not None and (accumulate.dtype != torch.uint8 or not accumulate.is_cuda):
            raise QPhiXAMDError("accumulate is a uint8 device tensor")
        self.data, self.layout, self.accumulate, self.cb = data, layout, accumulate, cb

    @property
    def prec(self):
        return self.layout.ft

    def ptr(self):
        return ctypes.c_void_p(self.data.data_ptr())


def _same_layout(a, b):
    return all(getattr(a, f) == getattr(b, f) for f, _ in QPhiXLayout._fields_)


def _check_args(res, psi, u, isign, cb, geom=None, chi=None):
    if isign not in (1, -1) or cb not in (0, 1):
        raise QPhiXAMDError("isign must be +-1 and cb 0 or 1")
    if not (res.prec == psi.prec == u.prec):
        raise QPhiXAMDError("res, psi and u must share a storage precision")
    if u.kind != FIELD_GAUGE:
        raise QPhiXAMDError("u must be a packed gauge field")
    if u.cb is not None and u.cb != cb:
        raise QPhiXAMDError(f"gauge packed for cb {u.cb}, operator applied on cb {cb}")
    spinors = [res, psi] + ([chi] if chi is not None else [])
    if isinstance(res, QPhiXSpinor):
        if not all(isinstance(f, QPhiXSpinor) and _same_layout(f.layout, res.layout) for f in spinors):
            raise QPhiXAMDError("res, psi and chi must be QPhiX arrays of one layout")
        n = geom.qphix_field_bytes(res.layout, FIELD_SPINOR) // res.data.element_size()
        if any(f.data.numel() < n for f in spinors):
            raise QPhiXAMDError("QPhiX array smaller than the layout requires")
        if res.accumulate is not None and res.accumulate.numel() < n // 24:
            raise QPhiXAMDError("accumulate needs one byte per site slot")
    elif any(isinstance(f, QPhiXSpinor) for f in spinors):
        raise QPhiXAMDError("res, psi and chi must be all native or all QPhiX arrays")
    res.cb = cb


def _run(geom, name, res, psi, chi, u, clov, isign, cb, scalars):
    """One operator call: the native entry qphix_amd_<name>, or, for QPhiX
    arrays, qphix_amd_qphix_<name> on them in place."""
    _check_args(res, psi, u, isign, cb, geom, chi)
    q = isinstance(res, QPhiXSpinor)
    args = [geom.ctx, ctypes.byref(res.layout) if q else res.prec, u.recon, isign, cb, res.ptr(), psi.ptr()]
    if chi is not None:
        args.append(chi.ptr())
    args.append(u.ptr())
    if clov is not None:
        args.append(clov.ptr())
    if q:
        args.append(_ptr(res.accumulate))
    fn = ("qphix_amd_qphix_" if q else "qphix_amd_") + name
    check(getattr(lib(), fn)(*args, *[float(v) for v in scalars], _stream()), fn)


class Dslash:
    """Wilson hopping term (QPhiX Dslash)."""

    def __init__(self, geom, aniso_coeff_s=1.0, aniso_coeff_t=1.0):
        self.geom = geom
        self.aniso_s = float(aniso_coeff_s)
        self.aniso_t = float(aniso_coeff_t)

    def dslash(self, res, psi, u, isign, cb):
        _run(self.geom, "dslash", res, psi, None, u, None, isign, cb,
             (self.aniso_s, self.aniso_t, self.aniso_t))

    def dslashAChiMinusBDPsi(self, res, psi, chi, u, alpha, beta, isign, cb):
        b = float(beta)
        _run(self.geom, "dslash_achimbdpsi", res, psi, chi, u, None, isign, cb,
             (alpha, b * self.aniso_s, b * self.aniso_t, b * self.aniso_t))


class ClovDslash(Dslash):
    """Wilson-clover (QPhiX ClovDslash): A^-1 applied after the hop."""

    def dslash(self, res, psi, u, invclov, isign, cb):
        _run(self.geom, "clov_dslash", res, psi, None, u, invclov, isign, cb,
             (self.aniso_s, self.aniso_t, self.aniso_t))

    def dslashAChiMinusBDPsi(self, res, psi, chi, u, clov, beta, isign, cb, prec_mass_rho=0.0):
        b = float(beta)
        _run(self.geom, "clov_dslash_achimbdpsi", res, psi, chi, u, clov, isign, cb,
             (b * self.aniso_s, b * self.aniso_t, b * self.aniso_t, prec_mass_rho))


class TMDslash(Dslash):
    """Twisted-mass Wilson (QPhiX TMDslash).  alpha = 4 + mass; the kernels
    take mu = twisted_mass/alpha and muinv = alpha/(alpha^2 + twisted_mass^2)
    (codegen/dslash_common.cc:1256-1266)."""

    def __init__(self, geom, mass, twisted_mass, aniso_coeff_s=1.0, aniso_coeff_t=1.0):
        super().__init__(geom, aniso_coeff_s, aniso_coeff_t)
        self.alpha = 4.0 + float(mass)
        self.mu_phys = float(twisted_mass)
        self.mu = self.mu_phys / self.alpha
        self.muinv = self.alpha / (self.alpha ** 2 + self.mu_phys ** 2)

    def dslash(self, res, psi, u, isign, cb):
        _run(self.geom, "tm_dslash", res, psi, None, u, None, isign, cb,
             (self.aniso_s, self.aniso_t, self.aniso_t, self.mu, self.muinv))

    def dslashAChiMinusBDPsi(self, res, psi, chi, u, alpha, beta, isign, cb):
        b = float(beta)
        _run(self.geom, "tm_dslash_achimbdpsi", res, psi, chi, u, None, isign, cb,
             (alpha, b * self.aniso_s, b * self.aniso_t, b * self.aniso_t, self.mu_phys / float(alpha)))


class TMClovDslash(Dslash):
    """Twisted-mass clover (QPhiX TMClovDslash): full complex 6x6 blocks."""

    def dslash(self, res, psi, u, invclov, isign, cb):
        _run(self.geom, "tm_clov_dslash", res, psi, None, u, invclov, isign, cb,
             (self.aniso_s, self.aniso_t, self.aniso_t))

    def dslashAChiMinusBDPsi(self, res, psi, chi, u, clov, beta, isign, cb, prec_mass_rho=0.0):
        b = float(beta)
        _run(self.geom, "tm_clov_dslash_achimbdpsi", res, psi, chi, u, clov, isign, cb,
             (b * self.aniso_s, b * self.aniso_t, b * self.aniso_t, prec_mass_rho))
