"""Even-odd preconditioned clover operator and CG over the C ABI.

QPhiX builds its even-odd (twisted-mass) clover operator from the two
generated kernels this library replaces -- clov_dslash (A_ee^-1 D_eo) and
clov_dslash_achimbdpsi (A_oo chi - 1/4 D_oe t) -- and solves M^dag M x = b
with InvCG; here both are single C ABI calls (qphix_amd_eo_op_*,
qphix_amd_cg_solve).  The operator acts on the odd checkerboard (cb 1):
    M = F_oo - 1/4 D_oe F_ee^-1 D_eo,  F = A + i mu g5.
"""
import ctypes

import torch

from ._lib import BF16, F64, FIELD_SPINOR, CGParams, CGResult, QPhiXAMDError, check, lib
from .dslash import Field, _ptr, _stream

# reals per lattice site of the lexicographic FP64 fields
LEX_GAUGE, LEX_CLOVER, LEX_FULLCLOVER, LEX_SPINOR = 72, 72, 144, 24


def _lex64(geom, *args):
    """Check lexicographic FP64 device fields before their pointers reach
    the C ABI, which reads (or writes, for outputs) exactly 2 * vcb * reals
    doubles of each.  args: (tensor, reals per site, name, is_output)
    tuples.  Every argument's dtype, size and (for outputs, written in place)
    contiguity is checked before any argument's device, so the error names
    the first malformed argument.  Returns the inputs made contiguous (and
    the outputs as they are), in order; one tensor for one argument."""
    for t, reals, what, out in args:
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64:
            raise QPhiXAMDError(f"{what}: must be a float64 device tensor")
        if t.numel() != 2 * geom.vcb * reals:
            raise QPhiXAMDError(f"{what}: {t.numel()} elements, expected {2 * geom.vcb * reals} "
                                f"({reals} per site of the lattice {geom.dims})")
        if out and not t.is_contiguous():
            raise QPhiXAMDError(f"{what}: written in place, must be contiguous")
    for t, _, what, _ in args:
        if not t.is_cuda:
            raise QPhiXAMDError(f"{what}: must be a float64 device tensor")
    res = [t if out else t.contiguous() for t, _, _, out in args]
    return res[0] if len(res) == 1 else res


def _native_f64_cb1(geom, f, what, check_device=True):
    """An FP64 native spinor on checkerboard 1 (or of unset checkerboard);
    check_device=False leaves the device check to a later call."""
    if not isinstance(f, Field) or f.kind != FIELD_SPINOR or f.prec != F64:
        raise QPhiXAMDError(f"{what}: must be an FP64 native spinor")
    if f.cb not in (None, 1):
        raise QPhiXAMDError(f"{what}: must live on checkerboard 1 (it is on {f.cb})")
    if f.data.numel() != 24 * geom.vcb:
        raise QPhiXAMDError(f"{what}: not a spinor of this geometry")
    if check_device and not f.data.is_cuda:
        raise QPhiXAMDError(f"{what}: must be a device tensor")
    return f


def reunitarize(geom, u_lex):
    """Project every link of u_lex (lex FP64, in place) back onto SU(3)."""
    _lex64(geom, (u_lex, LEX_GAUGE, "u_lex", True))
    check(lib().qphix_amd_reunitarize(geom.ctx, _ptr(u_lex), _stream()), "qphix_amd_reunitarize")
    return u_lex


def clover_from_links(geom, u_lex, alpha, csw):
    """Packed Hermitian clover blocks A(x) (lex FP64, (V,2,36)) from links."""
    u_lex = _lex64(geom, (u_lex, LEX_GAUGE, "u_lex", False))
    out = torch.empty(geom.vcb * 2, 2, 36, dtype=torch.float64, device=u_lex.device)
    check(lib().qphix_amd_clover_from_links(geom.ctx, _ptr(out), _ptr(u_lex),
                                            float(alpha), float(csw), _stream()),
          "qphix_amd_clover_from_links")
    return out


def clover_invert(geom, clov_lex):
    clov_lex = _lex64(geom, (clov_lex, LEX_CLOVER, "clov_lex", False))
    out = torch.empty_like(clov_lex)
    check(lib().qphix_amd_clover_invert(geom.ctx, _ptr(out), _ptr(clov_lex), _stream()),
          "qphix_amd_clover_invert")
    return out


def tmclover_blocks(geom, clov_lex, mu, sign, invert):
    clov_lex = _lex64(geom, (clov_lex, LEX_CLOVER, "clov_lex", False))
    if sign not in (1, -1):
        raise QPhiXAMDError("sign must be +1 or -1")
    out = torch.empty(geom.vcb * 2, 2, 6, 6, 2, dtype=torch.float64, device=clov_lex.device)
    check(lib().qphix_amd_tmclover_blocks(geom.ctx, _ptr(out), _ptr(clov_lex), float(mu),
                                          int(sign), int(bool(invert)), _stream()),
          "qphix_amd_tmclover_blocks")
    return out


def clover_force(geom, force, chi_lex, psi_lex, u_lex, csw, scale_even=1.0, scale_odd=1.0, project=True):
    """Accumulate into force the link gradient of Re sum_x w(x) chi^dag A psi
    (A = clover_from_links(.., csw); w = scale_even / scale_odd by parity)."""
    force, chi_lex, psi_lex, u_lex = _lex64(
        geom, (force, LEX_GAUGE, "force", True), (chi_lex, LEX_SPINOR, "chi_lex", False),
        (psi_lex, LEX_SPINOR, "psi_lex", False), (u_lex, LEX_GAUGE, "u_lex", False))
    check(lib().qphix_amd_clover_force(geom.ctx, _ptr(force), _ptr(chi_lex),
                                       _ptr(psi_lex), _ptr(u_lex), float(csw),
                                       float(scale_even), float(scale_odd), int(bool(project)), _stream()),
          "qphix_amd_clover_force")
    return force


class EvenOddCloverOperator:
    """M = F_oo - 1/4 D_oe F_ee^-1 D_eo on cb 1 with an FP64 copy (12-real
    links) and a sloppy copy (prec_sloppy storage, recon_sloppy links)."""

    def __init__(self, geom, u_lex, clov_lex, mu=0.0, prec_sloppy=F64, recon_sloppy=12):
        self._op = None
        if prec_sloppy not in (F64, 1, 2, BF16) or recon_sloppy not in (12, 18):
            raise QPhiXAMDError("prec_sloppy must be F64/F32/F16/BF16 and recon_sloppy 12 or 18")
        u_lex, clov_lex = _lex64(geom, (u_lex, LEX_GAUGE, "u_lex", False),
                                 (clov_lex, LEX_CLOVER, "clov_lex", False))
        self.geom = geom
        self.prec_sloppy = prec_sloppy
        self.mu = float(mu)
        op = ctypes.c_void_p()
        check(lib().qphix_amd_eo_op_create(ctypes.byref(op), geom.ctx, _ptr(u_lex),
                                           _ptr(clov_lex), self.mu, prec_sloppy,
                                           recon_sloppy, _stream()),
              "qphix_amd_eo_op_create")
        self._op = op

    def close(self):
        if getattr(self, "_op", None):
            lib().qphix_amd_eo_op_destroy(self._op)
            self._op = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, out, inp, dagger=False):
        if out.prec != inp.prec:
            raise QPhiXAMDError("out and in must share a precision")
        if inp.prec not in (F64, self.prec_sloppy):
            raise QPhiXAMDError("the operator holds FP64 and its sloppy precision only")
        for f, what in ((out, "out"), (inp, "in")):
            if f.kind != FIELD_SPINOR or f.data.numel() != 24 * self.geom.vcb:
                raise QPhiXAMDError(f"{what}: not a spinor of this geometry")
        check(lib().qphix_amd_eo_op_apply(self._op, inp.prec, int(bool(dagger)), out.ptr(), inp.ptr(),
                                          _stream()),
              "qphix_amd_eo_op_apply")
        out.cb = 1


    def force(self, force, x, u_lex, scale=1.0, project=True, csw=None):
        """Accumulate the gradient of phi^dag (M^dag M)^-1 phi at
        x = (M^dag M)^-1 phi (FP64, cb 1) into force (lex FP64 (V,4,3,3,2)):
        the hopping term, plus the clover term when csw is given (the
        operator's clover field then being clover_from_links(.., csw))."""
        _native_f64_cb1(self.geom, x, "x", check_device=False)
        force, u_lex = _lex64(self.geom, (force, LEX_GAUGE, "force", True),
                              (u_lex, LEX_GAUGE, "u_lex", False))
        _native_f64_cb1(self.geom, x, "x")
        u = _ptr(u_lex)
        if csw is None:
            st = lib().qphix_amd_eo_op_force(self._op, _ptr(force), x.ptr(), u, float(scale),
                                             int(bool(project)), _stream())
        else:
            st = lib().qphix_amd_eo_op_force_clover(self._op, _ptr(force), x.ptr(), u, float(csw),
                                                    float(scale), int(bool(project)), _stream())
        check(st, "qphix_amd_eo_op_force")
        return force


class InvCG:
    """CG on M^dag M (QPhiX InvCG semantics); mixed=True runs the sloppy
    iterations with FP64 reliable updates."""

    def __init__(self, op, max_iters=10000):
        self.op = op
        self.max_iters = int(max_iters)

    def __call__(self, x, b, rsd_target, mixed=False, delta=0.0, raise_on_noconv=True):
        for f, what in ((x, "x"), (b, "b")):
            _native_f64_cb1(self.op.geom, f, what, check_device=False)
        for f, what in ((x, "x"), (b, "b")):
            _native_f64_cb1(self.op.geom, f, what)
        p = CGParams(float(rsd_target), self.max_iters, float(delta), int(bool(mixed)))
        r = CGResult()
        st = lib().qphix_amd_cg_solve(self.op._op, x.ptr(), b.ptr(), ctypes.byref(p), ctypes.byref(r),
                                      _stream())
        res = dict(iters=r.iters, reliable_updates=r.reliable_updates, rsd_final=r.rsd_final,
                   seconds=r.seconds, mv_apps=r.mv_apps, status=st)
        if st != 0 and (st != -5 or raise_on_noconv):
            check(st, f"qphix_amd_cg_solve ({res})")
        x.cb = 1
        return res
