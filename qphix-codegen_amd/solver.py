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

from ._lib import F64, CGParams, CGResult, QPhiXAMDError, check, lib
from .dslash import Field, _ptr, _stream


def reunitarize(geom, u_lex):
    """Project every link of u_lex (lex FP64, in place) back onto SU(3)."""
    check(lib().qphix_amd_reunitarize(geom.ctx, _ptr(u_lex), _stream()), "qphix_amd_reunitarize")
    return u_lex


def clover_from_links(geom, u_lex, alpha, csw):
    """Packed Hermitian clover blocks A(x) (lex FP64, (V,2,36)) from links."""
    out = torch.empty(geom.vcb * 2, 2, 36, dtype=torch.float64, device=u_lex.device)
    check(lib().qphix_amd_clover_from_links(geom.ctx, _ptr(out), _ptr(u_lex.contiguous()),
                                            float(alpha), float(csw), _stream()),
          "qphix_amd_clover_from_links")
    return out


def clover_invert(geom, clov_lex):
    out = torch.empty_like(clov_lex)
    check(lib().qphix_amd_clover_invert(geom.ctx, _ptr(out), _ptr(clov_lex.contiguous()), _stream()),
          "qphix_amd_clover_invert")
    return out


def tmclover_blocks(geom, clov_lex, mu, sign, invert):
    out = torch.empty(geom.vcb * 2, 2, 6, 6, 2, dtype=torch.float64, device=clov_lex.device)
    check(lib().qphix_amd_tmclover_blocks(geom.ctx, _ptr(out), _ptr(clov_lex.contiguous()), float(mu),
                                          int(sign), int(bool(invert)), _stream()),
          "qphix_amd_tmclover_blocks")
    return out


def clover_force(geom, force, chi_lex, psi_lex, u_lex, csw, scale_even=1.0, scale_odd=1.0, project=True):
    """Accumulate into force the link gradient of Re sum_x w(x) chi^dag A psi
    (A = clover_from_links(.., csw); w = scale_even / scale_odd by parity)."""
    check(lib().qphix_amd_clover_force(geom.ctx, _ptr(force), _ptr(chi_lex.contiguous()),
                                       _ptr(psi_lex.contiguous()), _ptr(u_lex.contiguous()), float(csw),
                                       float(scale_even), float(scale_odd), int(bool(project)), _stream()),
          "qphix_amd_clover_force")
    return force


class EvenOddCloverOperator:
    """M = F_oo - 1/4 D_oe F_ee^-1 D_eo on cb 1 with an FP64 copy (12-real
    links) and a sloppy copy (prec_sloppy storage, recon_sloppy links)."""

    def __init__(self, geom, u_lex, clov_lex, mu=0.0, prec_sloppy=F64, recon_sloppy=12):
        self.geom = geom
        self.prec_sloppy = prec_sloppy
        self.mu = float(mu)
        op = ctypes.c_void_p()
        check(lib().qphix_amd_eo_op_create(ctypes.byref(op), geom.ctx, _ptr(u_lex.contiguous()),
                                           _ptr(clov_lex.contiguous()), self.mu, prec_sloppy,
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
        check(lib().qphix_amd_eo_op_apply(self._op, inp.prec, int(bool(dagger)), out.ptr(), inp.ptr(),
                                          _stream()),
              "qphix_amd_eo_op_apply")
        out.cb = 1


    def force(self, force, x, u_lex, scale=1.0, project=True, csw=None):
        """Accumulate the gradient of phi^dag (M^dag M)^-1 phi at
        x = (M^dag M)^-1 phi (FP64, cb 1) into force (lex FP64 (V,4,3,3,2)):
        the hopping term, plus the clover term when csw is given (the
        operator's clover field then being clover_from_links(.., csw))."""
        u = _ptr(u_lex.contiguous())
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
        if x.prec != F64 or b.prec != F64:
            raise QPhiXAMDError("x and b are FP64 native spinors on cb 1")
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
