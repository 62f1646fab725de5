"""torch.autograd bindings of the differentiation kernels (for checking).

DslashFunction       out = D psi (Wilson hop, lexicographic FP64 in/out).
                     backward: dL/dpsi = D^dagger grad (qphix_amd_dslash_adjoint),
                     dL/dU = G^dagger (qphix_amd_dslash_force, project = 0).
PseudofermionAction  S = phi^dagger (M^dagger M)^-1 phi with the even-odd
                     operator M = F_oo - 1/4 D_oe F_ee^-1 D_eo on cb 1,
                     F = A + i mu g5, A = clover_from_links(1/(2 kappa), csw)
                     (csw = mu = 0: M = alpha - 1/(4 alpha) D_oe D_eo).
                     backward recomputes the solve (nothing solver-sized is
                     kept between forward and backward) and calls
                     qphix_amd_eo_op_force(_clover):
                       dS = 1/2 Re[<Y, dD_oe Z> + <W, dD_eo X>]
                            - 2 Re <Y, dA_oo X> - 1/2 Re <W, dA_ee Z>,
                     X = (M^dag M)^-1 phi, Y = M X, Z = F_ee^-1 D_eo X,
                     W = (F_ee^-1)^dag D_oe^dag Y.
su3_force(...)       the traceless anti-Hermitian su(3) force of either loss,
                     d/de L(exp(i e T) U) = sum_links Re tr(i T F).
"""

import torch

from ._lib import F32, F64, FIELD_SPINOR, QPhiXAMDError, check, lib
from .dslash import Dslash, _ptr, _stream  # noqa: F401
from .solver import LEX_GAUGE, EvenOddCloverOperator, InvCG, _lex64, clover_from_links


def dslash_force(geom, force, chi, psi, U, isign, cb, scale=1.0, project=True):
    """Accumulate scale * (link gradient of Re<chi, D psi>) into force (lex
    FP64 (V,4,3,3,2)); chi, psi native fields on cb, 1-cb (FP64 or FP32)."""
    if chi.prec != psi.prec or chi.prec not in (F64, F32):
        raise QPhiXAMDError("chi and psi must share a precision, FP64 or FP32")
    for f, what in ((chi, "chi"), (psi, "psi")):
        if f.kind != FIELD_SPINOR or f.data.numel() != 24 * geom.vcb:
            raise QPhiXAMDError(f"{what}: not a spinor of this geometry")
    if chi.cb not in (None, cb) or psi.cb not in (None, 1 - cb):
        raise QPhiXAMDError("chi lives on cb and psi on 1 - cb")
    force, U = _lex64(geom, (force, LEX_GAUGE, "force", True), (U, LEX_GAUGE, "U", False))
    check(lib().qphix_amd_dslash_force(geom.ctx, chi.prec, isign, cb, _ptr(force), chi.ptr(), psi.ptr(),
                                       _ptr(U), 1.0, 1.0, 1.0, float(scale), int(bool(project)),
                                       _stream()),
          "qphix_amd_dslash_force")
    return force


def dslash_adjoint(geom, grad_psi, chi, gauge_other_cb, isign, cb):
    """grad_psi (on 1-cb) = D^dagger chi for D with output checkerboard cb."""
    check(lib().qphix_amd_dslash_adjoint(geom.ctx, chi.prec, gauge_other_cb.recon, isign, cb,
                                         grad_psi.ptr(), chi.ptr(), gauge_other_cb.ptr(), 1.0, 1.0,
                                         1.0, _stream()),
          "qphix_amd_dslash_adjoint")
    grad_psi.cb = 1 - cb
    return grad_psi


class DslashFunction(torch.autograd.Function):
    """out = D psi; psi read on 1-cb, out written on cb (other sites zero)."""

    @staticmethod
    def forward(ctx, psi, U, geom, isign, cb, prec=F64, recon=12):
        g = geom.pack_gauge(U.detach(), prec, recon, cb)
        ps = geom.pack_spinor(psi.detach(), prec, 1 - cb)
        res = geom.new_spinor(prec, cb=cb)
        Dslash(geom).dslash(res, ps, g, isign, cb)
        ctx.save_for_backward(psi, U)
        ctx.cfg = (geom, isign, cb, prec, recon)
        return geom.unpack_spinor(res)

    @staticmethod
    def backward(ctx, grad_out):
        psi, U = ctx.saved_tensors
        geom, isign, cb, prec, recon = ctx.cfg
        chi = geom.pack_spinor(grad_out.contiguous(), prec, cb)
        grad_psi = grad_U = None
        if ctx.needs_input_grad[0]:
            g1 = geom.pack_gauge(U.detach(), prec, recon, 1 - cb)
            gp = geom.new_spinor(prec, cb=1 - cb)
            dslash_adjoint(geom, gp, chi, g1, isign, cb)
            grad_psi = geom.unpack_spinor(gp)
        if ctx.needs_input_grad[1]:
            ps = geom.pack_spinor(psi.detach(), prec, 1 - cb)
            grad_U = torch.zeros_like(U)
            dslash_force(geom, grad_U, chi, ps, U.detach().contiguous(), isign, cb, project=False)
        return grad_psi, grad_U, None, None, None, None, None


def _eo_operator(geom, U, kappa, csw=0.0, mu=0.0):
    alpha = 1.0 / (2.0 * kappa)
    A = clover_from_links(geom, U, alpha, csw)  # alpha * identity for csw = 0
    return EvenOddCloverOperator(geom, U, A, mu=mu, prec_sloppy=F64), alpha


def _pf_solution(geom, U, phi, kappa, rsd, csw=0.0, mu=0.0):
    op, alpha = _eo_operator(geom, U, kappa, csw, mu)
    b = geom.pack_spinor(phi, F64, 1)
    x = geom.new_spinor(F64, cb=1)
    InvCG(op, max_iters=100000)(x, b, rsd)
    return op, alpha, b, x


def _pf_force(geom, U, phi, kappa, rsd, scale, project, csw=0.0, mu=0.0):
    op, alpha, b, X = _pf_solution(geom, U, phi, kappa, rsd, csw, mu)
    F = op.force(torch.zeros_like(U), X, U, scale=scale, project=project, csw=csw if csw else None)
    op.close()
    return F


class PseudofermionAction(torch.autograd.Function):
    """S = phi^dag (M^dag M)^-1 phi, M the even-odd clover / twisted-mass
    clover operator with A = clover_from_links(U, 1/(2 kappa), csw) (FP64 CG
    to rsd); csw = 0, mu = 0 is plain Wilson."""

    @staticmethod
    def forward(ctx, U, phi, geom, kappa, rsd=1e-13, csw=0.0, mu=0.0):
        U = U.detach().contiguous()
        op, alpha, b, x = _pf_solution(geom, U, phi.detach(), kappa, rsd, csw, mu)
        S = torch.dot(b.data, x.data)
        op.close()
        ctx.save_for_backward(U, phi)
        ctx.cfg = (geom, kappa, rsd, csw, mu)
        return S

    @staticmethod
    def backward(ctx, gS):
        U, phi = ctx.saved_tensors
        geom, kappa, rsd, csw, mu = ctx.cfg
        F = _pf_force(geom, U, phi.detach(), kappa, rsd, float(gS), False, csw, mu)
        return F, None, None, None, None, None, None


def su3_force_dslash(geom, chi_lex, psi_lex, U, isign, cb, prec=F64):
    """Traceless anti-Hermitian force of L = Re<chi, D psi>."""
    chi = geom.pack_spinor(chi_lex, prec, cb)
    psi = geom.pack_spinor(psi_lex, prec, 1 - cb)
    return dslash_force(geom, torch.zeros_like(U), chi, psi, U.contiguous(), isign, cb, project=True)


def su3_force_pseudofermion(geom, U, phi, kappa, rsd=1e-13, csw=0.0, mu=0.0):
    """Traceless anti-Hermitian force of S = phi^dag (M^dag M)^-1 phi."""
    return _pf_force(geom, U.contiguous(), phi, kappa, rsd, 1.0, True, csw, mu)
