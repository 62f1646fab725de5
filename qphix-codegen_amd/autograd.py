"""torch.autograd bindings of the differentiation kernels (for checking).

DslashFunction       out = D psi (Wilson hop, lexicographic FP64 in/out).
                     backward: dL/dpsi = D^dagger grad (qphix_amd_dslash_adjoint),
                     dL/dU = G^dagger (qphix_amd_dslash_force, project = 0).
PseudofermionAction  S = phi^dagger (M^dagger M)^-1 phi with the even-odd Wilson
                     operator M = alpha - 1/(4 alpha) D_oe D_eo on cb 1,
                     alpha = 1/(2 kappa).  backward recomputes the solve
                     (nothing solver-sized is kept between forward and
                     backward) and calls qphix_amd_eo_op_force:
                       dS = 1/2 Re[<Y, dD_oe Z> + <W, dD_eo X>],
                     X = (M^dag M)^-1 phi, Y = M X, Z = D_eo X / alpha,
                     W = D_oe^dag Y / alpha.
su3_force(...)       the traceless anti-Hermitian su(3) force of either loss,
                     d/de L(exp(i e T) U) = sum_links Re tr(i T F).
"""

import torch

from ._lib import F64, check, lib
from .dslash import Dslash, _ptr, _stream  # noqa: F401
from .solver import EvenOddCloverOperator, InvCG, clover_from_links


def dslash_force(geom, force, chi, psi, U, isign, cb, scale=1.0, project=True):
    """Accumulate scale * (link gradient of Re<chi, D psi>) into force (lex
    FP64 (V,4,3,3,2)); chi, psi native fields on cb, 1-cb (FP64 or FP32)."""
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


def _wilson_eo_operator(geom, U, kappa):
    alpha = 1.0 / (2.0 * kappa)
    A = clover_from_links(geom, U, alpha, 0.0)  # alpha * identity
    return EvenOddCloverOperator(geom, U, A, prec_sloppy=F64), alpha


def _pf_solution(geom, U, phi, kappa, rsd):
    op, alpha = _wilson_eo_operator(geom, U, kappa)
    b = geom.pack_spinor(phi, F64, 1)
    x = geom.new_spinor(F64, cb=1)
    InvCG(op, max_iters=100000)(x, b, rsd)
    return op, alpha, b, x


def _pf_force(geom, U, phi, kappa, rsd, scale, project):
    op, alpha, b, X = _pf_solution(geom, U, phi, kappa, rsd)
    F = op.force(torch.zeros_like(U), X, U, scale=scale, project=project)
    op.close()
    return F


class PseudofermionAction(torch.autograd.Function):
    """S = phi^dag (M^dag M)^-1 phi (even-odd Wilson, FP64 CG to rsd)."""

    @staticmethod
    def forward(ctx, U, phi, geom, kappa, rsd=1e-13):
        U = U.detach().contiguous()
        op, alpha, b, x = _pf_solution(geom, U, phi.detach(), kappa, rsd)
        S = torch.dot(b.data, x.data)
        op.close()
        ctx.save_for_backward(U, phi)
        ctx.cfg = (geom, kappa, rsd)
        return S

    @staticmethod
    def backward(ctx, gS):
        U, phi = ctx.saved_tensors
        geom, kappa, rsd = ctx.cfg
        F = _pf_force(geom, U, phi.detach(), kappa, rsd, float(gS), project=False)
        return F, None, None, None, None


def su3_force_dslash(geom, chi_lex, psi_lex, U, isign, cb, prec=F64):
    """Traceless anti-Hermitian force of L = Re<chi, D psi>."""
    chi = geom.pack_spinor(chi_lex, prec, cb)
    psi = geom.pack_spinor(psi_lex, prec, 1 - cb)
    return dslash_force(geom, torch.zeros_like(U), chi, psi, U.contiguous(), isign, cb, project=True)


def su3_force_pseudofermion(geom, U, phi, kappa, rsd=1e-13):
    """Traceless anti-Hermitian force of S = phi^dag (M^dag M)^-1 phi."""
    return _pf_force(geom, U.contiguous(), phi, kappa, rsd, 1.0, project=True)
