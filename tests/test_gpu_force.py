"""GPU tests of the differentiation kernels (config 3 semantics).

* the link-gradient kernel (qphix_amd_dslash_force) vs the FP64 oracle
  (FP64 1e-13, FP32 1e-6 relative), projected and Euclidean, both isigns
  and checkerboards, twisted boundary phases;
* dL/dpsi = D^dagger chi (qphix_amd_dslash_adjoint) vs the oracle;
* the torch.autograd binding: gradients of L = Re<chi, D psi> vs the oracle;
* on 16^3 x 32, central finite differences (e = 1e-5, FP64) along 32 random
  su(3) link perturbations of L = Re<chi, D_W psi> and of the pseudofermion
  action S = phi^dag (M^dag M)^-1 phi (even-odd Wilson, FP64 CG), against
  sum Re tr(i T F) from the kernels (1e-6 relative); the FP32 force against
  the FP64 one.
"""
import numpy as np
import pytest
import torch

from _util import cb_mask, cplx, on_cb, real

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


@pytest.fixture(scope="module")
def small(qamd):
    dims = (8, 8, 8, 8)
    L = qamd.lattice
    return dims, L.random_gauge(dims, 0.4, 61), L.random_spinor(dims, 62), L.random_spinor(dims, 63)


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-6)], ids=["f64", "f32"])
def test_force_kernel_vs_oracle(qamd, orc, small, prec, tol):
    dims, U, chi, psi = small
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    ph = [np.exp(0.3j), 1.0, np.exp(-1.1j), np.exp(0.5j)]
    geom.set_tbc([1, 0, 1, 1], ph)
    Ud = U.to(DEV)
    for isign in (1, -1):
        for cb in (0, 1):
            c = geom.pack_spinor(chi.to(DEV), prec, cb)
            p = geom.pack_spinor(psi.to(DEV), prec, 1 - cb)
            for project in (True, False):
                F = qamd.dslash_force(geom, torch.zeros_like(Ud), c, p, Ud, isign, cb, 1.0, project)
                ref = orc.force(dims, on_cb(chi.numpy(), dims, cb), on_cb(psi.numpy(), dims, 1 - cb),
                                U.numpy(), isign, cb, tbc=[1, 0, 1, 1], tbc_phases=ph, project=project)
                err = np.abs(F.cpu().numpy() - ref).max() / np.abs(ref).max()
                assert err < tol, (isign, cb, project, err)


@pytest.mark.parametrize("offset", [0, 1], ids=["aligned", "offset8"])
def test_force_kernel_partial_groups_vs_oracle(qamd, orc, offset):
    """6^4: 1296 sites, so the last work-group of 32 sites is half full and
    Lz is not a multiple of 8 (the other traversal); offset8 passes force
    and link arrays 8 bytes past a 16-byte boundary, so the work-groups'
    link and force runs move element by element."""
    dims = (6, 6, 6, 6)
    print(f"case force partial groups {dims} offset={offset}", flush=True)
    L = qamd.lattice
    U, chi, psi = L.random_gauge(dims, 0.4, 64), L.random_spinor(dims, 65), L.random_spinor(dims, 66)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    ubuf = torch.zeros(U.numel() + 2, dtype=torch.float64, device=DEV)
    Ud = ubuf[offset:offset + U.numel()].view(U.shape)
    Ud.copy_(U.to(DEV))
    assert (Ud.data_ptr() % 16 == 0) == (offset == 0)
    for isign, cb, project in ((1, 0, True), (-1, 1, True), (1, 1, False)):
        c = geom.pack_spinor(chi.to(DEV), 0, cb)
        p = geom.pack_spinor(psi.to(DEV), 0, 1 - cb)
        fbuf = torch.zeros(U.numel() + 2, dtype=torch.float64, device=DEV)
        F = fbuf[offset:offset + U.numel()].view(U.shape)
        F = qamd.dslash_force(geom, F, c, p, Ud, isign, cb, 1.0, project)
        ref = orc.force(dims, on_cb(chi.numpy(), dims, cb), on_cb(psi.numpy(), dims, 1 - cb),
                        U.numpy(), isign, cb, project=project)
        err = np.abs(F.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (isign, cb, project, err)
        # the guard elements around the array are untouched
        assert fbuf[:offset].count_nonzero() == 0 and fbuf[offset + U.numel():].count_nonzero() == 0


def test_adjoint_kernel_vs_oracle(qamd, orc, small):
    dims, U, chi, _ = small
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    for isign in (1, -1):
        for cb in (0, 1):
            c = geom.pack_spinor(chi.to(DEV), 0, cb)
            g = geom.pack_gauge(U.to(DEV), 0, 18, 1 - cb)
            out = geom.new_spinor(0, cb=1 - cb)
            lib = qamd.lib()
            import ctypes
            assert lib.qphix_amd_dslash_adjoint(geom.ctx, 0, 18, isign, cb, out.ptr(), c.ptr(), g.ptr(),
                                                0.9, 1.3, 0.7, None) == 0
            torch.cuda.synchronize()
            got = geom.unpack_spinor(out).cpu().numpy()
            # D^dag chi with D(coeff_t_f=1.3, coeff_t_b=0.7) = D_{-s}(coeff_t_f=0.7, coeff_t_b=1.3)
            ref = orc.dslash(dims, on_cb(chi.numpy(), dims, cb), U.numpy(), -isign, 1 - cb,
                             coeff_s=0.9, coeff_t_f=0.7, coeff_t_b=1.3)
            m = cb_mask(dims, 1 - cb)
            assert np.abs(got[m] - ref[m]).max() / np.abs(ref[m]).max() < 1e-13
            # and it is the adjoint: <chi, D psi> = <D^dag chi, psi>
            rng = np.random.default_rng(5 + isign + 3 * cb)
            psi = on_cb(rng.normal(size=chi.shape), dims, 1 - cb)
            d = orc.dslash(dims, psi, U.numpy(), isign, cb, coeff_s=0.9, coeff_t_f=1.3, coeff_t_b=0.7)
            lhs = np.sum(cplx(on_cb(chi.numpy(), dims, cb)).conj() * cplx(d))
            rhs = np.sum(cplx(got).conj() * cplx(psi))
            assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_autograd_dslash_vs_oracle(qamd, orc, small):
    dims, U, chi, psi = small
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    cb, isign = 1, 1
    Ud = U.to(DEV).requires_grad_(True)
    pd = psi.to(DEV).requires_grad_(True)
    out = qamd.DslashFunction.apply(pd, Ud, geom, isign, cb)
    chid = torch.from_numpy(on_cb(chi.numpy(), dims, cb)).to(DEV)
    L = (chid * out).sum()
    L.backward()
    ref_psi = orc.dslash(dims, on_cb(chi.numpy(), dims, cb), U.numpy(), -isign, 1 - cb)
    ref_U = orc.force(dims, on_cb(chi.numpy(), dims, cb), on_cb(psi.numpy(), dims, 1 - cb), U.numpy(),
                      isign, cb, project=False)
    gp = pd.grad.cpu().numpy()
    m = cb_mask(dims, 1 - cb)
    assert np.abs(gp[m] - ref_psi[m]).max() / np.abs(ref_psi[m]).max() < 1e-13
    assert np.abs(gp[~m]).max() == 0
    assert np.abs(Ud.grad.cpu().numpy() - ref_U).max() / np.abs(ref_U).max() < 1e-13


def _su3_dirs(qamd, V, n, seed):
    g = torch.Generator().manual_seed(seed)
    lam = qamd.lattice.gellmann(DEV)
    for k in range(n):
        w = torch.randn(V, 4, 8, generator=g, dtype=torch.float64).to(DEV)
        yield torch.einsum("vma,aij->vmij", w.to(torch.complex128), lam) * 0.5


def _rotate(U, T, eps):
    Uc = torch.view_as_complex(U)
    return torch.view_as_real((torch.linalg.matrix_exp(1j * eps * T) @ Uc).contiguous()).contiguous()


def _directional(F, T):
    Fc = torch.view_as_complex(F.contiguous())
    return torch.real(1j * torch.einsum("vmij,vmji->", T, Fc)).item()


def test_force_finite_differences_dslash_16x32(qamd):
    dims = (16, 16, 16, 32)
    L = qamd.lattice
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    U = L.random_gauge(dims, 0.3, 71, DEV)
    chi = L.random_spinor(dims, 72, DEV)
    psi = L.random_spinor(dims, 73, DEV)
    cb, isign = 0, 1
    chic = geom.pack_spinor(chi, 0, cb)

    def loss(Uv):
        g = geom.pack_gauge(Uv, 0, 18, cb)
        out = geom.new_spinor(0, cb=cb)
        qamd.Dslash(geom).dslash(out, geom.pack_spinor(psi, 0, 1 - cb), g, isign, cb)
        return torch.dot(chic.data, out.data).item()

    F = qamd.su3_force_dslash(geom, chi, psi, U, isign, cb)
    F32 = qamd.su3_force_dslash(geom, chi, psi, U, isign, cb, prec=1)
    assert ((F32 - F).abs().max() / F.abs().max()).item() < 1e-6
    eps = 1e-5
    worst = 0.0
    for T in _su3_dirs(qamd, U.shape[0], 32, 74):
        fd = (loss(_rotate(U, T, eps)) - loss(_rotate(U, T, -eps))) / (2 * eps)
        an = _directional(F, T)
        worst = max(worst, abs(fd - an) / abs(an))
    print("dslash force FD worst rel err", worst)
    assert worst < 1e-6


def test_force_finite_differences_pseudofermion_16x32(qamd):
    dims = (16, 16, 16, 32)
    L = qamd.lattice
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    U = L.random_gauge(dims, 0.3, 81, DEV)
    phi = L.random_spinor(dims, 82, DEV)
    kappa, rsd = 0.12, 1e-14

    def action(Uv):
        return qamd.PseudofermionAction.apply(Uv, phi, geom, kappa, rsd).item()

    F = qamd.su3_force_pseudofermion(geom, U, phi, kappa, rsd)
    # the autograd (Euclidean) gradient is consistent with the projected force:
    # TA(U G) with G^dag = dS/dU
    Ur = U.clone().requires_grad_(True)
    S = qamd.PseudofermionAction.apply(Ur, phi, geom, kappa, rsd)
    S.backward()
    Gd = torch.view_as_complex(Ur.grad.contiguous())
    W = torch.view_as_complex(U) @ Gd.conj().transpose(-1, -2)
    TA = 0.5 * (W - W.conj().transpose(-1, -2))
    tr = torch.diagonal(TA, dim1=-2, dim2=-1).sum(-1) / 3
    TA = TA - tr[..., None, None] * torch.eye(3, dtype=TA.dtype, device=DEV)
    Fc = torch.view_as_complex(F.contiguous())
    assert ((TA - Fc).abs().max() / Fc.abs().max()).item() < 1e-12
    eps = 1e-5
    worst = 0.0
    for T in _su3_dirs(qamd, U.shape[0], 32, 84):
        fd = (action(_rotate(U, T, eps)) - action(_rotate(U, T, -eps))) / (2 * eps)
        an = _directional(F, T)
        worst = max(worst, abs(fd - an) / abs(an))
    print("pseudofermion force FD worst rel err", worst)
    assert worst < 1e-6


def test_clover_force_vs_autograd(qamd):
    """qphix_amd_clover_force against torch autograd of the torch restatement
    of A(U) (tests/clover_torch.py, itself equal to the oracle's builder):
    L = Re sum_x w(x) chi^dag A psi with parity weights, Euclidean gradient
    and su(3) projection, 1e-12 relative."""
    import clover_torch as ct
    dims = (4, 4, 4, 8)
    L = qamd.lattice
    V = 512
    U = L.random_gauge(dims, 0.5, 91)
    chi, psi = L.random_spinor(dims, 92), L.random_spinor(dims, 93)
    csw, we, wo = 1.3, 0.7, -1.9
    Ur = U.clone().requires_grad_(True)
    A = ct.clover_matrix(torch.complex(Ur[..., 0], Ur[..., 1]), dims, 3.7, csw)
    c = torch.complex(chi[..., 0], chi[..., 1]).reshape(V, 12)
    p = torch.complex(psi[..., 0], psi[..., 1]).reshape(V, 12)
    w = torch.where(L.parity(dims) == 1, torch.tensor(wo, dtype=torch.float64),
                    torch.tensor(we, dtype=torch.float64))
    loss = (w * torch.einsum("vi,vij,vj->v", c.conj(), A, p).real).sum()
    loss.backward()
    ref = Ur.grad  # real-form gradient = G^dagger
    geom = qamd.Geometry(dims)
    got = qamd.clover_force(geom, torch.zeros_like(U).to(DEV), chi.to(DEV), psi.to(DEV), U.to(DEV), csw, we, wo,
                            project=False).cpu()
    assert ((got - ref).abs().max() / ref.abs().max()).item() < 1e-12
    # projected: F = TA(U G), G = (G^dagger)^dagger
    Gd = torch.complex(ref[..., 0], ref[..., 1])
    W = torch.complex(U[..., 0], U[..., 1]) @ Gd.conj().transpose(-1, -2)
    TA = 0.5 * (W - W.conj().transpose(-1, -2))
    TA = TA - (torch.diagonal(TA, dim1=-2, dim2=-1).sum(-1) / 3)[..., None, None] * torch.eye(3, dtype=TA.dtype)
    gotP = qamd.clover_force(geom, torch.zeros_like(U).to(DEV), chi.to(DEV), psi.to(DEV), U.to(DEV), csw, we, wo,
                             project=True).cpu()
    gp = torch.complex(gotP[..., 0], gotP[..., 1])
    assert ((gp - TA).abs().max() / TA.abs().max()).item() < 1e-12


@pytest.mark.parametrize("csw,mu", [(1.0, 0.0), (1.5, 0.1)], ids=["clover", "tmclover"])
def test_force_finite_differences_pseudofermion_clover(qamd, csw, mu):
    """The whole pseudofermion force (hopping + clover term) of the even-odd
    clover / twisted-mass clover operator with A = clover_from_links(U, ..,
    csw) against central differences of S along 16 su(3) directions
    (eps = 1e-5, FP64): |fd - an| <= 1e-6 of the largest directional
    derivative (a per-direction ratio is meaningless where a direction's
    derivative is itself near zero); the hopping term alone misses by far
    more, so the clover term is what is being checked; and the Euclidean
    autograd gradient is consistent with the projected force."""
    dims = (8, 8, 8, 8)
    L = qamd.lattice
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    U = L.random_gauge(dims, 0.3, 101, DEV)
    phi = L.random_spinor(dims, 102, DEV)
    kappa, rsd = 0.12, 1e-14

    def action(Uv):
        return qamd.PseudofermionAction.apply(Uv, phi, geom, kappa, rsd, csw, mu).item()

    F = qamd.su3_force_pseudofermion(geom, U, phi, kappa, rsd, csw, mu)
    Ur = U.clone().requires_grad_(True)
    qamd.PseudofermionAction.apply(Ur, phi, geom, kappa, rsd, csw, mu).backward()
    Gd = torch.view_as_complex(Ur.grad.contiguous())
    W = torch.view_as_complex(U) @ Gd.conj().transpose(-1, -2)
    TA = 0.5 * (W - W.conj().transpose(-1, -2))
    tr = torch.diagonal(TA, dim1=-2, dim2=-1).sum(-1) / 3
    TA = TA - tr[..., None, None] * torch.eye(3, dtype=TA.dtype, device=DEV)
    Fc = torch.view_as_complex(F.contiguous())
    assert ((TA - Fc).abs().max() / Fc.abs().max()).item() < 1e-12
    # the hopping term alone (the operator's force without the clover term)
    op, _, _, X = qamd.autograd._pf_solution(geom, U, phi, kappa, rsd, csw, mu)
    Fh = op.force(torch.zeros_like(U), X, U, project=True)
    op.close()
    eps = 1e-5
    errs, errs_hop, scale = [], [], 0.0
    for T in _su3_dirs(qamd, U.shape[0], 16, 103):
        fd = (action(_rotate(U, T, eps)) - action(_rotate(U, T, -eps))) / (2 * eps)
        an = _directional(F, T)
        scale = max(scale, abs(an))
        errs.append(abs(fd - an))
        errs_hop.append(abs(fd - _directional(Fh, T)))
    worst, worst_hop = max(errs) / scale, max(errs_hop) / scale
    print(f"clover pseudofermion force csw={csw} mu={mu}: FD worst err {worst:.3g} of the largest "
          f"derivative (hopping term alone: {worst_hop:.3g})", flush=True)
    assert worst < 1e-6
    assert worst_hop > 1e-3
