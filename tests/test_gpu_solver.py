"""GPU tests of the clover builder, the even-odd operator and the CG.

* clover from links, its inverse and the twisted-mass blocks vs the FP64
  oracle (1e-13 relative);
* M and M^dag (FP64, FP32, FP16, BF16 storage) vs the oracle's composition
  F_oo - 1/4 D_oe F_ee^-1 D_eo of its own dslash;
* CG on M^dag M, FP64 and mixed precision (FP32/FP16/BF16 sloppy), to a
  true relative residual of 1e-10 recomputed by the FP64 oracle on the CPU,
  and mixed vs FP64 solution agreement;
* repeated solves bit-identical; the CG on a lattice whose checkerboard is
  not a whole number of 128-site work-groups (separate dot product instead
  of the norm fused into the operator kernel); the CG at 16^4, where the
  operator kernels run traversal order 1.
"""
import numpy as np
import pytest
import torch

from _util import cb_mask, on_cb

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
KAPPA = 0.1350
ALPHA = 1.0 / (2 * KAPPA)
CSW = 1.0


@pytest.fixture(scope="module")
def lat(qamd):
    dims = (8, 8, 8, 8)
    L = qamd.lattice
    U = L.random_gauge(dims, 0.2, 2000)
    b = L.random_spinor(dims, 2001)
    return dims, U, b


def oracle_M(orc, dims, U, A, psi, dagger, mu=0.0, lat=None):
    isign = -1 if dagger else 1
    if mu == 0.0:
        inv = orc.clover_invert(A)
        t = orc.dslash(dims, psi, U, isign, 0, kind=orc.CLOVER, clov=inv)
        return orc.dslash(dims, t, U, isign, 1, kind=orc.CLOVER, achi=True, chi=psi, clov=A,
                          coeff_s=0.25, coeff_t_f=0.25, coeff_t_b=0.25)
    B = orc.tmclover_blocks(A, mu, isign, True)
    F = orc.tmclover_blocks(A, mu, isign, False)
    t = orc.dslash(dims, psi, U, isign, 0, kind=orc.TMCLOVER, clov=B)
    return orc.dslash(dims, t, U, isign, 1, kind=orc.TMCLOVER, achi=True, chi=psi, clov=F,
                      coeff_s=0.25, coeff_t_f=0.25, coeff_t_b=0.25)


def test_clover_build_vs_oracle(qamd, orc, lat):
    dims, U, _ = lat
    geom = qamd.Geometry(dims)
    A = qamd.clover_from_links(geom, U.to(DEV), ALPHA, CSW).cpu().numpy()
    ref = orc.clover_from_links(dims, U.numpy(), ALPHA, CSW)
    assert np.abs(A - ref).max() / np.abs(ref).max() < 1e-13
    inv = qamd.clover_invert(geom, torch.from_numpy(ref).to(DEV)).cpu().numpy()
    rinv = orc.clover_invert(ref)
    assert np.abs(inv - rinv).max() / np.abs(rinv).max() < 1e-13
    for sign in (1, -1):
        for invert in (0, 1):
            bl = qamd.tmclover_blocks(geom, torch.from_numpy(ref).to(DEV), 0.1, sign, invert).cpu().numpy()
            rb = orc.tmclover_blocks(ref, 0.1, sign, invert)
            assert np.abs(bl - rb).max() / np.abs(rb).max() < 1e-13


@pytest.mark.parametrize("mu", [0.0, 0.1], ids=["clover", "tmclover"])
def test_eo_operator_vs_oracle(qamd, orc, lat, mu):
    dims, U, b = lat
    geom = qamd.Geometry(dims)
    Ud = U.to(DEV)
    A = qamd.clover_from_links(geom, Ud, ALPHA, CSW)
    An = A.cpu().numpy()
    psi = on_cb(b.numpy(), dims, 1)
    m = cb_mask(dims, 1)
    for prec_s, tol in ((qamd.F32, 1e-6), (qamd.F16, 5e-3), (qamd.BF16, 4e-2)):
        op = qamd.EvenOddCloverOperator(geom, Ud, A, mu=mu, prec_sloppy=prec_s, recon_sloppy=12)
        for prec, tl in ((qamd.F64, 1e-13), (prec_s, tol)):
            for dagger in (False, True):
                src = geom.pack_spinor(b.to(DEV), prec, 1)
                out = geom.new_spinor(prec, cb=1)
                op.apply(out, src, dagger)
                got = geom.unpack_spinor(out).cpu().numpy()
                ref = oracle_M(orc, dims, U.numpy(), An, psi, dagger, mu)
                err = np.abs(got[m] - ref[m]).max() / np.abs(ref[m]).max()
                assert err < tl, (prec, dagger, err)
        op.close()


def _true_residual(orc, dims, U, A, x, b, mu):
    y = oracle_M(orc, dims, U, A, x, False, mu)
    y = oracle_M(orc, dims, U, A, on_cb(y, dims, 1), True, mu)
    m = cb_mask(dims, 1)
    return np.linalg.norm((b - y)[m]) / np.linalg.norm(b[m])


@pytest.mark.parametrize("mu", [0.0, 0.1], ids=["clover", "tmclover"])
def test_cg_fp64_and_mixed(qamd, orc, lat, mu):
    dims, U, b = lat
    geom = qamd.Geometry(dims)
    Ud = U.to(DEV)
    A = qamd.clover_from_links(geom, Ud, ALPHA, CSW)
    An = A.cpu().numpy()
    bn = on_cb(b.numpy(), dims, 1)
    bf = geom.pack_spinor(b.to(DEV), qamd.F64, 1)
    sols = {}
    for prec_s in (qamd.F64, qamd.F32, qamd.F16, qamd.BF16):
        op = qamd.EvenOddCloverOperator(geom, Ud, A, mu=mu, prec_sloppy=prec_s, recon_sloppy=12)
        x = geom.new_spinor(qamd.F64, cb=1)
        res = qamd.InvCG(op, max_iters=5000)(x, bf, 1e-10, mixed=prec_s != qamd.F64)
        print(f"mu={mu} sloppy={qamd.PREC_NAME[prec_s]}: {res}", flush=True)
        assert res["rsd_final"] < 1e-10
        xn = geom.unpack_spinor(x).cpu().numpy()
        r = _true_residual(orc, dims, U.numpy(), An, xn, bn, mu)
        assert r < 1.5e-10, (prec_s, r)
        sols[prec_s] = xn
        op.close()
    ref = sols[qamd.F64]
    for prec_s, xs in sols.items():
        d = np.linalg.norm(xs - ref) / np.linalg.norm(ref)
        assert d < 1e-8, (prec_s, d)


def test_cg_traversal_order1_16x4(qamd, orc):
    """16^4: Lt <= Lz, so the operator kernels run with the XCDs sweeping t
    together over z-ranges (order 1) and the x/y rows staged in LDS, and the
    CG takes <p, M^dag M p> and the residual update from the operator's
    fused epilogues.  Twisted-mass clover, FP64 and FP16/BF16 mixed, to a
    true residual of 1e-10 recomputed by the oracle."""
    dims = (16, 16, 16, 16)
    L = qamd.lattice
    U = L.random_gauge(dims, 0.2, 2200)
    b = L.random_spinor(dims, 2201)
    geom = qamd.Geometry(dims)
    assert geom.traversal == 1
    Ud = U.to(DEV)
    A = qamd.clover_from_links(geom, Ud, ALPHA, CSW)
    An = A.cpu().numpy()
    bn = on_cb(b.numpy(), dims, 1)
    bf = geom.pack_spinor(b.to(DEV), qamd.F64, 1)
    for prec_s in (qamd.F64, qamd.F16, qamd.BF16):
        op = qamd.EvenOddCloverOperator(geom, Ud, A, mu=0.05, prec_sloppy=prec_s, recon_sloppy=12)
        x = geom.new_spinor(qamd.F64, cb=1)
        res = qamd.InvCG(op, max_iters=5000)(x, bf, 1e-10, mixed=prec_s != qamd.F64)
        print(f"16^4 order 1 sloppy={qamd.PREC_NAME[prec_s]}: {res}", flush=True)
        xn = geom.unpack_spinor(x).cpu().numpy()
        r = _true_residual(orc, dims, U.numpy(), An, xn, bn, 0.05)
        assert res["rsd_final"] < 1e-10 and r < 1.5e-10, (prec_s, res["rsd_final"], r)
        op.close()


def test_cg_repeatable(qamd, lat):
    """Deterministic reductions: two solves of the same system give the same
    iterations and bit-identical solutions, for FP64 and the mixed solves."""
    dims, U, b = lat
    geom = qamd.Geometry(dims)
    Ud = U.to(DEV)
    A = qamd.clover_from_links(geom, Ud, ALPHA, CSW)
    bf = geom.pack_spinor(b.to(DEV), qamd.F64, 1)
    for prec_s in (qamd.F64, qamd.F16, qamd.BF16):
        op = qamd.EvenOddCloverOperator(geom, Ud, A, prec_sloppy=prec_s, recon_sloppy=12)
        runs = []
        for _ in range(2):
            x = geom.new_spinor(qamd.F64, cb=1)
            res = qamd.InvCG(op, max_iters=5000)(x, bf, 1e-10, mixed=prec_s != qamd.F64)
            runs.append((res["iters"], res["reliable_updates"], x.data.clone()))
        assert runs[0][:2] == runs[1][:2], prec_s
        assert torch.equal(runs[0][2], runs[1][2]), prec_s
        op.close()


def test_workspace_reuse_across_calls(qamd, lat):
    """One operator keeps its CG and force workspaces between calls: a
    mixed FP16 solve, an FP64 solve (larger sloppy vectors), the force
    (which shares the FP64 vectors) and another mixed solve give exactly
    what fresh operators give."""
    dims, U, b = lat
    geom = qamd.Geometry(dims)
    Ud = U.to(DEV)
    A = qamd.clover_from_links(geom, Ud, ALPHA, CSW)
    bf = geom.pack_spinor(b.to(DEV), qamd.F64, 1)

    def solve(op, mixed):
        x = geom.new_spinor(qamd.F64, cb=1)
        qamd.InvCG(op, max_iters=5000)(x, bf, 1e-10, mixed=mixed)
        return x

    def fresh(mixed):
        op = qamd.EvenOddCloverOperator(geom, Ud, A, mu=0.1, prec_sloppy=qamd.F16, recon_sloppy=12)
        x = solve(op, mixed)
        F = op.force(torch.zeros_like(Ud), x, Ud, project=True, csw=CSW) if not mixed else None
        op.close()
        return x.data.clone(), F

    x16, _ = fresh(True)
    x64, F64 = fresh(False)
    op = qamd.EvenOddCloverOperator(geom, Ud, A, mu=0.1, prec_sloppy=qamd.F16, recon_sloppy=12)
    assert torch.equal(solve(op, True).data, x16)
    y = solve(op, False)
    assert torch.equal(y.data, x64)
    assert torch.equal(op.force(torch.zeros_like(Ud), y, Ud, project=True, csw=CSW), F64)
    assert torch.equal(solve(op, True).data, x16)
    op.close()


def test_cg_without_fused_norm(qamd, orc):
    """4x4x4x6: 192 sites per checkerboard, not a whole number of 128-site
    work-groups, so the CG takes <p, M^dag M p> from a separate dot product
    instead of the norm fused into the operator kernel."""
    dims = (4, 4, 4, 6)
    L = qamd.lattice
    U = L.random_gauge(dims, 0.2, 2100)
    b = L.random_spinor(dims, 2101)
    geom = qamd.Geometry(dims)
    Ud = U.to(DEV)
    A = qamd.clover_from_links(geom, Ud, ALPHA, CSW)
    An = A.cpu().numpy()
    bn = on_cb(b.numpy(), dims, 1)
    bf = geom.pack_spinor(b.to(DEV), qamd.F64, 1)
    for prec_s in (qamd.F64, qamd.F16):
        op = qamd.EvenOddCloverOperator(geom, Ud, A, prec_sloppy=prec_s, recon_sloppy=12)
        x = geom.new_spinor(qamd.F64, cb=1)
        res = qamd.InvCG(op, max_iters=5000)(x, bf, 1e-10, mixed=prec_s != qamd.F64)
        assert res["rsd_final"] < 1e-10
        xn = geom.unpack_spinor(x).cpu().numpy()
        r = _true_residual(orc, dims, U.numpy(), An, xn, bn, 0.0)
        assert r < 1.5e-10, (prec_s, r)
        op.close()


def test_cg_reports_nonconvergence(qamd, lat):
    dims, U, b = lat
    geom = qamd.Geometry(dims)
    Ud = U.to(DEV)
    A = qamd.clover_from_links(geom, Ud, ALPHA, CSW)
    op = qamd.EvenOddCloverOperator(geom, Ud, A, prec_sloppy=qamd.F16)
    x = geom.new_spinor(qamd.F64, cb=1)
    bf = geom.pack_spinor(b.to(DEV), qamd.F64, 1)
    res = qamd.InvCG(op, max_iters=5)(x, bf, 1e-10, mixed=True, raise_on_noconv=False)
    assert res["status"] == -5 and res["iters"] == 5 and res["rsd_final"] > 1e-10


def test_reunitarize(qamd):
    """Perturbed links projected back onto SU(3): unitary, det 1, equal to the
    host Gram-Schmidt projection, and exact SU(3) links are left unchanged."""
    dims = (8, 8, 8, 8)
    L = qamd.lattice
    geom = qamd.Geometry(dims)
    U = L.random_gauge(dims, 0.4, 91)
    noisy = U + 1e-3 * torch.randn(U.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    got = qamd.reunitarize(geom, noisy.clone().to(DEV)).cpu()
    ref = L.cplx_to_real(L.reunitarize(L.real_to_cplx(noisy)))
    assert (got - ref).abs().max().item() < 1e-14
    z = L.real_to_cplx(got)
    eye = torch.eye(3, dtype=torch.complex128)
    assert (z @ z.conj().transpose(-1, -2) - eye).abs().max().item() < 1e-14
    assert (torch.linalg.det(z) - 1).abs().max().item() < 1e-14
    same = qamd.reunitarize(geom, U.clone().to(DEV)).cpu()
    assert (same - U).abs().max().item() < 1e-14


def test_block_inverses_of_indefinite_hermitian_blocks(qamd):
    """The block inverses take no positive-definiteness for granted:
    Hermitian blocks with eigenvalues of both signs invert to FP64 accuracy,
    plain and twisted (the kernels form (A - i m)(A^2 + m^2)^-1, whose error
    grows like cond(A)^2 eps; these blocks have cond(A) = 4, clover blocks
    near alpha 1 about 1.5-3)."""
    dims = (4, 4, 4, 4)
    L = qamd.lattice
    geom = qamd.Geometry(dims)
    g = torch.Generator().manual_seed(77)
    V = 256
    q, _ = torch.linalg.qr(torch.randn(V, 2, 6, 6, dtype=torch.complex128, generator=g))
    ev = torch.tensor([3.0, -2.0, 1.0, -1.5, 2.5, -4.0], dtype=torch.float64)
    dense = q @ torch.diag_embed(ev.to(torch.complex128)) @ q.conj().transpose(-1, -2)
    clov = L.clover_pack(0.5 * (dense + dense.conj().transpose(-1, -2)))
    ref = torch.linalg.inv(L.clover_dense(clov))
    got = L.clover_dense(qamd.clover_invert(geom, clov.to(DEV)).cpu())
    assert (got - ref).abs().max().item() / ref.abs().max().item() < 1e-13
    for sign in (1, -1):
        blk = qamd.tmclover_blocks(geom, clov.to(DEV), 0.3, sign, 1).cpu()
        g5 = torch.tensor([1.0, -1.0], dtype=torch.float64).view(1, 2, 1, 1)
        twisted = L.clover_dense(clov) + 1j * sign * 0.3 * g5 * torch.eye(6, dtype=torch.complex128)
        rf = torch.linalg.inv(twisted)
        gz = torch.complex(blk[..., 0], blk[..., 1]).reshape(V, 2, 6, 6)
        assert (gz - rf).abs().max().item() / rf.abs().max().item() < 1e-13


def test_split_after_create_refused(qamd, lat):
    """An operator built on a whole lattice refuses to run once the context
    is marked split (qphix_amd_set_comm): it would otherwise drop the
    wrap-around hops and return a wrong M.  Clearing the mask restores it."""
    import ctypes
    dims, U, b = lat
    geom = qamd.Geometry(dims)
    Ud = U.to(DEV)
    A = qamd.clover_from_links(geom, Ud, ALPHA, CSW)
    op = qamd.EvenOddCloverOperator(geom, Ud, A, prec_sloppy=qamd.F32, recon_sloppy=12)
    bf = geom.pack_spinor(b.to(DEV), qamd.F64, 1)
    ref = geom.new_spinor(qamd.F64, cb=1)
    op.apply(ref, bf)
    lib = qamd.lib()
    geom.set_comm((0, 0, 0, 1))
    out = geom.new_spinor(qamd.F64, cb=1)
    force = torch.zeros_like(Ud)
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib.qphix_amd_eo_op_apply(op._op, qamd.F64, 0, out.ptr(), bf.ptr(), st) == -1
    assert lib.qphix_amd_eo_op_force(op._op, ctypes.c_void_p(force.data_ptr()), bf.ptr(),
                                     ctypes.c_void_p(Ud.data_ptr()), 1.0, 1, st) == -1
    assert lib.qphix_amd_eo_op_force_clover(op._op, ctypes.c_void_p(force.data_ptr()), bf.ptr(),
                                            ctypes.c_void_p(Ud.data_ptr()), CSW, 1.0, 1, st) == -1
    prm = qamd._lib.CGParams(1e-10, 100, 0.0, 0)
    res = qamd._lib.CGResult()
    assert lib.qphix_amd_cg_solve(op._op, out.ptr(), bf.ptr(), ctypes.byref(prm), ctypes.byref(res), st) == -1
    geom.set_comm((0, 0, 0, 0))
    op.apply(out, bf)
    torch.cuda.synchronize()
    assert torch.equal(out.data, ref.data)
    assert torch.count_nonzero(force) == 0
    op.close()
