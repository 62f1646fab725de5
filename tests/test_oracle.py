"""Pins the FP64 C oracle (oracle/qphix_oracle.c) on the CPU.

Each test is chosen so that a plausible mistake in the oracle (a dropped
direction, a wrong sign of isign, a transposed link, a wrong boundary
factor, a wrong gamma matrix) fails at least one of them:

* the gamma matrices the oracle uses equal the ones implied by the
  reference's own projector/reconstruct tables (codegen/dslash_common.cc:748-890);
* free field: D on a plane wave has the closed form sum_mu [2 cos p - 2 i s g_mu sin p];
* gamma_5 hermiticity: D_+^dag = D_- = g5 D_+ g5, per operator family;
* gauge covariance: D[U^g](g psi) = g D[U] psi;
* twisted BC phase e^{i theta} per hop equals a U(1) rotation of psi;
* a brute-force dense matrix built with numpy translations on 2^4 and 4^4;
* clover with identity blocks = Wilson; TM with mu = 0 = hop/alpha;
  A A^-1 = 1; (A + i mu g5) B = 1;
* clover from links: unit gauge -> alpha*1, gauge covariance, and the closed
  form of a constant abelian field strength.
"""
import numpy as np
import pytest
import torch

from _util import GAMMA, GAMMA5, cb_mask, coords, cplx, g5, inner, on_cb, real, rel_err

DIMS = (4, 4, 4, 8)


def _lat(qamd, dims=DIMS, sigma=0.3, seed=1):
    L = qamd.lattice
    U = L.random_gauge(dims, sigma, seed).numpy()
    psi = L.random_spinor(dims, seed + 1).numpy()
    chi = L.random_spinor(dims, seed + 2).numpy()
    return U, psi, chi


# ---------------------------------------------------------------------------
# The reference's projector and reconstruct tables, as data
# (codegen/dslash_common.cc:748-766 proj_ops_plus/minus,
#  :784-836 rec_plus_pbeta_ops / rec_minus_pbeta_ops).
# addCVec -> +1, subCVec -> -1, addiCVec -> +i, subiCVec -> -i.
PROJ_PLUS = [((0, 3), (1, 2), 1j, 1j), ((0, 3), (1, 2), -1j, -1j),
             ((0, 3), (1, 2), -1, 1), ((0, 3), (1, 2), 1, -1),
             ((0, 2), (1, 3), 1j, -1j), ((0, 2), (1, 3), -1j, 1j),
             ((0, 2), (1, 3), 1, 1), ((0, 2), (1, 3), -1, -1)]
PROJ_MINUS = [((0, 3), (1, 2), -1j, -1j), ((0, 3), (1, 2), 1j, 1j),
              ((0, 3), (1, 2), 1, -1), ((0, 3), (1, 2), -1, 1),
              ((0, 2), (1, 3), -1j, 1j), ((0, 2), (1, 3), 1j, -1j),
              ((0, 2), (1, 3), -1, -1), ((0, 2), (1, 3), 1, 1)]
# (s2, s3, Func1 coefficient into spin 2, Func2 coefficient into spin 3)
REC_PLUS = [(1, 0, -1j, -1j), (1, 0, 1, -1), (0, 1, -1j, 1j), (0, 1, 1, 1)]
REC_MINUS = [(1, 0, 1j, 1j), (1, 0, -1, 1), (0, 1, 1j, -1j), (0, 1, -1, -1)]


def _table_matrix(proj, rec):
    (a0, b0), (a1, b1), c0, c1 = proj
    H = np.zeros((2, 4), complex)
    H[0, a0], H[0, b0] = 1, c0
    H[1, a1], H[1, b1] = 1, c1
    s2, s3, f1, f2 = rec
    Rm = np.zeros((4, 2), complex)
    Rm[0, 0] = Rm[1, 1] = 1
    Rm[2, s2] = f1
    Rm[3, s3] = f2
    return Rm @ H


def test_gamma_basis_from_reference_tables():
    one = np.eye(4)
    for mu in range(4):
        # isign = +1: back hop uses proj_ops_plus[2mu] + rec_plus, forward
        # proj_ops_plus[2mu+1] + rec_minus (codegen/dslash_common.cc:1789-1799)
        assert np.allclose(_table_matrix(PROJ_PLUS[2 * mu], REC_PLUS[mu]), one + GAMMA[mu])
        assert np.allclose(_table_matrix(PROJ_PLUS[2 * mu + 1], REC_MINUS[mu]), one - GAMMA[mu])
        assert np.allclose(_table_matrix(PROJ_MINUS[2 * mu], REC_MINUS[mu]), one - GAMMA[mu])
        assert np.allclose(_table_matrix(PROJ_MINUS[2 * mu + 1], REC_PLUS[mu]), one + GAMMA[mu])
    for mu in range(4):
        assert np.allclose(GAMMA[mu], GAMMA[mu].conj().T)
        for nu in range(4):
            ac = GAMMA[mu] @ GAMMA[nu] + GAMMA[nu] @ GAMMA[mu]
            assert np.allclose(ac, 2 * (mu == nu) * one)
    assert np.allclose(GAMMA5, np.diag([1, 1, -1, -1]))


@pytest.mark.parametrize("isign", [1, -1])
@pytest.mark.parametrize("cb", [0, 1])
def test_free_field_plane_wave(orc, isign, cb):
    dims = DIMS
    Lx, Ly, Lz, Lt = dims
    rng = np.random.default_rng(5)
    p = np.array([2 * np.pi * 1 / Lx, 2 * np.pi * 3 / Ly, 2 * np.pi * 2 / Lz, np.pi * 3 / Lt])
    u = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    X = coords(dims)
    ph = np.exp(1j * X @ p)
    psi = real(ph[:, None, None] * u[None])
    V = Lx * Ly * Lz * Lt
    U = np.zeros((V, 4, 3, 3, 2))
    U[..., 0] = np.eye(3)
    out = orc.dslash(dims, psi, U, isign, cb, t_boundary=-1.0)  # antiperiodic t
    K = sum(2 * np.cos(p[m]) * np.eye(4) - 2j * isign * np.sin(p[m]) * GAMMA[m] for m in range(4))
    expect = real(ph[:, None, None] * (K @ u)[None])
    m = cb_mask(dims, cb)
    assert rel_err(out[m], expect[m]) < 1e-13
    assert np.all(out[~m] == 0)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_gamma5_hermiticity(qamd, orc, kind):
    """D_+^dag = D_- = g5 D_+ g5 (gamma_mu Hermitian), so for each family
    <chi, E_+ D_+ psi> = <D_- E_+^dag chi, psi> with E the epilogue."""
    U, psi, chi = _lat(qamd)
    L = qamd.lattice
    clov = L.random_clover(DIMS, 7)
    psi1, chi0 = on_cb(psi, DIMS, 1), on_cb(chi, DIMS, 0)
    if kind == orc.WILSON:
        a = inner(chi0, orc.dslash(DIMS, psi1, U, 1, 0))
        t = chi0
    elif kind == orc.CLOVER:
        a = inner(chi0, orc.dslash(DIMS, psi1, U, 1, 0, kind=kind, clov=clov.numpy()))
        t = orc.clover_apply(DIMS, clov.numpy(), chi0, 0)  # packed blocks are Hermitian
    elif kind == orc.TM:
        a = inner(chi0, orc.dslash(DIMS, psi1, U, 1, 0, kind=kind, mu=0.3, muinv=0.7))
        z = cplx(chi0)
        t = real(0.7 * (z + 1j * 0.3 * np.einsum("ab,vbc->vac", GAMMA5, z)))
    else:
        Bp = L.tmclover_blocks(clov, 0.2, +1, True).numpy()
        Bm = L.tmclover_blocks(clov, 0.2, -1, True).numpy()  # = Bp^dag
        a = inner(chi0, orc.dslash(DIMS, psi1, U, 1, 0, kind=kind, clov=Bp))
        t = orc.clover_apply(DIMS, Bm, chi0, 0, full=True)
    b = inner(orc.dslash(DIMS, t, U, -1, 1), psi1)
    assert abs(a - b) / abs(a) < 1e-13
    # D_- = g5 D_+ g5
    m = orc.dslash(DIMS, psi1, U, -1, 0)
    p = g5(orc.dslash(DIMS, g5(psi1), U, 1, 0))
    assert rel_err(m, p) < 1e-15


def test_gauge_covariance(qamd, orc):
    U, psi, _ = _lat(qamd)
    L = qamd.lattice
    V = psi.shape[0]
    g = L.cplx_to_real(L.su3_from_algebra(torch.randn(V, 8, dtype=torch.float64) * 0.8)).numpy()
    gc = cplx(g)
    Uc = cplx(U)
    X = coords(DIMS)
    Lv = np.array(DIMS)
    Ug = np.empty_like(Uc)
    for mu in range(4):
        Y = X.copy()
        Y[:, mu] = (Y[:, mu] + 1) % Lv[mu]
        idx = ((Y[:, 3] * DIMS[2] + Y[:, 2]) * DIMS[1] + Y[:, 1]) * DIMS[0] + Y[:, 0]
        Ug[:, mu] = gc @ Uc[:, mu] @ gc[idx].conj().transpose(0, 2, 1)
    gpsi = real(np.einsum("vij,vsj->vsi", gc, cplx(psi)))
    for cb in (0, 1):
        a = orc.dslash(DIMS, gpsi, real(Ug), 1, cb)
        b = real(np.einsum("vij,vsj->vsi", gc, cplx(orc.dslash(DIMS, psi, U, 1, cb))))
        m = cb_mask(DIMS, cb)
        assert rel_err(a[m], b[m]) < 1e-13


@pytest.mark.parametrize("mu", [0, 1, 2, 3])
def test_twisted_bc_is_u1_rotation(qamd, orc, mu):
    U, psi, _ = _lat(qamd)
    Lm = DIMS[mu]
    theta = 2 * np.pi / Lm
    tbc = [m == mu for m in range(4)]
    phases = [np.exp(1j * theta) if m == mu else 1.0 for m in range(4)]
    X = coords(DIMS)
    om = np.exp(-1j * theta * X[:, mu])  # Omega = diag(e^{-i theta x_mu})
    # D_phi psi = Omega D Omega^dag psi
    a = orc.dslash(DIMS, psi, U, 1, 0, tbc=tbc, tbc_phases=phases, t_boundary=-1.0)
    w = real(np.conj(om)[:, None, None] * cplx(psi))
    b = real(om[:, None, None] * cplx(orc.dslash(DIMS, w, U, 1, 0, t_boundary=-1.0)))
    m = cb_mask(DIMS, 0)
    assert rel_err(a[m], b[m]) < 1e-13
    # theta = 0 phase equals the plain operator
    c = orc.dslash(DIMS, psi, U, 1, 0, tbc=tbc, tbc_phases=[1.0] * 4)
    assert rel_err(c[m], orc.dslash(DIMS, psi, U, 1, 0)[m]) == 0.0


@pytest.mark.parametrize("dims", [(2, 2, 2, 2), (4, 4, 4, 4)], ids=["2x4", "4x4"])
def test_dense_bruteforce(qamd, orc, dims):
    """The oracle against a dense matrix built site by site with numpy
    translations: on 2^4 the forward and backward neighbours of a direction
    are one site, on 4^4 (3072 x 3072) all eight are distinct."""
    L = qamd.lattice
    U = L.random_gauge(dims, 0.5, 3).numpy()
    psi = L.random_spinor(dims, 4).numpy()
    V = int(np.prod(dims))
    X = coords(dims)
    Uc = cplx(U)
    Lv = np.array(dims)
    cs, ctf, ctb, tb = 0.7, 1.3, 0.9, -1.0
    for isign in (1, -1):
        D = np.zeros((V * 12, V * 12), complex)
        for x in range(V):
            for mu in range(4):
                for fb in (0, 1):
                    Y = X[x].copy()
                    Y[mu] = (Y[mu] + (1 if fb else -1)) % Lv[mu]
                    y = ((Y[3] * dims[2] + Y[2]) * dims[1] + Y[1]) * dims[0] + Y[0]
                    if mu < 3:
                        beta = cs
                    elif fb:
                        beta = ctf * (tb if X[x][3] == dims[3] - 1 else 1)
                    else:
                        beta = ctb * (tb if X[x][3] == 0 else 1)
                    S = np.eye(4) + (isign if fb == 0 else -isign) * GAMMA[mu]
                    C = Uc[x, mu] if fb else Uc[y, mu].conj().T
                    D[x * 12:(x + 1) * 12, y * 12:(y + 1) * 12] += beta * np.kron(S, C)
        dense = (D @ cplx(psi).reshape(-1)).reshape(V, 4, 3)
        for cb in (0, 1):
            src = on_cb(psi, dims, 1 - cb)
            out = orc.dslash(dims, src, U, isign, cb, coeff_s=cs, coeff_t_f=ctf, coeff_t_b=ctb,
                             t_boundary=tb)
            d = (D @ cplx(src).reshape(-1)).reshape(V, 4, 3)
            m = cb_mask(dims, cb)
            assert rel_err(out[m], real(d)[m]) < 1e-13
        assert dense.shape == (V, 4, 3)


def test_clover_identity_and_tm_mu0(qamd, orc):
    U, psi, chi = _lat(qamd)
    V = psi.shape[0]
    ident = np.zeros((V, 2, 36))
    ident[:, :, :6] = 1.0
    w = orc.dslash(DIMS, psi, U, 1, 0)
    c = orc.dslash(DIMS, psi, U, 1, 0, kind=orc.CLOVER, clov=ident)
    assert rel_err(c, w) == 0.0
    alpha = 4.2
    t = orc.dslash(DIMS, psi, U, -1, 0, kind=orc.TM, mu=0.0, muinv=1.0 / alpha)
    m = cb_mask(DIMS, 0)
    assert rel_err(t[m], w[m] * 0 + orc.dslash(DIMS, psi, U, -1, 0)[m] / alpha) < 1e-15
    # achimbdpsi forms with trivial epilogues reduce to alpha chi - D psi
    a = orc.dslash(DIMS, psi, U, 1, 0, achi=True, chi=chi, alpha=alpha)
    b = orc.dslash(DIMS, psi, U, 1, 0, kind=orc.TM, achi=True, chi=chi, alpha=alpha, mu=0.0)
    assert rel_err(a[m], alpha * chi[m] - w[m]) < 1e-14
    assert rel_err(b[m], a[m]) < 1e-15


def test_clover_inverse_and_tm_blocks(qamd, orc):
    L = qamd.lattice
    clov = L.random_clover((2, 2, 2, 4), 9).numpy()
    inv = orc.clover_invert(clov)
    A = L.clover_dense(torch.from_numpy(clov)).numpy()
    Ai = L.clover_dense(torch.from_numpy(inv)).numpy()
    assert np.abs(A @ Ai - np.eye(6)).max() < 1e-13
    assert np.abs(Ai - L.clover_dense(L.invert_clover(torch.from_numpy(clov))).numpy()).max() < 1e-13
    for sign in (1, -1):
        B = cplx(orc.tmclover_blocks(clov, 0.1, sign, True))
        Bf = cplx(orc.tmclover_blocks(clov, 0.1, sign, False))
        g5b = np.array([1.0, -1.0])[None, :, None, None]
        assert np.abs((A + 1j * sign * 0.1 * g5b * np.eye(6)) @ B - np.eye(6)).max() < 1e-13
        assert np.abs(Bf - (A + 1j * sign * 0.1 * g5b * np.eye(6))).max() < 1e-15


def test_clover_achimbdpsi_rho_sign(qamd, orc):
    U, psi, chi = _lat(qamd)
    V = psi.shape[0]
    ident = np.zeros((V, 2, 36))
    ident[:, :, :6] = 1.0
    m = cb_mask(DIMS, 0)
    for isign in (1, -1):
        a = orc.dslash(DIMS, psi, U, isign, 0, kind=orc.CLOVER, achi=True, chi=chi, clov=ident, rho=0.25)
        d = orc.dslash(DIMS, psi, U, isign, 0)
        z = cplx(chi) + 1j * isign * 0.25 * np.einsum("ab,vbc->vac", GAMMA5, cplx(chi))
        assert rel_err(a[m], (real(z) - d)[m]) < 1e-14


def _sigma(mu, nu):
    return 0.5j * (GAMMA[mu] @ GAMMA[nu] - GAMMA[nu] @ GAMMA[mu])


def test_clover_from_links(qamd, orc):
    L = qamd.lattice
    dims = (4, 4, 4, 4)
    V = 256
    alpha, csw = 3.7, 1.3
    # unit gauge: A = alpha
    c = orc.clover_from_links(dims, L.unit_gauge(dims).numpy(), alpha, csw)
    A = L.clover_dense(torch.from_numpy(c)).numpy()
    assert np.abs(A - alpha * np.eye(6)).max() < 1e-15
    # constant abelian field in the x-y plane: U_x = diag(e^{-iBy q}), q = (1,-1,0)
    B = 2 * np.pi / dims[1]
    X = coords(dims)
    U = np.zeros((V, 4, 3, 3), complex)
    U[:] = np.eye(3)
    q = np.array([1.0, -1.0, 0.0])
    for v in range(V):
        U[v, 0] = np.diag(np.exp(-1j * B * X[v, 1] * q))
    c = orc.clover_from_links(dims, real(U), alpha, csw)
    A = L.clover_dense(torch.from_numpy(c)).numpy()
    F = np.diag(np.sin(B * q))
    full = alpha * np.eye(12) + 0.5 * csw * np.kron(_sigma(0, 1), F)
    for b in range(2):
        assert np.abs(A[:, b] - full[6 * b:6 * b + 6, 6 * b:6 * b + 6][None]).max() < 1e-14
    # off-block part of sigma_xy vanishes (the clover is block diagonal)
    assert np.abs(full[:6, 6:]).max() == 0
    # gauge covariance: A[U^g](x) = g(x) A[U](x) g(x)^dag (colour part)
    Ur = L.random_gauge(dims, 0.4, 21).numpy()
    g = L.su3_from_algebra(torch.randn(V, 8, dtype=torch.float64)).numpy()
    Uc = cplx(Ur)
    Ug = np.empty_like(Uc)
    Lv = np.array(dims)
    for mu in range(4):
        Y = X.copy()
        Y[:, mu] = (Y[:, mu] + 1) % Lv[mu]
        idx = ((Y[:, 3] * dims[2] + Y[:, 2]) * dims[1] + Y[:, 1]) * dims[0] + Y[:, 0]
        Ug[:, mu] = g @ Uc[:, mu] @ g[idx].conj().transpose(0, 2, 1)
    A0 = L.clover_dense(torch.from_numpy(orc.clover_from_links(dims, Ur, alpha, csw))).numpy()
    A1 = L.clover_dense(torch.from_numpy(orc.clover_from_links(dims, real(Ug), alpha, csw))).numpy()
    G = np.zeros((V, 6, 6), complex)
    G[:, :3, :3] = g
    G[:, 3:, 3:] = g
    assert np.abs(A1 - G[:, None] @ A0 @ G[:, None].conj().transpose(0, 1, 3, 2)).max() < 1e-13
