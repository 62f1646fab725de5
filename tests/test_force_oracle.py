"""Pins the FP64 oracle's link gradient (oracle_force) by finite
differences of the oracle's own L = Re<chi, D psi>:

* su(3) directions: U -> exp(i e T) U with random Hermitian traceless T on
  every link; central differences at e = 1e-5 against sum Re tr(i T F);
* Euclidean directions: U -> U + e E with random complex E (L is linear in
  U, so the central difference is exact up to rounding) against G^dagger;
* the projected force is traceless anti-Hermitian.
Both isigns, both checkerboards, anisotropic coefficients, antiperiodic time
and twisted boundary phases.
"""
import numpy as np
import pytest
import torch

from _util import cplx, on_cb, real

DIMS = (4, 4, 4, 4)


def loss(orc, U, chi, psi, isign, cb, kw):
    out = orc.dslash(DIMS, psi, U, isign, cb, **kw)
    return float(np.sum(chi * out))


def su3_direction(qamd, V, seed):
    g = torch.Generator().manual_seed(seed)
    w = torch.randn(V, 4, 8, generator=g, dtype=torch.float64)
    lam = qamd.lattice.gellmann()
    return torch.einsum("vma,aij->vmij", w.to(torch.complex128), lam).numpy() * 0.5


def expm_i(eps, T):
    w, v = np.linalg.eigh(T)
    return (v * np.exp(1j * eps * w)[..., None, :]) @ np.conj(np.swapaxes(v, -1, -2))


@pytest.mark.parametrize("isign,cb", [(1, 0), (-1, 1), (1, 1)])
def test_oracle_force_su3_finite_differences(qamd, orc, isign, cb):
    L = qamd.lattice
    V = 256
    U = L.random_gauge(DIMS, 0.4, 41).numpy()
    chi = on_cb(L.random_spinor(DIMS, 42).numpy(), DIMS, cb)
    psi = on_cb(L.random_spinor(DIMS, 43).numpy(), DIMS, 1 - cb)
    kw = dict(coeff_s=0.9, coeff_t_f=1.2, coeff_t_b=0.7, t_boundary=-1.0, tbc=[1, 0, 0, 1],
              tbc_phases=[np.exp(0.4j), 1.0, 1.0, np.exp(-0.9j)])
    F = cplx(orc.force(DIMS, chi, psi, U, isign, cb, project=True, **kw))
    # traceless anti-Hermitian
    assert np.abs(F + np.conj(np.swapaxes(F, -1, -2))).max() < 1e-13
    assert np.abs(np.trace(F, axis1=-2, axis2=-1)).max() < 1e-13
    Uc = cplx(U)
    eps = 1e-5
    for k in range(8):
        T = su3_direction(qamd, V, 100 + k)
        up = real(expm_i(eps, T) @ Uc)
        um = real(expm_i(-eps, T) @ Uc)
        fd = (loss(orc, up, chi, psi, isign, cb, kw) - loss(orc, um, chi, psi, isign, cb, kw)) / (2 * eps)
        an = float(np.sum(np.real(1j * np.trace(T @ F, axis1=-2, axis2=-1))))
        assert abs(fd - an) <= 1e-7 * max(abs(an), 1.0), (k, fd, an)


def test_oracle_force_euclidean_gradient(qamd, orc):
    L = qamd.lattice
    U = L.random_gauge(DIMS, 0.4, 51).numpy()
    cb, isign = 1, -1
    chi = on_cb(L.random_spinor(DIMS, 52).numpy(), DIMS, cb)
    psi = on_cb(L.random_spinor(DIMS, 53).numpy(), DIMS, 1 - cb)
    kw = dict(coeff_s=1.1, coeff_t_f=0.8, coeff_t_b=1.3, t_boundary=-1.0)
    G = orc.force(DIMS, chi, psi, U, isign, cb, project=False, **kw)
    rng = np.random.default_rng(5)
    for _ in range(4):
        E = rng.normal(size=U.shape)
        fd = (loss(orc, U + 1e-3 * E, chi, psi, isign, cb, kw)
              - loss(orc, U - 1e-3 * E, chi, psi, isign, cb, kw)) / 2e-3
        an = float(np.sum(G * E))
        assert abs(fd - an) <= 1e-10 * max(abs(an), 1.0)


def test_torch_clover_restatement_equals_oracle_builder(qamd, orc):
    """tests/clover_torch.py (whose autograd is the reference gradient of the
    clover-term force) builds the same A(U) as the oracle, and A is
    Hermitian and block diagonal in spin (blocks {0,1} and {2,3})."""
    import clover_torch as ct
    dims = (4, 4, 6, 4)
    L = qamd.lattice
    U = L.random_gauge(dims, 0.5, 61)
    Uc = torch.complex(U[..., 0], U[..., 1])
    A = ct.clover_matrix(Uc, dims, 3.7, 1.3)
    assert (A - A.conj().transpose(-1, -2)).abs().max().item() < 1e-14
    assert A[:, :6, 6:].abs().max().item() < 1e-15
    ref = orc.clover_from_links(dims, U.numpy(), 3.7, 1.3)
    assert np.abs(ct.pack_blocks(A).numpy() - ref).max() < 1e-14


def test_oracle_sampled_entry_points_equal_whole_lattice(qamd, orc):
    """oracle.force_links and oracle.clover_from_links_sites (used by the
    48^3x96 GPU checks, where whole-lattice oracle calls are too large) give
    exactly the whole-lattice results at the listed links / sites, with
    anisotropic coefficients, antiperiodic time and twisted phases."""
    L = qamd.lattice
    dims = (4, 6, 4, 8)
    U = L.random_gauge(dims, 0.3, 71).numpy()
    chi = L.random_spinor(dims, 72).numpy()
    psi = L.random_spinor(dims, 73).numpy()
    V = U.shape[0]
    rng = np.random.default_rng(7)
    links = rng.choice(4 * V, 200, replace=False)
    kw = dict(coeff_s=0.9, coeff_t_f=1.3, coeff_t_b=0.7, t_boundary=-1.0, tbc=(1, 0, 1, 1),
              tbc_phases=(np.exp(0.3j), 1, np.exp(-0.7j), np.exp(1.1j)))
    for isign, cb, project in ((1, 0, True), (-1, 1, True), (1, 1, False)):
        full = orc.force(dims, chi, psi, U, isign, cb, project=project, **kw).reshape(4 * V, 3, 3, 2)
        part = orc.force_links(dims, links, chi, psi, U, isign, cb, project=project, **kw)
        assert np.array_equal(part, full[links])
    sites = rng.choice(V, 100, replace=False)
    A = orc.clover_from_links(dims, U, 3.7, 1.5)
    assert np.array_equal(orc.clover_from_links_sites(dims, sites, U, 3.7, 1.5), A[sites])
