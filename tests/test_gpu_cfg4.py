"""GPU checks of the large single-GPU workload: 48^3x96 (5,308,416 sites per
checkerboard), twisted-mass clover with the clover term built from the links,
links sigma = 0.2, kappa = 0.1355, csw = 1.5, mu = 0.05, antiperiodic t.

* clover from links: the GPU builder vs the FP64 oracle at 512 sampled sites;
* tm_clov_dslash (isign +-1, cb 0, the blocks of (A + i s mu g5)^-1) and
  tm_clov_dslash_achimbdpsi (isign +-1, cb 1, A + i s mu g5, beta 1/4) in FP32
  and FP16 storage (12-real links) vs the FP64 oracle at 4096 sampled sites,
  the block inverses formed by the oracle from the GPU's clover field;
* FP64 CG and FP16-mixed reliable-update CG on M^dag M of the even-odd
  operator to a true relative residual of 1e-10, each solution's residual
  recomputed outside the solver (b - M^dag M x through the FP64 operator),
  and the two solutions' agreement;
* the pseudofermion force at the solutions: from the mixed solution against
  the one from the FP64 solution; the hopping term of the FP64 force against
  the oracle's link gradient at 96 sampled links, with Y = M X, Z = F_ee^-1
  D_eo X and W = (F_ee^-1)^dag D_oe^dag Y recomputed by the oracle at every
  site those links need; the clover term against central finite differences
  of its functional -2 Re<Y, A_oo X> - 1/2 Re<W, A_ee Z> along random su(3)
  directions on 8 links.

Whole-lattice host copies of these fields would take tens of GB, so the
oracle reads full-size host arrays that hold only the rows its sampled
sites touch (numpy leaves the pages it never writes uncommitted).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DIMS = (48, 48, 48, 96)
V = DIMS[0] * DIMS[1] * DIMS[2] * DIMS[3]
KAPPA, MU, CSW, SIGMA = 0.1355, 0.05, 1.5, 0.2
ALPHA = 1.0 / (2 * KAPPA)


def coords(idx):
    Lx, Ly, Lz, _ = DIMS
    idx = np.asarray(idx, dtype=np.int64)
    return np.stack([idx % Lx, (idx // Lx) % Ly, (idx // (Lx * Ly)) % Lz, idx // (Lx * Ly * Lz)])


def lex(c):
    Lx, Ly, Lz, _ = DIMS
    return ((c[3] * Lz + c[2]) * Ly + c[1]) * Lx + c[0]


def parity_of(idx):
    return coords(idx).sum(axis=0) & 1


def shift(idx, mu, d):
    """Sites idx + d * mu-hat (mu an int or an array like idx), periodic."""
    c = coords(idx)
    mu = np.broadcast_to(np.asarray(mu), c.shape[1:])
    L = np.asarray(DIMS)[mu]
    sel = np.arange(4)[:, None] == mu[None, :]
    c = np.where(sel, (c + d) % L, c)
    return lex(c)


def nbrs(idx):
    return np.unique(np.concatenate([shift(idx, m, d) for m in range(4) for d in (1, -1)]))


def sample_cb(rng, cb, n):
    """n distinct random lexicographic sites of checkerboard cb."""
    Lx, Ly, Lz, Lt = DIMS
    out = set()
    while len(out) < n:
        xh, y, z, t = (rng.integers(0, L) for L in (Lx // 2, Ly, Lz, Lt))
        out.add(int(lex(np.array([2 * xh + ((cb + y + z + t) & 1), y, z, t]))))
    return np.array(sorted(out), dtype=np.int64)


def host_rows(t, idx):
    """A full-lattice host array with t's rows idx and zeros elsewhere."""
    out = np.zeros(tuple(t.shape), dtype=np.float64)
    ii = np.unique(np.asarray(idx, dtype=np.int64))
    out[ii] = t[torch.from_numpy(ii).to(t.device)].cpu().numpy()
    return out


def scattered(rows, idx, shape):
    out = np.zeros(shape, dtype=np.float64)
    out[idx] = rows
    return out


def oracle_blocks(orc, A, sites, sign, invert):
    """Full-lattice host array of the twisted-mass clover blocks (A + i sign mu
    g5) or its inverse at `sites`, formed by the oracle from the GPU's A."""
    rows = orc.tmclover_blocks(A[torch.from_numpy(sites).to(A.device)].cpu().numpy(), MU, sign, invert)
    return scattered(rows, sites, (V, 2, 6, 6, 2))


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


@pytest.fixture(scope="module")
def lat(qamd):
    L = qamd.lattice
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    print("cfg4 lattice 48^3x96: links, clover from links, spinors", flush=True)
    U = L.random_gauge(DIMS, SIGMA, 4000, device=DEV)
    A = qamd.clover_from_links(geom, U, ALPHA, CSW)
    phi = L.random_spinor(DIMS, 4001, device=DEV)
    chi = L.random_spinor(DIMS, 4002, device=DEV)
    torch.cuda.synchronize()
    yield dict(geom=geom, U=U, A=A, phi=phi, chi=chi)
    torch.cuda.empty_cache()


def test_cfg4_clover_from_links_sampled(qamd, orc, lat):
    rng = np.random.default_rng(21)
    sites = np.sort(rng.choice(V, 512, replace=False))
    near = nbrs(np.unique(np.concatenate([sites, nbrs(sites)])))  # every plaquette corner
    Uh = host_rows(lat["U"], np.concatenate([sites, near]))
    ref = orc.clover_from_links_sites(DIMS, sites, Uh, ALPHA, CSW)
    got = lat["A"][torch.from_numpy(sites).to(DEV)].cpu().numpy()
    assert rel(got, ref) < 1e-13


@pytest.mark.parametrize("prec,tol", [(1, 1e-6), (2, 5e-3)], ids=["f32", "f16"])
def test_cfg4_tmclover_dslash_sampled(qamd, orc, lat, prec, tol):
    geom, U, A = lat["geom"], lat["U"], lat["A"]
    rng = np.random.default_rng(22 + prec)
    op = qamd.TMClovDslash(geom)
    for achi, isign in ((False, 1), (False, -1), (True, 1), (True, -1)):
        cb = 1 if achi else 0
        print(f"case cfg4 prec={prec} achi={achi} isign={isign}", flush=True)
        sites = sample_cb(rng, cb, 4096)
        blocks = qamd.tmclover_blocks(geom, A, MU, isign, not achi)
        g = geom.pack_gauge(U, prec, 12, cb)
        cl = geom.pack_full_clover(blocks, prec, cb)
        del blocks
        psi = geom.pack_spinor(lat["phi"], prec, 1 - cb)
        res = geom.new_spinor(prec, cb=cb)
        if achi:
            op.dslashAChiMinusBDPsi(res, psi, geom.pack_spinor(lat["chi"], prec, cb), g, cl, 0.25, isign, cb)
        else:
            op.dslash(res, psi, g, cl, isign, cb)
        got = geom.unpack_spinor(res)[torch.from_numpy(sites).to(DEV)].cpu().numpy()
        del g, cl, psi, res
        torch.cuda.empty_cache()
        c = 0.25 if achi else 1.0
        kw = dict(kind=3, achi=achi, coeff_s=c, coeff_t_f=c, coeff_t_b=c, t_boundary=-1.0, rho=0.0,
                  clov=oracle_blocks(orc, A, sites, isign, not achi))
        if achi:
            kw["chi"] = host_rows(lat["chi"], sites)
        ref = orc.dslash_sites(DIMS, sites, host_rows(lat["phi"], nbrs(sites)),
                               host_rows(U, np.concatenate([sites, nbrs(sites)])), isign, **kw)
        err = rel(got, ref)
        print(f"  max error / max |ref| = {err:.3e}", flush=True)
        assert err < tol, (prec, achi, isign, err)


def su3_generator(rng):
    """A random traceless Hermitian 3x3 matrix (complex128)."""
    m = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
    h = 0.5 * (m + m.conj().T)
    return h - np.trace(h) / 3 * np.eye(3)


def dense_blocks(packed):
    """(n, 2, 36) packed Hermitian blocks -> (n, 2, 6, 6) complex."""
    n = packed.shape[0]
    out = torch.zeros(n, 2, 6, 6, dtype=torch.complex128, device=packed.device)
    for r in range(6):
        out[:, :, r, r] = packed[:, :, r].to(torch.complex128)
        for c in range(r):
            od = r * (r - 1) // 2 + c
            v = torch.complex(packed[:, :, 6 + 2 * od], packed[:, :, 6 + 2 * od + 1])
            out[:, :, r, c] = v
            out[:, :, c, r] = v.conj()
    return out


def test_cfg4_cg_and_force(qamd, orc, lat):
    geom, U, A, phi = lat["geom"], lat["U"], lat["A"], lat["phi"]
    F64 = qamd.F64
    b = geom.pack_spinor(phi, F64, 1)
    # the FP64 solve uses only the operator's FP64 copy; its sloppy copy is
    # FP32 (not a second FP64 one) to keep the memory of this workload down
    op64 = qamd.EvenOddCloverOperator(geom, U, A, mu=MU, prec_sloppy=qamd.F32, recon_sloppy=12)
    sols = {}
    for name, prec in (("f64", F64), ("f16", qamd.F16)):
        op = op64 if prec == F64 else qamd.EvenOddCloverOperator(geom, U, A, mu=MU, prec_sloppy=prec,
                                                                   recon_sloppy=12)
        x = geom.new_spinor(F64, cb=1)
        res = qamd.InvCG(op, max_iters=2000)(x, b, 1e-10, mixed=prec != F64)
        print(f"cfg4 CG {name}: {res['iters']} iterations, {res['reliable_updates']} reliable updates, "
              f"{res['seconds']:.3f} s, rsd {res['rsd_final']:.3e}", flush=True)
        if op is not op64:
            op.close()
        # the true residual, recomputed outside the solver's loop
        mx, r = geom.new_spinor(F64, cb=1), geom.new_spinor(F64, cb=1)
        op64.apply(mx, x)
        op64.apply(r, mx, dagger=True)
        rsd = float(((b.data - r.data).norm() / b.data.norm()).item())
        print(f"  recomputed true residual {rsd:.3e}", flush=True)
        assert rsd < 1e-10
        assert abs(rsd - res["rsd_final"]) <= 1e-3 * rsd
        sols[name] = x
    x64, x16 = sols["f64"], sols["f16"]
    dx = float(((x16.data - x64.data).norm() / x64.data.norm()).item())
    print(f"cfg4 mixed vs FP64 solution: {dx:.3e}", flush=True)
    assert dx < 1e-8

    # --- force: mixed vs FP64 solution
    F_full = op64.force(torch.zeros_like(U), x64, U, project=True, csw=CSW)
    F_mixed = op64.force(torch.zeros_like(U), x16, U, project=True, csw=CSW)
    df = float(((F_mixed - F_full).abs().max() / F_full.abs().max()).item())
    print(f"cfg4 force from the mixed solution vs from the FP64 one: {df:.3e}", flush=True)
    assert df < 1e-7
    del F_mixed, x16
    F_hop = op64.force(torch.zeros_like(U), x64, U, project=True)
    torch.cuda.empty_cache()

    # --- the hopping term against the oracle at sampled links
    rng = np.random.default_rng(31)
    ys = rng.choice(V, 96, replace=False)
    mus = rng.integers(0, 4, len(ys))
    links = 4 * ys + mus
    s0 = np.unique(np.concatenate([ys, shift(ys, mus, 1)]))
    s0e, s0o = s0[parity_of(s0) == 0], s0[parity_of(s0) == 1]
    ysites = np.unique(np.concatenate([s0o, nbrs(s0e)]))            # odd
    zsites = np.unique(np.concatenate([s0e, nbrs(ysites)]))         # even
    xsites = np.unique(np.concatenate([ysites, nbrs(zsites)]))      # odd
    usites = np.unique(np.concatenate([ys, zsites, ysites, nbrs(zsites), nbrs(ysites)]))
    print(f"cfg4 force oracle chain: {len(links)} links, sites Y {len(ysites)} Z {len(zsites)} "
          f"X {len(xsites)}", flush=True)
    Uh = host_rows(U, usites)
    Xh = host_rows(geom.unpack_spinor(x64), xsites)
    Zr = orc.dslash_sites(DIMS, zsites, Xh, Uh, 1, kind=3, t_boundary=-1.0,
                          clov=oracle_blocks(orc, A, zsites, 1, True))
    Zh = scattered(Zr, zsites, (V, 4, 3, 2))
    Yr = orc.dslash_sites(DIMS, ysites, Zh, Uh, 1, kind=3, achi=True, chi=Xh, coeff_s=0.25,
                          coeff_t_f=0.25, coeff_t_b=0.25, t_boundary=-1.0, rho=0.0,
                          clov=oracle_blocks(orc, A, ysites, 1, False))
    Yh = scattered(Yr, ysites, (V, 4, 3, 2))
    Wr = orc.dslash_sites(DIMS, s0e, Yh, Uh, -1, kind=3, t_boundary=-1.0,
                          clov=oracle_blocks(orc, A, s0e, -1, True))
    Wh = scattered(Wr, s0e, (V, 4, 3, 2))
    ref = 0.5 * (orc.force_links(DIMS, links, Yh, Zh, Uh, 1, 1, t_boundary=-1.0)
                 + orc.force_links(DIMS, links, Wh, Xh, Uh, 1, 0, t_boundary=-1.0))
    got = F_hop.reshape(V * 4, 3, 3, 2)[torch.from_numpy(links).to(DEV)].cpu().numpy()
    err = rel(got, ref)
    print(f"cfg4 hopping force vs oracle at {len(links)} links: {err:.3e}", flush=True)
    assert err < 1e-11

    # --- Y, Z, W on the GPU (FP64), as the force forms them, vs the oracle
    tmk = qamd.TMClovDslash(geom)
    g0 = geom.pack_gauge(U, F64, 12, 0)
    Yf, Zf, Wf = geom.new_spinor(F64, cb=1), geom.new_spinor(F64, cb=0), geom.new_spinor(F64, cb=0)
    op64.apply(Yf, x64)
    blk = geom.pack_full_clover(qamd.tmclover_blocks(geom, A, MU, 1, True), F64, 0)
    tmk.dslash(Zf, x64, g0, blk, 1, 0)
    blk = geom.pack_full_clover(qamd.tmclover_blocks(geom, A, MU, -1, True), F64, 0)
    tmk.dslash(Wf, Yf, g0, blk, -1, 0)
    del blk, g0
    torch.cuda.empty_cache()
    Yl, Wl = geom.unpack_spinor(Yf), geom.unpack_spinor(Wf)
    Xl, Zl = geom.unpack_spinor(x64), geom.unpack_spinor(Zf)
    for name, t, sites, r in (("Y", Yl, ysites, Yr), ("Z", Zl, zsites, Zr), ("W", Wl, s0e, Wr)):
        e = rel(t[torch.from_numpy(sites).to(DEV)].cpu().numpy(), r)
        assert e < 1e-12, (name, e)
    chi_l, psi_l = Yl + Wl, Xl + Zl
    del Yl, Wl, Xl, Zl, Yf, Zf, Wf

    # --- the clover term against central finite differences of its functional
    F_clov = (F_full - F_hop).reshape(V, 4, 3, 3, 2)
    del F_full, F_hop
    eps = 1e-4
    fd, an = [], []
    for k in range(8):
        y, mu = int(rng.integers(0, V)), int(rng.integers(0, 4))
        T = su3_generator(rng)
        u0 = U[y, mu].clone()
        u0c = torch.complex(u0[..., 0], u0[..., 1])
        vals = []
        for sgn in (1, -1):
            E = torch.linalg.matrix_exp(torch.from_numpy(1j * sgn * eps * T).to(DEV))
            m = E @ u0c
            U[y, mu] = torch.stack([m.real, m.imag], dim=-1)
            vals.append(qamd.clover_from_links(geom, U, ALPHA, CSW))
        U[y, mu] = u0
        dA = vals[0] - vals[1]
        del vals
        nz = torch.nonzero(dA.abs().amax(dim=(1, 2)) > 0).squeeze(1)
        w = torch.from_numpy(np.where(parity_of(nz.cpu().numpy()) == 1, -2.0, -0.5)).to(DEV)
        D = dense_blocks(dA[nz])
        chi = torch.complex(chi_l[nz][..., 0], chi_l[nz][..., 1]).reshape(-1, 2, 6)
        psi = torch.complex(psi_l[nz][..., 0], psi_l[nz][..., 1]).reshape(-1, 2, 6)
        dL = torch.einsum("nbr,nbrc,nbc->n", chi.conj(), D, psi).real
        fd.append(float((w * dL).sum().item()) / (2 * eps))
        Fk = F_clov[y, mu].cpu().numpy()
        Fk = Fk[..., 0] + 1j * Fk[..., 1]
        an.append(float(np.real(np.trace(1j * T @ Fk))))
        del dA, D
    torch.cuda.synchronize()
    fd, an = np.array(fd), np.array(an)
    err = float(np.abs(fd - an).max() / np.abs(an).max())
    print(f"cfg4 clover force vs finite differences on 8 links: {err:.3e} "
          f"(largest derivative {np.abs(an).max():.3e})", flush=True)
    assert err < 1e-6
    op64.close()
