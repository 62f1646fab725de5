"""GPU parity of every dslash family against the FP64 oracle.

8^4 (2048 sites/cb), every family x {dslash, achimbdpsi} x isign +-1 x cb 0/1
x 12/18 links, in FP64 (tolerance 1e-13 relative) and FP32 (1e-6 relative);
FP16 and BF16 storage with FP32 arithmetic at tolerances set by the storage
rounding (FP16 unit roundoff 2^-11, BF16 2^-8).  Relative error = max-abs
difference over all checkerboard sites / max-abs oracle value.

32^3x64 (the benchmark lattice): the oracle at 4096 sampled sites, 12- vs
18-link agreement, and D_+^dag = D_- via FP64 inner products.
"""
import numpy as np
import pytest
import torch

from _util import cb_mask, inner, on_cb, parity

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
TOL = {0: 1e-13, 1: 1e-6, 2: 5e-3, 3: 4e-2}
ALPHA = 1.0 / (2 * 0.1350)   # 4 + m at kappa = 0.1350
MU = 0.1                      # twisted mass
BETA = 0.25                   # b in a chi - b D psi
RHO = 0.05                    # prec_mass_rho


def make_lattice(qamd, dims, seed, device="cpu"):
    L = qamd.lattice
    U = L.random_gauge(dims, 0.3, seed, device)
    psi = L.random_spinor(dims, seed + 1, device)
    chi = L.random_spinor(dims, seed + 2, device)
    clov = L.random_clover(dims, seed + 3, device)
    lat = dict(U=U, psi=psi, chi=chi, clov=clov, inv=L.invert_clover(clov))
    for s in (1, -1):
        lat[f"B{s}"] = L.tmclover_blocks(clov, MU, s, True)   # (A + i s mu g5)^-1
        lat[f"F{s}"] = L.tmclover_blocks(clov, MU, s, False)  # A + i s mu g5
    return lat


@pytest.fixture(scope="module")
def lat8(qamd):
    return make_lattice(qamd, (8, 8, 8, 8), 1000)


def epilogue_field(kind, achi, isign, lat):
    """The clover-like field the kernel takes for this family."""
    if kind == 1:
        return lat["clov"] if achi else lat["inv"]
    if kind == 3:
        return lat[f"F{isign}"] if achi else lat[f"B{isign}"]
    return None


def run_gpu(qamd, geom, kind, achi, prec, recon, isign, cb, lat):
    dev = DEV
    g = geom.pack_gauge(lat["U"].to(dev), prec, recon, cb)
    psi = geom.pack_spinor(lat["psi"].to(dev), prec, 1 - cb)
    chi = geom.pack_spinor(lat["chi"].to(dev), prec, cb)
    res = geom.new_spinor(prec, cb=cb)
    ep = epilogue_field(kind, achi, isign, lat)
    if kind == 0:
        op = qamd.Dslash(geom)
        if achi:
            op.dslashAChiMinusBDPsi(res, psi, chi, g, ALPHA, BETA, isign, cb)
        else:
            op.dslash(res, psi, g, isign, cb)
    elif kind == 2:
        op = qamd.TMDslash(geom, ALPHA - 4.0, MU)
        if achi:
            op.dslashAChiMinusBDPsi(res, psi, chi, g, ALPHA, BETA, isign, cb)
        else:
            op.dslash(res, psi, g, isign, cb)
    else:
        full = kind == 3
        cl = (geom.pack_full_clover if full else geom.pack_clover)(ep.to(dev), prec, cb)
        op = (qamd.TMClovDslash if full else qamd.ClovDslash)(geom)
        if achi:
            op.dslashAChiMinusBDPsi(res, psi, chi, g, cl, BETA, isign, cb, RHO)
        else:
            op.dslash(res, psi, g, cl, isign, cb)
    torch.cuda.synchronize()
    return geom.unpack_spinor(res).cpu().numpy()


def oracle_args(kind, achi, isign, lat, dims):
    coeff = BETA if achi else 1.0
    kw = dict(kind=kind, achi=achi, coeff_s=coeff, coeff_t_f=coeff, coeff_t_b=coeff,
              t_boundary=-1.0, alpha=ALPHA, rho=RHO)
    if achi:
        kw["chi"] = lat["chi"].numpy()
    ep = epilogue_field(kind, achi, isign, lat)
    if ep is not None:
        kw["clov"] = ep.numpy()
    if kind == 2:
        kw["mu"] = MU / ALPHA
        kw["muinv"] = ALPHA / (ALPHA ** 2 + MU ** 2)
    return kw


def run_oracle(orc, dims, kind, achi, isign, cb, lat):
    kw = oracle_args(kind, achi, isign, lat, dims)
    return orc.dslash(dims, lat["psi"].numpy(), lat["U"].numpy(), isign, cb, **kw)


@pytest.mark.parametrize("prec", [0, 1], ids=["f64", "f32"])
@pytest.mark.parametrize("achi", [False, True], ids=["dslash", "achimbdpsi"])
@pytest.mark.parametrize("kind", [0, 1, 2, 3], ids=["wilson", "clover", "tm", "tmclover"])
def test_parity_8x4(qamd, orc, lat8, kind, achi, prec):
    dims = (8, 8, 8, 8)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    for recon in (12, 18):
        for isign in (1, -1):
            for cb in (0, 1):
                print(f"case kind={kind} achi={achi} prec={prec} recon={recon} isign={isign} cb={cb}",
                      flush=True)
                out = run_gpu(qamd, geom, kind, achi, prec, recon, isign, cb, lat8)
                ref = run_oracle(orc, dims, kind, achi, isign, cb, lat8)
                m = cb_mask(dims, cb)
                err = np.abs(out[m] - ref[m]).max() / np.abs(ref[m]).max()
                assert err < TOL[prec], (recon, isign, cb, err)
                assert np.all(out[~m] == 0)


@pytest.mark.parametrize("prec", [2, 3], ids=["f16", "bf16"])
@pytest.mark.parametrize("kind", [0, 1, 2, 3], ids=["wilson", "clover", "tm", "tmclover"])
def test_parity_low_precision_storage(qamd, orc, lat8, kind, prec):
    dims = (8, 8, 8, 8)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    for achi in (False, True):
        for isign in (1, -1):
            cb = 1
            out = run_gpu(qamd, geom, kind, achi, prec, 12, isign, cb, lat8)
            ref = run_oracle(orc, dims, kind, achi, isign, cb, lat8)
            m = cb_mask(dims, cb)
            err = np.abs(out[m] - ref[m]).max() / np.abs(ref[m]).max()
            assert err < TOL[prec], (achi, isign, err)


def test_tbc_anisotropy_periodic(qamd, orc, lat8):
    """Twisted BC phases in every dimension, anisotropic coefficients and a
    periodic time boundary (FP64)."""
    dims = (8, 8, 8, 8)
    geom = qamd.Geometry(dims, t_boundary=1.0)
    phases = [np.exp(1j * t) for t in (0.3, -0.7, 1.1, 2.0)]
    geom.set_tbc([1, 1, 0, 1], phases)
    U = lat8["U"].to(DEV)
    for isign in (1, -1):
        cb = 0
        g = geom.pack_gauge(U, 0, 18, cb)
        psi = geom.pack_spinor(lat8["psi"].to(DEV), 0, 1)
        res = geom.new_spinor(0, cb=cb)
        op = qamd.Dslash(geom, aniso_coeff_s=0.8, aniso_coeff_t=1.3)
        op.dslash(res, psi, g, isign, cb)
        out = geom.unpack_spinor(res).cpu().numpy()
        ref = orc.dslash(dims, lat8["psi"].numpy(), lat8["U"].numpy(), isign, cb, coeff_s=0.8,
                         coeff_t_f=1.3, coeff_t_b=1.3, t_boundary=1.0, tbc=[1, 1, 0, 1],
                         tbc_phases=phases)
        m = cb_mask(dims, cb)
        assert np.abs(out[m] - ref[m]).max() / np.abs(ref[m]).max() < 1e-13


@pytest.mark.parametrize("prec", [0, 1, 2, 3], ids=["f64", "f32", "f16", "bf16"])
@pytest.mark.parametrize("dims", [(32, 16, 4, 4), (48, 16, 4, 4), (24, 32, 4, 4), (24, 24, 4, 4),
                                  (40, 8, 4, 4)],
                         ids=["rows16", "rows24", "rows12", "planes288", "planes160"])
def test_lds_rows_with_phases(qamd, orc, dims, prec):
    """Lattices whose work-groups read their x and y neighbours from rows
    staged in LDS.  32x16x4x4: 8 whole x rows per work-group, two work-groups
    per plane, the y wrap crossing between them; 48x16x4x4 and 24x32x4x4:
    work-groups that start and end inside a row (5.3 and 10.7 rows);
    24x24x4x4 and 40x8x4x4: planes of 288 and 160 sites, where the
    work-groups inside one plane use LDS and those straddling two planes
    read globally, in the same launch.  Twisted phases on x, y and t,
    antiperiodic time, every family, both isigns and checkerboards."""
    lat = make_lattice(qamd, dims, 1300)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    tbc = [1, 1, 0, 1]
    phases = [np.exp(1j * t) for t in (0.4, -0.9, 0.6, 1.7)]
    geom.set_tbc(tbc, phases)
    for kind in range(4):
        for achi in (False, True):
            for isign, cb in ((1, 0), (-1, 1)):
                for recon in ((12, 18) if prec < 2 else (12,)):
                    out = run_gpu(qamd, geom, kind, achi, prec, recon, isign, cb, lat)
                    kw = oracle_args(kind, achi, isign, lat, dims)
                    kw.update(tbc=tbc, tbc_phases=phases)
                    ref = orc.dslash(dims, lat["psi"].numpy(), lat["U"].numpy(), isign, cb, **kw)
                    m = cb_mask(dims, cb)
                    err = np.abs(out[m] - ref[m]).max() / np.abs(ref[m]).max()
                    assert err < TOL[prec], (kind, achi, isign, cb, recon, err)


@pytest.mark.parametrize("dims", [(32, 16, 16, 16), (48, 16, 8, 8)], ids=["rows16", "rows24"])
def test_traversal_orders_bit_identical(qamd, dims):
    """Every work-group traversal (a t-slab per XCD; XCDs sweeping t
    together over z-ranges; z-major t-slabs; t-slabs in z-halves) gives the
    same bits: each site's arithmetic does not depend on the schedule.  Both
    lattices take the LDS path; 48x16x8x8 has work-groups that start inside
    an x row."""
    lat = make_lattice(qamd, dims, 1400)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    geom.set_tbc([1, 1, 1, 1], [np.exp(1j * t) for t in (0.2, -0.5, 0.9, 1.3)])
    for prec in (0, 1, 2):
        for kind in range(4):
            for achi in (False, True):
                isign, cb = (1, 0) if achi else (-1, 1)
                outs = []
                for order in range(4):
                    geom.set_traversal(order)
                    assert geom.traversal == order
                    outs.append(run_gpu(qamd, geom, kind, achi, prec, 12, isign, cb, lat))
                for order in range(1, 4):
                    assert np.array_equal(outs[order], outs[0]), (prec, kind, achi, order)
    geom.set_traversal(-1)


@pytest.mark.parametrize("dims", [(4, 2, 4, 2), (2, 2, 2, 2), (2, 4, 2, 6)])
def test_small_lattice_4x4(qamd, orc, dims):
    """Smallest extents, FP64 clover: 2 in y and t; Lx = 2 (one site per
    checkerboard row, both x neighbours the same site); the 2^4 lattice."""
    lat = make_lattice(qamd, dims, 77)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    for isign in (1, -1):
        for cb in (0, 1):
            out = run_gpu(qamd, geom, 1, True, 0, 12, isign, cb, lat)
            ref = run_oracle(orc, dims, 1, True, isign, cb, lat)
            m = cb_mask(dims, cb)
            assert np.abs(out[m] - ref[m]).max() / np.abs(ref[m]).max() < 1e-13, (dims, isign, cb)


def test_bad_arguments_raise(qamd):
    geom = qamd.Geometry((8, 8, 8, 8))
    with pytest.raises(qamd.QPhiXAMDError):
        qamd.Geometry((8, 8, 8, 7))
    a = geom.new_spinor(1, cb=0)
    b = geom.new_spinor(1, cb=1)
    g = geom.new_field(qamd.FIELD_GAUGE, 1, 12, cb=0)
    with pytest.raises(qamd.QPhiXAMDError):
        qamd.Dslash(geom).dslash(a, b, g, 2, 0)
    with pytest.raises(qamd.QPhiXAMDError):
        qamd.Dslash(geom).dslash(a, a, g, 1, 0)   # out aliases in


# ---------------------------------------------------------------------------
# the benchmark lattice, 32^3 x 64
@pytest.fixture(scope="module")
def lat32(qamd):
    return make_lattice(qamd, (32, 32, 32, 64), 1100, device=DEV)


def test_parity_32x64_sampled_sites(qamd, orc, lat32):
    dims = (32, 32, 32, 64)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    rng = np.random.default_rng(3)
    cpu = {k: v.cpu() for k, v in lat32.items()}
    par = parity(dims)
    for kind, achi in ((1, False), (1, True), (3, False), (2, True), (0, False)):
        for isign in (1, -1):
            cb = 0 if isign > 0 else 1
            print(f"case 32^3x64 kind={kind} achi={achi} isign={isign}", flush=True)
            out = run_gpu(qamd, geom, kind, achi, 1, 12, isign, cb, lat32)
            sites = rng.choice(np.nonzero(par == cb)[0], 4096, replace=False)
            kw = oracle_args(kind, achi, isign, cpu, dims)
            ref = orc.dslash_sites(dims, sites, cpu["psi"].numpy(), cpu["U"].numpy(), isign, **kw)
            err = np.abs(out[sites] - ref).max() / np.abs(ref).max()
            assert err < 1e-6, (kind, achi, isign, err)


def test_parity_48_slices_sampled_sites(qamd, orc):
    """48x48x48x16: t-slices of 55,296 sites per checkerboard as at 48^3x96,
    so the automatic traversal has the XCDs sweep t together over z-ranges
    (order 1), with the x/y rows staged in LDS.  Twisted-mass clover dslash
    and achimbdpsi and the clover dslash, FP32, FP16 and BF16 storage,
    against the FP64 oracle at 4096 sampled sites of the output checkerboard."""
    dims = (48, 48, 48, 16)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    assert geom.traversal == 1
    lat = make_lattice(qamd, dims, 1500, device=DEV)
    cpu = {k: v.cpu() for k, v in lat.items()}
    rng = np.random.default_rng(5)
    par = parity(dims)
    for prec, tol in ((1, TOL[1]), (2, TOL[2]), (3, TOL[3])):
        for kind, achi, isign in ((3, False, 1), (3, True, 1), (3, False, -1), (1, False, 1)):
            cb = 0 if not achi else 1
            print(f"case 48^3x16 prec={prec} kind={kind} achi={achi} isign={isign}", flush=True)
            out = run_gpu(qamd, geom, kind, achi, prec, 12, isign, cb, lat)
            sites = rng.choice(np.nonzero(par == cb)[0], 4096, replace=False)
            kw = oracle_args(kind, achi, isign, cpu, dims)
            ref = orc.dslash_sites(dims, sites, cpu["psi"].numpy(), cpu["U"].numpy(), isign, **kw)
            err = np.abs(out[sites] - ref).max() / np.abs(ref).max()
            assert err < tol, (prec, kind, achi, isign, err)


def test_recon12_equals_recon18_32x64(qamd, lat32):
    dims = (32, 32, 32, 64)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    op = qamd.ClovDslash(geom)
    cb = 1
    outs = []
    for recon in (12, 18):
        g = geom.pack_gauge(lat32["U"], 1, recon, cb)
        cl = geom.pack_clover(lat32["inv"], 1, cb)
        psi = geom.pack_spinor(lat32["psi"], 1, 0)
        res = geom.new_spinor(1, cb=cb)
        op.dslash(res, psi, g, cl, 1, cb)
        outs.append(res.data.double())
    d = (outs[0] - outs[1]).abs().max().item() / outs[1].abs().max().item()
    assert d < 2e-6, d


def test_adjointness_32x64(qamd, lat32):
    """<chi, D_+ psi> = <D_- chi, psi> with FP64 storage on the full lattice."""
    dims = (32, 32, 32, 64)
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    op = qamd.Dslash(geom)
    g0 = geom.pack_gauge(lat32["U"], 0, 12, 0)
    g1 = geom.pack_gauge(lat32["U"], 0, 12, 1)
    psi1 = geom.pack_spinor(lat32["psi"], 0, 1)
    chi0 = geom.pack_spinor(lat32["chi"], 0, 0)
    a0 = geom.new_spinor(0, cb=0)
    b1 = geom.new_spinor(0, cb=1)
    op.dslash(a0, psi1, g0, 1, 0)
    op.dslash(b1, chi0, g1, -1, 1)
    za = torch.view_as_complex(geom.unpack_spinor(a0))
    zb = torch.view_as_complex(geom.unpack_spinor(b1))
    zc = torch.view_as_complex(geom.unpack_spinor(chi0))
    zp = torch.view_as_complex(geom.unpack_spinor(psi1))
    lhs = torch.vdot(zc.reshape(-1), za.reshape(-1)).item()
    rhs = torch.vdot(zb.reshape(-1), zp.reshape(-1)).item()
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


# ---------------------------------------------------------------------------
# directly against the reference's own kernels (committed fixtures)
def _fixture(name):
    import os
    return dict(np.load(os.path.join(os.path.dirname(__file__), "golden", name), allow_pickle=False))


def _fixture_lat(qamd, f, isign, kind, achi):
    L = qamd.lattice
    d = {k: torch.from_numpy(np.asarray(f[k], dtype=np.float64)) for k in ("U", "psi", "chi", "clov")}
    lat = dict(U=d["U"], psi=d["psi"], chi=d["chi"], clov=d["clov"])
    lat["inv"] = (torch.from_numpy(np.asarray(f["inv"], dtype=np.float64)) if "inv" in f
                  else L.invert_clover(d["clov"]))
    mu = float(f["tm_mu"])
    lat[f"B{isign}"] = L.tmclover_blocks(d["clov"], mu, isign, True)
    lat[f"F{isign}"] = L.tmclover_blocks(d["clov"], mu, isign, False)
    return lat


def _run_fixture_case(qamd, f, geom, prec, case):
    kind, achi, isign, cb, recon = (int(c) for c in case)
    p = {str(k): float(v) for k, v in zip(f["param_names"], f["params"])}
    lat = _fixture_lat(qamd, f, isign, kind, achi)
    g = geom.pack_gauge(lat["U"].to(DEV), prec, recon, cb)
    psi = geom.pack_spinor(lat["psi"].to(DEV), prec, 1 - cb)
    chi = geom.pack_spinor(lat["chi"].to(DEV), prec, cb)
    res = geom.new_spinor(prec, cb=cb)
    lib = qamd.lib()
    ctx = geom.ctx
    st = None
    cs, ctf, ctb = p["coeff_s"], p["coeff_t_f"], p["coeff_t_b"]
    ep = epilogue_field(kind, achi, isign, lat)
    cl = None
    if kind == 1:
        cl = geom.pack_clover(ep.to(DEV), prec, cb)
    elif kind == 3:
        cl = geom.pack_full_clover(ep.to(DEV), prec, cb)
    import ctypes
    P = ctypes.c_void_p
    if kind == 0 and not achi:
        rc = lib.qphix_amd_dslash(ctx, prec, recon, isign, cb, res.ptr(), psi.ptr(), g.ptr(), cs, ctf, ctb, st)
    elif kind == 0:
        rc = lib.qphix_amd_dslash_achimbdpsi(ctx, prec, recon, isign, cb, res.ptr(), psi.ptr(), chi.ptr(),
                                             g.ptr(), p["alpha"], cs, ctf, ctb, st)
    elif kind == 1 and not achi:
        rc = lib.qphix_amd_clov_dslash(ctx, prec, recon, isign, cb, res.ptr(), psi.ptr(), g.ptr(), cl.ptr(),
                                       cs, ctf, ctb, st)
    elif kind == 1:
        rc = lib.qphix_amd_clov_dslash_achimbdpsi(ctx, prec, recon, isign, cb, res.ptr(), psi.ptr(),
                                                  chi.ptr(), g.ptr(), cl.ptr(), cs, ctf, ctb, p["rho"], st)
    elif kind == 2 and not achi:
        rc = lib.qphix_amd_tm_dslash(ctx, prec, recon, isign, cb, res.ptr(), psi.ptr(), g.ptr(), cs, ctf,
                                     ctb, p["mu"], p["muinv"], st)
    elif kind == 2:
        rc = lib.qphix_amd_tm_dslash_achimbdpsi(ctx, prec, recon, isign, cb, res.ptr(), psi.ptr(),
                                                chi.ptr(), g.ptr(), p["alpha"], cs, ctf, ctb, p["mu"], st)
    elif not achi:
        rc = lib.qphix_amd_tm_clov_dslash(ctx, prec, recon, isign, cb, res.ptr(), psi.ptr(), g.ptr(),
                                          cl.ptr(), cs, ctf, ctb, st)
    else:
        rc = lib.qphix_amd_tm_clov_dslash_achimbdpsi(ctx, prec, recon, isign, cb, res.ptr(), psi.ptr(),
                                                     chi.ptr(), g.ptr(), cl.ptr(), cs, ctf, ctb, p["rho"],
                                                     st)
    assert rc == 0
    torch.cuda.synchronize()
    return geom.unpack_spinor(res).cpu().numpy()


def test_gpu_f64_matches_reference_scalar_fixture(qamd):
    """Every operator family through the C ABI vs the reference's scalar
    FP64 kernels, at the fixture's sampled sites (1e-13 relative)."""
    f = _fixture("ref_scalar_f64_4x4.npz")
    dims = tuple(int(d) for d in f["dims"])
    p = {str(k): float(v) for k, v in zip(f["param_names"], f["params"])}
    geom = qamd.Geometry(dims, t_boundary=p["t_boundary"])
    for case, sites, ref_out in zip(f["cases"], f["sites"], f["outs"]):
        out = _run_fixture_case(qamd, f, geom, 0, case)
        err = np.abs(out[sites] - ref_out).max() / np.abs(ref_out).max()
        assert err < 1e-13, (case, err)


def test_gpu_f32_matches_reference_avx512_fixture(qamd):
    """FP32 clover family vs the reference's AVX-512 FP32 kernels: the same
    bits (the kernels follow the reference's FP32 rounding sequence)."""
    f = _fixture("ref_avx512_f32_32x2.npz")
    dims = tuple(int(d) for d in f["dims"])
    p = {str(k): float(v) for k, v in zip(f["param_names"], f["params"])}
    geom = qamd.Geometry(dims, t_boundary=p["t_boundary"])
    for case, ref_out in zip(f["cases"], f["outs"]):
        cb = int(case[3])
        out = _run_fixture_case(qamd, f, geom, 1, case)
        m = cb_mask(dims, cb)
        err = np.abs(out[m] - ref_out).max() / np.abs(ref_out).max()
        assert err < 1e-6, (case, err)
        assert np.array_equal(out[m], np.asarray(ref_out, dtype=np.float64)), (case, err)


def test_gpu_f16_matches_reference_lowprec_fixture(qamd):
    """FP16 storage, FP32 arithmetic, every family and kernel, 12 and 18
    reals, against the reference's own low-precision AVX-512 kernels
    (golden fixture): the same bits in every element (FP32 arithmetic in the
    reference's rounding sequence, each result rounded to FP32 and then to
    FP16)."""
    from _util import f16_ulps, f16_within
    f = _fixture("ref_avx512_f16_32x2.npz")
    dims = tuple(int(d) for d in f["dims"])
    p = {str(k): float(v) for k, v in zip(f["param_names"], f["params"])}
    geom = qamd.Geometry(dims, t_boundary=p["t_boundary"])
    lex = {k: torch.from_numpy(np.asarray(f[k], dtype=np.float64)).to(DEV)
           for k in ("U", "psi", "chi", "clov", "inv", "fB", "fF")}
    F16 = qamd.F16
    cs, ctf, ctb = p["coeff_s"], p["coeff_t_f"], p["coeff_t_b"]
    lib = qamd.lib()
    for case, ref_out in zip(f["cases"], f["outs"]):
        kind, achi, isign, cb, recon = (int(c) for c in case)
        g = geom.pack_gauge(lex["U"], F16, recon, cb)
        psi = geom.pack_spinor(lex["psi"], F16, 1 - cb)
        chi = geom.pack_spinor(lex["chi"], F16, cb)
        res = geom.new_spinor(F16, cb=cb)
        a = (geom.ctx, F16, recon, isign, cb, res.ptr(), psi.ptr())
        if kind == 0:
            rc = (lib.qphix_amd_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), p["alpha"], cs, ctf, ctb, None)
                  if achi else lib.qphix_amd_dslash(*a, g.ptr(), cs, ctf, ctb, None))
        elif kind == 1:
            cl = geom.pack_clover(lex["clov" if achi else "inv"], F16, cb)
            rc = (lib.qphix_amd_clov_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), cl.ptr(), cs, ctf, ctb,
                                                       p["rho"], None)
                  if achi else lib.qphix_amd_clov_dslash(*a, g.ptr(), cl.ptr(), cs, ctf, ctb, None))
        elif kind == 2:
            rc = (lib.qphix_amd_tm_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), p["alpha"], cs, ctf, ctb,
                                                     p["mu"], None)
                  if achi else lib.qphix_amd_tm_dslash(*a, g.ptr(), cs, ctf, ctb, p["mu"], p["muinv"], None))
        else:
            cl = geom.pack_full_clover(lex["fF" if achi else "fB"], F16, cb)
            rc = (lib.qphix_amd_tm_clov_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), cl.ptr(), cs, ctf, ctb,
                                                          p["rho"], None)
                  if achi else lib.qphix_amd_tm_clov_dslash(*a, g.ptr(), cl.ptr(), cs, ctf, ctb, None))
        assert rc == 0
        out = geom.unpack_spinor(res).cpu().numpy()[cb_mask(dims, cb)]
        ulps = f16_ulps(out.astype(np.float16), ref_out)
        same = float(np.mean(ulps == 0))
        print(f"f16 case {tuple(int(c) for c in case)}: max {ulps.max()} ulp, identical {same:.4f}", flush=True)
        assert f16_within(out, ref_out).all() and same == 1.0, (case, ulps.max(), same)


def test_ndtm_entry_points_are_the_degenerate_kernels(qamd):
    """The reference's ndtm_ bodies are its degenerate twisted-mass bodies
    (codegen/dslash_common.cc:1599-1603,1643-1650): the ndtm_ entry points
    give bit-identical results to the tm_ ones."""
    dims = (8, 4, 4, 4)
    L = qamd.lattice
    geom = qamd.Geometry(dims)
    U = L.random_gauge(dims, 0.3, 901).to(DEV)
    psi_l, chi_l = L.random_spinor(dims, 902).to(DEV), L.random_spinor(dims, 903).to(DEV)
    fcl = L.tmclover_blocks(L.random_clover(dims, 904), 0.1, 1, True).to(DEV)
    lib = qamd.lib()
    for prec in (qamd.F64, qamd.F16):
        for isign, cb in ((1, 0), (-1, 1)):
            g = geom.pack_gauge(U, prec, 12, cb)
            psi, chi = geom.pack_spinor(psi_l, prec, 1 - cb), geom.pack_spinor(chi_l, prec, cb)
            cl = geom.pack_full_clover(fcl, prec, cb)
            outs = []
            for nd in ("tm", "ndtm"):
                a = (geom.ctx, prec, 12, isign, cb)
                r = [geom.new_spinor(prec, cb=cb) for _ in range(4)]
                assert getattr(lib, f"qphix_amd_{nd}_dslash")(*a, r[0].ptr(), psi.ptr(), g.ptr(), 0.9, 1.1, 1.2,
                                                             0.03, 0.26, None) == 0
                assert getattr(lib, f"qphix_amd_{nd}_dslash_achimbdpsi")(*a, r[1].ptr(), psi.ptr(), chi.ptr(),
                                                                         g.ptr(), 3.7, 0.9, 1.1, 1.2, 0.03,
                                                                         None) == 0
                assert getattr(lib, f"qphix_amd_{nd}_clov_dslash")(*a, r[2].ptr(), psi.ptr(), g.ptr(), cl.ptr(),
                                                                  0.9, 1.1, 1.2, None) == 0
                assert getattr(lib, f"qphix_amd_{nd}_clov_dslash_achimbdpsi")(*a, r[3].ptr(), psi.ptr(),
                                                                              chi.ptr(), g.ptr(), cl.ptr(),
                                                                              0.9, 1.1, 1.2, 0.05, None) == 0
                outs.append([x.data.cpu() for x in r])
            for a_, b_ in zip(*outs):
                assert torch.equal(a_, b_)
