"""GPU tests of the face operations (qphix_amd_set_comm, qphix_amd_face_pack,
qphix_amd_{,clov_,tm_,tm_clov_}face_finish): a lattice split over several
GPUs, emulated on one.

Each part of the split has its own context (its split dimensions marked with
set_comm), its own fields -- the gauge handed over in the QPhiX layout with
the back links of its boundary sites taken from across the cut, as a host
with a gauge halo provides them -- runs the operator (the interior part)
and packs its faces; the faces are exchanged between the parts on the host
side; the finish calls complete the boundary sites.  The assembled result
must equal the FP64 oracle's operator on the whole lattice (1e-13 in FP64,
1e-6 in FP32, the storage rounding in FP16 and BF16), for every family,
dslash and achimbdpsi, both isigns and checkerboards, antiperiodic time, and
with twisted boundary phases.
"""
import itertools

import numpy as np
import pytest
import torch

import qphix_blocks as qb
from _util import cb_mask

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
CS, CTF, CTB, TB = 0.9, 1.15, 1.05, -1.0
ALPHA, MU_PHYS, RHO = 4.2, 0.13, 0.07
MU, MUINV = MU_PHYS / ALPHA, ALPHA / (ALPHA ** 2 + MU_PHYS ** 2)

SPLITS = {  # name: (global dims, parts per dimension)
    "t2": ((8, 4, 4, 8), (1, 1, 1, 2)),
    "x2": ((16, 4, 4, 4), (2, 1, 1, 1)),
    "yz2x2": ((8, 8, 8, 4), (1, 2, 2, 1)),
    "xyzt16": ((16, 8, 8, 8), (2, 2, 2, 2)),
    # parts whose rows fill whole work-groups: the x and y hops come from
    # the rows staged in LDS
    "x2_rows": ((64, 8, 4, 4), (2, 1, 1, 1)),
    "y2_rows": ((32, 16, 4, 4), (1, 2, 1, 1)),
}


def _lattice(qamd, G, seed):
    L = qamd.lattice
    clov = L.random_clover(G, seed + 3)
    return dict(U=L.random_gauge(G, 0.35, seed).numpy(), psi=L.random_spinor(G, seed + 1).numpy(),
                chi=L.random_spinor(G, seed + 2).numpy(), clov=clov.numpy(),
                inv=L.invert_clover(clov).numpy(),
                Bp=L.tmclover_blocks(clov, MU_PHYS, 1, True).numpy(),
                Bm=L.tmclover_blocks(clov, MU_PHYS, -1, True).numpy(),
                Fp=L.tmclover_blocks(clov, MU_PHYS, 1, False).numpy(),
                Fm=L.tmclover_blocks(clov, MU_PHYS, -1, False).numpy())


def _epilogue(lat, kind, achi, isign):
    s = "p" if isign > 0 else "m"
    if kind == 1:
        return lat["clov"] if achi else lat["inv"]
    if kind == 3:
        return lat[("F" if achi else "B") + s]
    return None


def _global_index(G, Lloc, origin):
    """Global lex index of each local lex site."""
    t, z, y, x = np.meshgrid(*(np.arange(n) for n in Lloc[::-1]), indexing="ij")
    gx, gy, gz, gt = (c.ravel() + o for c, o in zip((x, y, z, t), origin))
    return ((gt * G[2] + gz) * G[1] + gy) * G[0] + gx


def _run(lib, geom, kind, achi, prec, recon, isign, cb, res, psi, chi, g, cl):
    a = (geom.ctx, prec, recon, isign, cb, res.ptr(), psi.ptr())
    if kind == 0:
        rc = (lib.qphix_amd_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), ALPHA, CS, CTF, CTB, None) if achi
              else lib.qphix_amd_dslash(*a, g.ptr(), CS, CTF, CTB, None))
    elif kind == 1:
        rc = (lib.qphix_amd_clov_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), cl.ptr(), CS, CTF, CTB, RHO, None)
              if achi else lib.qphix_amd_clov_dslash(*a, g.ptr(), cl.ptr(), CS, CTF, CTB, None))
    elif kind == 2:
        rc = (lib.qphix_amd_tm_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), ALPHA, CS, CTF, CTB, MU, None)
              if achi else lib.qphix_amd_tm_dslash(*a, g.ptr(), CS, CTF, CTB, MU, MUINV, None))
    else:
        rc = (lib.qphix_amd_tm_clov_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), cl.ptr(), CS, CTF, CTB, RHO,
                                                      None)
              if achi else lib.qphix_amd_tm_clov_dslash(*a, g.ptr(), cl.ptr(), CS, CTF, CTB, None))
    assert rc == 0


def _run_q(lib, geom, kind, achi, recon, isign, cb, res, psi, chi, g, cl):
    """_run on QPhiX arrays in place (qphix_amd_qphix_*), the same scalars."""
    import ctypes
    a = (geom.ctx, ctypes.byref(res.layout), recon, isign, cb, res.ptr(), psi.ptr())
    if kind == 0:
        rc = (lib.qphix_amd_qphix_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), None, ALPHA, CS, CTF, CTB, None)
              if achi else lib.qphix_amd_qphix_dslash(*a, g.ptr(), None, CS, CTF, CTB, None))
    elif kind == 1:
        rc = (lib.qphix_amd_qphix_clov_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), cl.ptr(), None, CS, CTF, CTB, RHO,
                                                         None)
              if achi else lib.qphix_amd_qphix_clov_dslash(*a, g.ptr(), cl.ptr(), None, CS, CTF, CTB, None))
    elif kind == 2:
        rc = (lib.qphix_amd_qphix_tm_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), None, ALPHA, CS, CTF, CTB, MU,
                                                       None)
              if achi else lib.qphix_amd_qphix_tm_dslash(*a, g.ptr(), None, CS, CTF, CTB, MU, MUINV, None))
    else:
        rc = (lib.qphix_amd_qphix_tm_clov_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), cl.ptr(), None, CS, CTF, CTB,
                                                            RHO, None)
              if achi else lib.qphix_amd_qphix_tm_clov_dslash(*a, g.ptr(), cl.ptr(), None, CS, CTF, CTB, None))
    assert rc == 0


def split_apply(qamd, lat, G, grid, kind, achi, isign, cb, prec, recon, tbc=None, qv=None):
    """The operator on the whole lattice G assembled from its parts.  qv =
    (veclen, soalen): the parts keep psi, chi and the result as QPhiX
    FourSpinorBlock arrays, which the operators, face_pack and face_finish
    read and write in place."""
    lib = qamd.lib()
    Lloc = tuple(n // p for n, p in zip(G, grid))
    comm = [int(p > 1) for p in grid]
    lay = qb.Layout(0 if prec == qamd.F64 else 1, 1, 1, compress12=recon == 12)
    parts = {}
    ep = _epilogue(lat, kind, achi, isign)
    for gc in itertools.product(*(range(p) for p in grid)):
        origin = tuple(c * n for c, n in zip(gc, Lloc))
        gi = _global_index(G, Lloc, origin)
        geom = qamd.Geometry(Lloc, t_boundary=TB)
        geom.set_comm(comm)
        if tbc is not None:
            geom.set_tbc(*tbc)
        garr = qb.pack_gauge(lat["U"], lay, Lloc, cb, origin=origin, gdims=G)
        g = geom.gauge_from_qphix(torch.from_numpy(garr).to(DEV), lay.ctype(qamd), prec, recon, cb)
        psi = geom.pack_spinor(torch.from_numpy(lat["psi"][gi]).to(DEV), prec, 1 - cb)
        chi = geom.pack_spinor(torch.from_numpy(lat["chi"][gi]).to(DEV), prec, cb)
        cl = None
        if kind == 1:
            cl = geom.pack_clover(torch.from_numpy(ep[gi]).to(DEV), prec, cb)
        elif kind == 3:
            cl = geom.pack_full_clover(torch.from_numpy(ep[gi]).to(DEV), prec, cb)
        res = geom.new_spinor(prec, cb=cb)
        if qv is None:
            _run(lib, geom, kind, achi, prec, recon, isign, cb, res, psi, chi, g, cl)
        else:
            ql = qamd.qphix_layout(prec, qv[0], qv[1], recon == 12)
            psi = qamd.QPhiXSpinor(geom.spinor_to_qphix(psi, ql), ql)
            chi = qamd.QPhiXSpinor(geom.spinor_to_qphix(chi, ql), ql)
            res = qamd.QPhiXSpinor(torch.zeros_like(psi.data), ql)
            _run_q(lib, geom, kind, achi, recon, isign, cb, res, psi, chi, g, cl)
        faces = {}
        for mu in range(4):
            if comm[mu]:
                for fb in (0, 1):
                    faces[mu, fb] = qamd.face_pack(geom, geom.new_face(prec, mu), psi, isign, cb, mu, fb)
        parts[gc] = dict(geom=geom, origin=origin, gi=gi, g=g, cl=cl, res=res, faces=faces)
    for gc, p in parts.items():
        for mu in range(4):
            if not comm[mu]:
                continue
            for fb in (0, 1):
                # fb 0: from the backward neighbour, which sent its slice L-1 forwards
                step = -1 if fb == 0 else 1
                nb = list(gc)
                nb[mu] = (nb[mu] + step) % grid[mu]
                buf = parts[tuple(nb)]["faces"][mu, 1 - fb]
                if mu < 3:
                    beta = CS
                else:
                    beta = CTB if fb == 0 else CTF
                    at_edge = p["origin"][3] == 0 if fb == 0 else p["origin"][3] + Lloc[3] == G[3]
                    if at_edge:
                        beta *= TB
                if achi:
                    qamd.face_finish(p["geom"], 0, p["res"], buf, p["g"], isign, cb, mu, fb, beta)
                else:
                    sgn = 1.0 if kind in (1, 3) else -1.0
                    qamd.face_finish(p["geom"], kind, p["res"], buf, p["g"], isign, cb, mu, fb, sgn * beta,
                                     clov=p["cl"], mu=MU, muinv=MUINV)
    out = np.zeros_like(lat["psi"])
    for p in parts.values():
        res = p["res"]
        if qv is not None:
            res = p["geom"].spinor_from_qphix(res.data, res.layout, prec, cb)
        loc = p["geom"].unpack_spinor(res).cpu().numpy()
        m = cb_mask(p["geom"].dims, cb)
        out[p["gi"][m]] = loc[m]
    return out


def oracle_apply(orc, lat, G, kind, achi, isign, cb, tbc=None):
    kw = dict(coeff_s=CS, coeff_t_f=CTF, coeff_t_b=CTB, t_boundary=TB, alpha=ALPHA, mu=MU, muinv=MUINV,
              rho=RHO, kind=kind, achi=bool(achi), chi=lat["chi"], clov=_epilogue(lat, kind, achi, isign))
    if tbc is not None:
        kw.update(tbc=tbc[0], tbc_phases=tbc[1])
    return orc.dslash(G, lat["psi"], lat["U"], isign, cb, **kw)


@pytest.mark.parametrize("split", sorted(SPLITS))
def test_split_lattice_equals_whole_lattice(qamd, orc, split):
    G, grid = SPLITS[split]
    lat = _lattice(qamd, G, 700)
    for kind, achi in itertools.product(range(4), (0, 1)):
        for isign, cb in ((1, 0), (-1, 1)):
            ref = oracle_apply(orc, lat, G, kind, achi, isign, cb)
            m = cb_mask(G, cb)
            for prec, recon, tol in ((qamd.F64, 18, 1e-13), (qamd.F32, 12, 1e-6)):
                got = split_apply(qamd, lat, G, grid, kind, achi, isign, cb, prec, recon)
                err = np.abs(got[m] - ref[m]).max() / np.abs(ref[m]).max()
                assert err < tol, (split, kind, achi, isign, cb, prec, err)


@pytest.mark.parametrize("split", ["t2", "xyzt16", "x2_rows"])
def test_split_low_precision_storage(qamd, orc, split):
    """FP16 and BF16 storage (FP32 arithmetic; the face buffers in the same
    storage): the assembled result equals the whole-lattice oracle within
    the storage rounding, 5e-3 (FP16) and 4e-2 (BF16) relative as for the
    unsplit kernels."""
    G, grid = SPLITS[split]
    lat = _lattice(qamd, G, 730)
    for kind, achi in ((0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (3, 0), (3, 1)):
        isign, cb = (1, 0) if achi == 0 else (-1, 1)
        ref = oracle_apply(orc, lat, G, kind, achi, isign, cb)
        m = cb_mask(G, cb)
        for prec, tol in ((qamd.F16, 5e-3), (qamd.BF16, 4e-2)):
            got = split_apply(qamd, lat, G, grid, kind, achi, isign, cb, prec, 12)
            err = np.abs(got[m] - ref[m]).max() / np.abs(ref[m]).max()
            assert err < tol, (split, kind, achi, prec, err)


@pytest.mark.parametrize("split", ["xyzt16", "x2_rows", "y2_rows"])
def test_split_with_twisted_boundary_phases(qamd, orc, split):
    G, grid = SPLITS[split]
    lat = _lattice(qamd, G, 710)
    tbc = ([1, 0, 1, 1], [np.exp(0.3j), 1.0, np.exp(-0.7j), np.exp(1.1j)])
    for kind, achi in ((0, 0), (1, 0), (2, 0), (3, 1)):
        for isign, cb in ((1, 1), (-1, 0)):
            ref = oracle_apply(orc, lat, G, kind, achi, isign, cb, tbc)
            got = split_apply(qamd, lat, G, grid, kind, achi, isign, cb, qamd.F64, 12, tbc)
            m = cb_mask(G, cb)
            err = np.abs(got[m] - ref[m]).max() / np.abs(ref[m]).max()
            assert err < 1e-13, (kind, achi, isign, cb, err)


def test_face_pack_is_the_projected_boundary(qamd, orc):
    """A face buffer holds (1 -+ isign gamma_mu) psi on the boundary slice,
    [spin][colour][re|im] per face site, in the documented face order."""
    from _util import GAMMA
    G = (8, 4, 6, 4)
    L = qamd.lattice
    psi = L.random_spinor(G, 720).numpy()
    geom = qamd.Geometry(G)
    Lx, Ly, Lz, Lt = G
    for isign, cb in ((1, 0), (-1, 1)):
        f = geom.pack_spinor(torch.from_numpy(psi).to(DEV), qamd.F64, 1 - cb)
        for mu in range(4):
            for fb in (0, 1):
                buf = qamd.face_pack(geom, geom.new_face(qamd.F64, mu), f, isign, cb, mu, fb)
                n = geom.face_sites(mu)
                W = 2
                h = buf.cpu().numpy().reshape(12 // W, n, W).transpose(1, 0, 2).reshape(n, 2, 3, 2)
                sgn = -isign if fb == 0 else isign
                P = np.eye(4) + sgn * GAMMA[mu]
                edge = 0 if fb == 0 else G[mu] - 1
                for fi in range(n):
                    if mu == 0:
                        yh, r = fi % (Ly // 2), fi // (Ly // 2)
                        z, t, x = r % Lz, r // Lz, edge
                        y = 2 * yh + ((1 - cb + x + z + t) & 1)
                    else:
                        xh, r = fi % (Lx // 2), fi // (Lx // 2)
                        if mu == 1:
                            z, t, y = r % Lz, r // Lz, edge
                        elif mu == 2:
                            y, t, z = r % Ly, r // Ly, edge
                        else:
                            y, z, t = r % Ly, r // Ly, edge
                        x = 2 * xh + ((1 - cb + y + z + t) & 1)
                    s = ((t * Lz + z) * Ly + y) * Lx + x
                    sp = psi[s, :, :, 0] + 1j * psi[s, :, :, 1]  # [spin][colour]
                    full = P @ sp
                    hc = h[fi, :, :, 0] + 1j * h[fi, :, :, 1]
                    # the projected spinor is determined by its upper two spins
                    np.testing.assert_allclose(hc, full[:2], atol=1e-14)


@pytest.mark.parametrize("split,ft,qv", [("t2", 1, (8, 4)), ("xyzt16", 0, (4, 2)), ("x2_rows", 1, (16, 16)),
                                         ("y2_rows", 2, (16, 8)), ("t2", 2, (4, 4))],
                         ids=["t2_f32_s4", "xyzt16_f64_s2", "x2rows_f32_s16", "y2rows_f16_s8", "t2_f16_s4"])
def test_split_lattice_in_place_on_qphix_arrays(qamd, split, ft, qv):
    """A split lattice whose parts keep their spinors in QPhiX
    FourSpinorBlock arrays: the in-place operators give the interior, the
    faces are packed from and finished into those arrays in place
    (qphix_amd_qphix_face_pack, qphix_amd_qphix_*face_finish).  The
    assembled result has the same bits as the split lattice on native
    spinors, for every family, dslash and achimbdpsi, both isigns."""
    G, grid = SPLITS[split]
    lat = _lattice(qamd, G, 710)
    prec = (qamd.F64, qamd.F32, qamd.F16)[ft]
    recon = 12
    for kind, achi in itertools.product(range(4), (0, 1)):
        for isign, cb in ((1, 0), (-1, 1)):
            print(f"case in-place split={split} ft={ft} qv={qv} kind={kind} achi={achi} isign={isign}", flush=True)
            want = split_apply(qamd, lat, G, grid, kind, achi, isign, cb, prec, recon)
            got = split_apply(qamd, lat, G, grid, kind, achi, isign, cb, prec, recon, qv=qv)
            m = cb_mask(G, cb)
            assert np.array_equal(got[m], want[m]), (split, kind, achi, isign, cb)


@pytest.mark.parametrize("prec", [0, 1, 2, 3], ids=["f64", "f32", "f16", "bf16"])
def test_both_faces_in_one_launch(qamd, prec):
    """fb = 2 packs and finishes both faces of a dimension in one launch, the
    buffer holding face 0 then face 1: the same bits as the fb = 0 and fb = 1
    calls, for every dimension and every finish epilogue (and, in FP32, on
    QPhiX arrays in place)."""
    dims = (8, 4, 6, 4)
    print(f"case both faces prec={prec} {dims}", flush=True)
    L = qamd.lattice
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    geom.set_comm([1, 1, 1, 1])
    U = L.random_gauge(dims, 0.3, 900, device=DEV)
    clov = L.random_clover(dims, 901, device=DEV)
    g0 = geom.pack_gauge(U, prec, 12, 0)
    cl = geom.pack_clover(clov, prec, 0)
    fcl = geom.pack_full_clover(L.tmclover_blocks(clov, 0.1, 1, True), prec, 0)
    psi = geom.pack_spinor(L.random_spinor(dims, 902, device=DEV), prec, 1)
    base = geom.pack_spinor(L.random_spinor(dims, 903, device=DEV), prec, 0)
    variants = [(psi, base, lambda t: t)]
    if prec == qamd.F32:
        lay = qamd.qphix_layout(prec, 4, 4, True)
        variants.append((qamd.QPhiXSpinor(geom.spinor_to_qphix(psi, lay), lay),
                         qamd.QPhiXSpinor(geom.spinor_to_qphix(base, lay), lay),
                         lambda t: qamd.QPhiXSpinor(t, lay)))
    for src, out0, wrap in variants:
        for dim in range(4):
            n = geom.face_sites(dim) * 12
            both = geom.new_face(prec, dim, both=True)
            qamd.face_pack(geom, both, src, 1, 0, dim, 2)
            for fb in (0, 1):
                one = geom.new_face(prec, dim)
                qamd.face_pack(geom, one, src, 1, 0, dim, fb)
                assert torch.equal(one, both[fb * n:(fb + 1) * n]), (dim, fb)
            recv = torch.randn(2 * n, device=DEV).to(both.dtype)
            for kind, extra in ((0, {}), (1, dict(clov=cl)), (2, dict(mu=0.1, muinv=0.9)), (3, dict(clov=fcl))):
                ref, got = wrap(out0.data.clone()), wrap(out0.data.clone())
                if isinstance(out0, qamd.Field):
                    ref = qamd.Field(ref, prec, out0.kind, cb=0)
                    got = qamd.Field(got, prec, out0.kind, cb=0)
                for fb in (0, 1):
                    qamd.face_finish(geom, kind, ref, recv[fb * n:(fb + 1) * n], g0, 1, 0, dim, fb, 0.5, **extra)
                qamd.face_finish(geom, kind, got, recv, g0, 1, 0, dim, 2, 0.5, **extra)
                torch.cuda.synchronize()
                assert torch.equal(ref.data, got.data), (dim, kind)
                assert not torch.equal(got.data, out0.data), (dim, kind)
