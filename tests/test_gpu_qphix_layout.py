"""GPU tests of the QPhiX block-layout converters
(qphix_amd_{spinor,gauge,clover,full_clover}_from_qphix, qphix_amd_spinor_to_qphix).

* Round trip: QPhiX spinor blocks -> native (same precision) -> QPhiX
  blocks is bit-exact, and padding blocks of the destination are untouched.
* Operator through the QPhiX layout: QPhiX-layout gauge (compressed and
  not), spinors and clover / full clover blocks converted on the device,
  clover and twisted-mass-clover dslash and achimbdpsi, result converted back
  to QPhiX blocks, against the FP64 oracle on the same (FT-rounded) inputs.
  Layouts cover soalen < veclen (one gauge vector spanning 2-8 rows of y),
  padding, FP64 / FP32 / FP16 storage.
* The reference's AVX-512 FP32 kernels' outputs (golden fixture, VECLEN =
  SOALEN = 16) are reproduced from inputs handed over in the QPhiX layout.
"""
import os

import numpy as np
import pytest
import torch

import qphix_blocks as qb
from _util import cb_mask

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DIMS = (32, 8, 4, 4)
NP2PREC = {0: 0, 1: 1, 2: 2}  # FT -> native precision of the same width
TOL = {0: 1e-13, 1: 2e-6, 2: 8e-3}

LAYOUTS = {
    "f64_v8s8_18": qb.Layout(0, 8, 8, compress12=False),
    "f64_v4s4_12_pad": qb.Layout(0, 4, 4, True, 2, 6),
    "f32_v16s8_12_pad": qb.Layout(1, 16, 8, True, 4, 8),
    "f32_v16s4_18": qb.Layout(1, 16, 4, False),
    "f16_v16s16_12": qb.Layout(2, 16, 16, True),
    "f16_v16s2_18_pad": qb.Layout(2, 16, 2, False, 8, 16),
}


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def _q(a, ft):
    """Round FP64 data to the layout's storage type and back."""
    return np.asarray(a).astype(qb.NP_FT[ft]).astype(np.float64)


@pytest.fixture(scope="module")
def lat(qamd):
    L = qamd.lattice
    clov = L.random_clover(DIMS, 603)
    return dict(U=L.random_gauge(DIMS, 0.3, 601).numpy(), psi=L.random_spinor(DIMS, 602).numpy(),
                chi=L.random_spinor(DIMS, 604).numpy(), clov=clov.numpy(),
                inv=L.invert_clover(clov).numpy(),
                B=L.tmclover_blocks(clov, 0.1, 1, True).numpy(),
                F=L.tmclover_blocks(clov, 0.1, 1, False).numpy())


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_spinor_round_trip_and_padding(qamd, lat, name):
    lay = LAYOUTS[name]
    geom = qamd.Geometry(DIMS)
    cl = lay.ctype(qamd)
    for cb in (0, 1):
        src = qb.pack_spinor(lat["psi"], lay, DIMS, cb, fill=np.nan)
        f = geom.spinor_from_qphix(_dev(src), cl, NP2PREC[lay.ft], cb)
        back = torch.full_like(_dev(src), 7.0)
        geom.spinor_to_qphix(f, cl, out=back)
        back = back.cpu().numpy()
        pad = np.isnan(src)
        assert np.all(back[pad] == 7.0), "padding blocks were written"
        assert np.array_equal(back[~pad], src[~pad])
        # the native field holds the same values as the lexicographic path
        lex = geom.unpack_spinor(f).cpu().numpy()
        m = cb_mask(DIMS, cb)
        np.testing.assert_array_equal(lex[m], _q(lat["psi"], lay.ft)[m])


@pytest.mark.parametrize("name", sorted(LAYOUTS))
def test_operators_through_qphix_layout(qamd, orc, lat, name):
    lay = LAYOUTS[name]
    prec = NP2PREC[lay.ft]
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    cl = lay.ctype(qamd)
    U = _q(lat["U"], lay.ft)
    cs, ct, rho = 0.9, 1.1, 0.05
    for recon in ((12, 18) if lay.compress12 else (18,)):
        for cb, isign in ((0, 1), (1, -1)):
            g = geom.gauge_from_qphix(_dev(qb.pack_gauge(lat["U"], lay, DIMS, cb)), cl, prec, recon, cb)
            psi = geom.spinor_from_qphix(_dev(qb.pack_spinor(lat["psi"], lay, DIMS, 1 - cb)), cl, prec, 1 - cb)
            chi = geom.spinor_from_qphix(_dev(qb.pack_spinor(lat["chi"], lay, DIMS, cb)), cl, prec, cb)
            psi_l, chi_l = _q(lat["psi"], lay.ft), _q(lat["chi"], lay.ft)
            inv = geom.clover_from_qphix(_dev(qb.pack_clover(lat["inv"], lay, DIMS, cb)), cl, prec, cb)
            A = geom.clover_from_qphix(_dev(qb.pack_clover(lat["clov"], lay, DIMS, cb)), cl, prec, cb)
            B = geom.full_clover_from_qphix(_dev(qb.pack_full_clover(lat["B"], lay, DIMS, cb)), cl, prec, cb)
            F = geom.full_clover_from_qphix(_dev(qb.pack_full_clover(lat["F"], lay, DIMS, cb)), cl, prec, cb)
            cd = qamd.ClovDslash(geom, cs, ct)
            runs = []
            res = geom.new_spinor(prec, cb=cb)
            cd.dslash(res, psi, g, inv, isign, cb)
            runs.append((res, dict(kind=orc.CLOVER, clov=_q(lat["inv"], lay.ft))))
            res = geom.new_spinor(prec, cb=cb)
            cd.dslashAChiMinusBDPsi(res, psi, chi, g, A, 1.0, isign, cb, prec_mass_rho=rho)
            runs.append((res, dict(kind=orc.CLOVER, achi=True, chi=chi_l, clov=_q(lat["clov"], lay.ft),
                                   rho=rho)))
            tcd = qamd.TMClovDslash(geom, cs, ct)
            res = geom.new_spinor(prec, cb=cb)
            tcd.dslash(res, psi, g, B, isign, cb)
            runs.append((res, dict(kind=orc.TMCLOVER, clov=_q(lat["B"], lay.ft))))
            res = geom.new_spinor(prec, cb=cb)
            tcd.dslashAChiMinusBDPsi(res, psi, chi, g, F, 1.0, isign, cb, prec_mass_rho=rho)
            runs.append((res, dict(kind=orc.TMCLOVER, achi=True, chi=chi_l, clov=_q(lat["F"], lay.ft),
                                   rho=rho)))
            m = cb_mask(DIMS, cb)
            for res, kw in runs:
                out = qb.unpack_spinor(geom.spinor_to_qphix(res, cl).cpu().numpy(), lay, DIMS, cb)
                ref = orc.dslash(DIMS, psi_l, U, isign, cb, coeff_s=cs, coeff_t_f=ct, coeff_t_b=ct, **kw)
                err = np.abs(out[m] - ref[m]).max() / np.abs(ref[m]).max()
                assert err < TOL[lay.ft], (name, recon, cb, isign, kw["kind"], kw.get("achi"), err)


def test_reference_avx512_fixture_through_qphix_layout(qamd):
    """Inputs in the reference kernels' own block layout (VECLEN = SOALEN = 16,
    FP32, the blocks its AVX-512 clover kernels read) -> device converters ->
    FP32 clover operators -> QPhiX blocks; the reference outputs to 1e-6."""
    f = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_avx512_f32_32x2.npz"),
                     allow_pickle=False))
    dims = tuple(int(d) for d in f["dims"])
    p = {str(k): float(v) for k, v in zip(f["param_names"], f["params"])}
    geom = qamd.Geometry(dims, t_boundary=p["t_boundary"])
    f64 = {k: np.asarray(f[k], dtype=np.float64) for k in ("U", "psi", "chi", "clov", "inv")}
    lib = qamd.lib()
    for case, ref_out in zip(f["cases"], f["outs"]):
        kind, achi, isign, cb, recon = (int(c) for c in case)
        assert kind == 1
        lay = qb.Layout(1, 16, 16, compress12=recon == 12)
        cl = lay.ctype(qamd)
        g = geom.gauge_from_qphix(_dev(qb.pack_gauge(f64["U"], lay, dims, cb)), cl, 1, recon, cb)
        psi = geom.spinor_from_qphix(_dev(qb.pack_spinor(f64["psi"], lay, dims, 1 - cb)), cl, 1, 1 - cb)
        res = geom.new_spinor(1, cb=cb)
        if achi:
            chi = geom.spinor_from_qphix(_dev(qb.pack_spinor(f64["chi"], lay, dims, cb)), cl, 1, cb)
            A = geom.clover_from_qphix(_dev(qb.pack_clover(f64["clov"], lay, dims, cb)), cl, 1, cb)
            rc = lib.qphix_amd_clov_dslash_achimbdpsi(geom.ctx, 1, recon, isign, cb, res.ptr(), psi.ptr(),
                                                      chi.ptr(), g.ptr(), A.ptr(), p["coeff_s"],
                                                      p["coeff_t_f"], p["coeff_t_b"], p["rho"], None)
        else:
            inv = geom.clover_from_qphix(_dev(qb.pack_clover(f64["inv"], lay, dims, cb)), cl, 1, cb)
            rc = lib.qphix_amd_clov_dslash(geom.ctx, 1, recon, isign, cb, res.ptr(), psi.ptr(), g.ptr(),
                                           inv.ptr(), p["coeff_s"], p["coeff_t_f"], p["coeff_t_b"], None)
        assert rc == 0
        out = qb.unpack_spinor(geom.spinor_to_qphix(res, cl).cpu().numpy(), lay, dims, cb)
        m = cb_mask(dims, cb)
        err = np.abs(out[m] - ref_out).max() / np.abs(ref_out).max()
        assert err < 1e-6, (case, err)
