"""GPU tests of the operators on QPhiX FourSpinorBlock arrays in place
(qphix_amd_qphix_*), called through the C ABI the way a QPhiX host would:
the spinor arrays are the reference driver's own blocks (oracle/ref_driver.cpp
packs them for the reference's kernels) handed over as they are, the links
and clover blocks are converted once from the driver's SU3MatrixBlock /
CloverBlock / FullCloverBlock arrays, and the result is written into a
FourSpinorBlock array.

* live against the reference's AVX-512 kernels (oracle/_ref) with
  non-trivial accumulate masks (x lane by lane, y too when soalen < veclen,
  z and t per vector, the patterns its bodies express): every family x
  {dslash, achimbdpsi} x isign x 12/18-real links, soalen == veclen and the
  xy-packed soalen = veclen / 2, FP64 and FP32, and FP16 storage against
  the reference's low-precision build: the same bits in every element;
* without a mask, bit-identical to the native-layout kernels on the same
  data (the in-place path changes addressing only).
"""
import ctypes

import numpy as np
import pytest
import torch

import qphix_blocks as qb
from _util import accumulate_mask, f16_ulps

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DIMS = (32, 8, 4, 4)
# scalars exact in FP16, so the low-precision build (which reads its scalars
# as FP16) sees the same values
KW0 = dict(alpha=3.75, mu=0.03125, muinv=0.25, rho=0.125, coeff_s=0.875, coeff_t_f=0.75, coeff_t_b=1.25,
           t_boundary=-1.0)
TORCH_FT = {0: torch.float64, 1: torch.float32, 2: torch.float16}
NATIVE = {(0, False): "qphix_amd_dslash", (0, True): "qphix_amd_dslash_achimbdpsi",
          (1, False): "qphix_amd_clov_dslash", (1, True): "qphix_amd_clov_dslash_achimbdpsi",
          (2, False): "qphix_amd_tm_dslash", (2, True): "qphix_amd_tm_dslash_achimbdpsi",
          (3, False): "qphix_amd_tm_clov_dslash", (3, True): "qphix_amd_tm_clov_dslash_achimbdpsi"}


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def call(qamd, geom, kind, achi, recon, isign, cb, out, psi, chi, g, cl, lay=None, acc=None):
    """One operator call through the C ABI: on QPhiX arrays in place when
    lay is given (qphix_amd_qphix_*), else on native spinors."""
    l, kw = qamd.lib(), KW0
    cs, ctf, ctb = kw["coeff_s"], kw["coeff_t_f"], kw["coeff_t_b"]
    if lay is not None:
        fn = getattr(l, NATIVE[(kind, achi)].replace("qphix_amd_", "qphix_amd_qphix_"))
        head = (geom.ctx, ctypes.byref(lay), recon, isign, cb, _p(out), _p(psi))
        a = _p(acc)
    else:
        fn = getattr(l, NATIVE[(kind, achi)])
        head = (geom.ctx, g.prec, recon, isign, cb, _p(out), _p(psi))
    mid = []
    if achi:
        mid.append(_p(chi))
    mid.append(g.ptr())
    if kind in (1, 3):
        mid.append(cl.ptr())
    if lay is not None:
        mid.append(a)
    if kind == 0:
        tail = ((kw["alpha"],) if achi else ()) + (cs, ctf, ctb)
    elif kind == 2:
        tail = (kw["alpha"], cs, ctf, ctb, kw["mu"]) if achi else (cs, ctf, ctb, kw["mu"], kw["muinv"])
    else:
        tail = (cs, ctf, ctb) + ((kw["rho"],) if achi else ())
    st = fn(*head, *mid, *tail, _stream())
    assert st == 0, (NATIVE[(kind, achi)], lay is not None, st)


def epilogue(L, clt, kind, achi, isign):
    if kind == 1:
        return clt.numpy() if achi else L.invert_clover(clt).numpy()
    if kind == 3:
        return L.tmclover_blocks(clt, 0.1, isign, not achi).numpy()
    return None


def qphix_mask(lay, dims, cb, mask):
    """The per-site mask in the output array's site order (block * soalen + lane)."""
    blk, lane, lex = qb.spinor_slots(lay, dims, cb)
    out = np.zeros(lay.n_spinor_blocks(dims) * lay.soalen, dtype=np.uint8)
    out[blk * lay.soalen + lane] = mask[lex]
    return out


def to_dev(a, ft):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV).view(TORCH_FT[ft])


@pytest.fixture(scope="module")
def fields(qamd):
    L = qamd.lattice
    return dict(U=L.random_gauge(DIMS, 0.3, 61).numpy(), psi=L.random_spinor(DIMS, 62).numpy(),
                chi=L.random_spinor(DIMS, 63).numpy(), clt=L.random_clover(DIMS, 64))


CONFIGS = ["avx512_f64", "avx512_f64_s4", "avx512_f32", "avx512_f32_s8", "avx512_f16"]


@pytest.mark.parametrize("config", CONFIGS)
def test_inplace_equals_reference_kernels_with_masks(qamd, fields, config):
    from oracle import ref
    if not ref.available(config):
        pytest.skip("oracle/_ref not built")
    L = qamd.lattice
    S, V = ref.soalen(config), ref.veclen(config)
    ft = 2 if ref.lib(config).ref_lowprec() else (1 if ref.lib(config).ref_precision() == 1 else 0)
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    U, psi, chi, clt = fields["U"], fields["psi"], fields["chi"], fields["clt"]
    worst = 0.0
    for kind in range(4):
        for achi in (False, True):
            for isign, cb in ((1, 0), (-1, 1)):
                clov = epilogue(L, clt, kind, achi, isign)
                kw = dict(KW0, kind=kind, achi=achi, chi=chi, clov=clov)
                mask = accumulate_mask(DIMS, cb, S, V, 7 * kind + 3 * achi + cb)
                lay_q = qb.Layout(ft, V, S)
                acc = torch.from_numpy(qphix_mask(lay_q, DIMS, cb, mask)).to(DEV)
                psi_q = to_dev(ref.pack(config, 0, DIMS, psi, 1 - cb), ft)
                chi_q = to_dev(ref.pack(config, 0, DIMS, chi, cb), ft)
                for recon in (12, 18):
                    print(f"case in-place {config} kind={kind} achi={achi} isign={isign} recon={recon}",
                          flush=True)
                    lay = qamd.qphix_layout(ft, V, S, compress12=recon == 12)
                    g = geom.gauge_from_qphix(to_dev(ref.pack(config, 1, DIMS, U, cb, recon), ft), lay, ft,
                                              recon, cb)
                    cl = None
                    if kind == 1:
                        cl = geom.clover_from_qphix(to_dev(ref.pack(config, 2, DIMS, clov, cb), ft), lay, ft, cb)
                    if kind == 3:
                        cl = geom.full_clover_from_qphix(to_dev(ref.pack(config, 3, DIMS, clov, cb), ft), lay,
                                                         ft, cb)
                    out_q = torch.zeros_like(psi_q)
                    call(qamd, geom, kind, achi, recon, isign, cb, out_q, psi_q, chi_q, g, cl, lay=lay, acc=acc)
                    got = out_q.cpu().numpy()
                    want_lex, _ = ref.apply(config, DIMS, psi, U, isign, cb, recon=recon, mask=mask, **kw)
                    want = ref.pack(config, 0, DIMS, want_lex, cb)
                    if ft == 2:
                        same = float(np.mean(f16_ulps(got, want) == 0))
                        worst = max(worst, 1.0 - same)
                    else:
                        err = float(np.abs(got.astype(np.float64) - want).max() / np.abs(want).max())
                        worst = max(worst, err)
                    assert np.array_equal(got, want), (kind, achi, isign, recon, worst)
    print(f"{config}: worst {'fraction of differing FP16 elements' if ft == 2 else 'relative error'} "
          f"{worst:.3e}", flush=True)


@pytest.mark.parametrize("ft,V,S", [(0, 8, 4), (1, 16, 16), (2, 16, 8)], ids=["f64_s4", "f32_s16", "f16_s8"])
def test_inplace_bit_identical_to_native(qamd, fields, ft, V, S):
    """No mask, padded layouts: the in-place kernels give exactly the bits of
    the native kernels on the same data (converted with the layout
    converters), for every family and kernel, both isigns, 12/18 links."""
    L = qamd.lattice
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    U, psi, chi, clt = (torch.from_numpy(fields[k]).to(DEV) if k != "clt" else fields[k]
                        for k in ("U", "psi", "chi", "clt"))
    lay = qamd.qphix_layout(ft, V, S, True, pad_xy=V // S, pad_xyz=2 * (V // S))
    for kind in range(4):
        for achi in (False, True):
            for isign, cb in ((1, 0), (-1, 1)):
                clov = epilogue(L, clt, kind, achi, isign)
                for recon in (12, 18):
                    g = geom.pack_gauge(U, ft, recon, cb)
                    cl = None
                    if kind == 1:
                        cl = geom.pack_clover(torch.from_numpy(clov).to(DEV), ft, cb)
                    if kind == 3:
                        cl = geom.pack_full_clover(torch.from_numpy(clov).to(DEV), ft, cb)
                    psi_n = geom.pack_spinor(psi, ft, 1 - cb)
                    chi_n = geom.pack_spinor(chi, ft, cb)
                    out_n = geom.new_spinor(ft, cb=cb)
                    call(qamd, geom, kind, achi, recon, isign, cb, out_n.data, psi_n.data, chi_n.data, g, cl)
                    psi_q = geom.spinor_to_qphix(psi_n, lay)
                    chi_q = geom.spinor_to_qphix(chi_n, lay)
                    out_q = torch.full_like(psi_q, float("nan"))
                    call(qamd, geom, kind, achi, recon, isign, cb, out_q, psi_q, chi_q, g, cl, lay=lay)
                    want = geom.spinor_to_qphix(out_n, lay, out=torch.full_like(psi_q, float("nan")))
                    # padding slots are neither read nor written: NaN in both
                    assert torch.equal(out_q.nan_to_num(7.0), want.nan_to_num(7.0)), (kind, achi, isign, recon)


def test_inplace_argument_checks(qamd):
    """Layouts the kernels cannot address and bad pointers are refused before
    any launch."""
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    l = qamd.lib()
    buf = torch.zeros(1 << 16, dtype=torch.float32, device=DEV)
    p = _p(buf)
    q = ctypes.c_void_p(buf.data_ptr() + 4096)
    r = ctypes.c_void_p(buf.data_ptr() + 8192)
    st = _stream()
    ok = qamd.qphix_layout(1, 16, 16, True)
    bf16 = qamd.qphix_layout(3, 16, 16, True)  # no QPhiX FT
    assert l.qphix_amd_qphix_dslash(geom.ctx, ctypes.byref(bf16), 12, 1, 0, p, q, r, None, 1.0, 1.0, 1.0, st) == -1
    g24 = qamd.Geometry((24, 8, 4, 4), t_boundary=-1.0)  # Lx/2 = 12: a valid layout, soalen 12 not 2^k
    s12 = qamd.qphix_layout(1, 12, 12, True)
    assert l.qphix_amd_qphix_dslash(g24.ctx, ctypes.byref(s12), 12, 1, 0, p, q, r, None, 1.0, 1.0, 1.0, st) == -1
    assert l.qphix_amd_qphix_dslash(geom.ctx, ctypes.byref(ok), 12, 1, 0, p, p, r, None, 1.0, 1.0, 1.0, st) == -1
    assert l.qphix_amd_qphix_dslash(geom.ctx, ctypes.byref(ok), 12, 2, 0, p, q, r, None, 1.0, 1.0, 1.0, st) == -1
    assert l.qphix_amd_qphix_clov_dslash(geom.ctx, ctypes.byref(ok), 12, 1, 0, p, q, r, None, None, 1.0, 1.0,
                                         1.0, st) == -1
    odd = ctypes.c_void_p(buf.data_ptr() + 2)
    assert l.qphix_amd_qphix_dslash(geom.ctx, ctypes.byref(ok), 12, 1, 0, odd, q, r, None, 1.0, 1.0, 1.0,
                                    st) == -6
    torch.cuda.synchronize()
    assert torch.count_nonzero(buf) == 0


def test_python_classes_take_qphix_arrays(qamd, fields):
    """The QPhiX-shaped classes (Dslash, ClovDslash, TMDslash, TMClovDslash)
    take QPhiXSpinor arrays where they take native spinors and give the
    native results' bits, with a mask exactly the C ABI's in-place result."""
    L = qamd.lattice
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    U = torch.from_numpy(fields["U"]).to(DEV)
    psi = torch.from_numpy(fields["psi"]).to(DEV)
    chi = torch.from_numpy(fields["chi"]).to(DEV)
    clt = fields["clt"]
    lay = qamd.qphix_layout(1, 16, 8, True)
    cb = 1
    g = geom.pack_gauge(U, 1, 12, cb)
    inv = geom.pack_clover(L.invert_clover(clt).to(DEV), 1, cb)
    clov = geom.pack_clover(clt.to(DEV), 1, cb)
    psi_n, chi_n = geom.pack_spinor(psi, 1, 1 - cb), geom.pack_spinor(chi, 1, cb)
    psi_q = qamd.QPhiXSpinor(geom.spinor_to_qphix(psi_n, lay), lay)
    chi_q = qamd.QPhiXSpinor(geom.spinor_to_qphix(chi_n, lay), lay)
    ops = [(qamd.Dslash(geom, 0.875, 1.25), lambda op, r, p, c: op.dslash(r, p, g, -1, cb)),
           (qamd.ClovDslash(geom), lambda op, r, p, c: op.dslashAChiMinusBDPsi(r, p, c, g, clov, 0.5, -1, cb, 0.125)),
           (qamd.ClovDslash(geom), lambda op, r, p, c: op.dslash(r, p, g, inv, 1, cb)),
           (qamd.TMDslash(geom, -0.25, 0.1), lambda op, r, p, c: op.dslashAChiMinusBDPsi(r, p, c, g, 3.75, 0.25, 1, cb))]
    for op, run in ops:
        out_n = geom.new_spinor(1, cb=cb)
        run(op, out_n, psi_n, chi_n)
        out_q = qamd.QPhiXSpinor(torch.zeros_like(psi_q.data), lay)
        run(op, out_q, psi_q, chi_q)
        assert torch.equal(out_q.data, geom.spinor_to_qphix(out_n, lay))
    # with a mask: the class call equals the direct C ABI call
    lay_q = qb.Layout(1, 16, 8)
    mask = accumulate_mask(DIMS, cb, 8, 16, 5)
    acc = torch.from_numpy(qphix_mask(lay_q, DIMS, cb, mask)).to(DEV)
    out_q = qamd.QPhiXSpinor(torch.zeros_like(psi_q.data), lay, accumulate=acc)
    qamd.ClovDslash(geom, 0.875, 1.25).dslash(out_q, psi_q, g, inv, 1, cb)
    direct = torch.zeros_like(psi_q.data)
    assert qamd.lib().qphix_amd_qphix_clov_dslash(geom.ctx, ctypes.byref(lay), 12, 1, cb, _p(direct), _p(psi_q.data),
                                                   g.ptr(), inv.ptr(), _p(acc), 0.875, 1.25, 1.25, _stream()) == 0
    torch.cuda.synchronize()
    assert torch.equal(out_q.data, direct)
    full = torch.zeros_like(psi_q.data)
    qamd.ClovDslash(geom, 0.875, 1.25).dslash(qamd.QPhiXSpinor(full, lay), psi_q, g, inv, 1, cb)
    assert not torch.equal(full, direct)  # the mask drops hops
    with pytest.raises(qamd.QPhiXAMDError):
        qamd.Dslash(geom).dslash(out_q, psi_n, g, -1, cb)  # mixed native / QPhiX spinors


@pytest.mark.parametrize("ft,V,S", [(1, 16, 16), (1, 16, 8), (1, 8, 4), (1, 4, 2), (0, 8, 8), (0, 8, 4), (0, 4, 2),
                                    (2, 16, 16), (2, 16, 8), (2, 8, 4), (2, 4, 2)],
                         ids=["f32_s16", "f32_s8", "f32_s4", "f32_s2", "f64_s8", "f64_s4", "f64_s2",
                              "f16_s16", "f16_s8", "f16_s4", "f16_s2"])
def test_inplace_vector_and_element_access_agree(qamd, fields, ft, V, S):
    """The in-place kernels read and write vectors per lane group where the
    layout and the arrays' alignment allow it (FP32 and FP16 soalen >= 4,
    FP64 soalen >= 2) and element by element otherwise.  The same arrays at
    an offset of one element (aligned to their element type, not to the
    vector) take the element path: both give the same bits, with and without
    a mask, and without a mask the native kernels' bits."""
    L = qamd.lattice
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    U, psi, chi, clt = (torch.from_numpy(fields[k]).to(DEV) if k != "clt" else fields[k]
                        for k in ("U", "psi", "chi", "clt"))
    # padding in whole vectors of veclen / soalen blocks
    pad = dict(pad_xy=V // S, pad_xyz=3 * (V // S))
    lay = qamd.qphix_layout(ft, V, S, True, **pad)
    # the mask holds one byte per slot of the padded array
    lay_q = qb.Layout(ft, V, S, True, **pad)

    def shifted(t):
        """t's values in a buffer one element past a 16-byte boundary"""
        buf = torch.full((t.numel() + 1,), float("nan"), dtype=t.dtype, device=DEV)
        view = buf[1:].view(t.shape)
        view.copy_(t)
        assert view.data_ptr() % (8 if ft == 2 else 16) != 0  # not aligned to the lane group's vector
        return view

    for kind, achi in ((1, True), (3, False), (2, True), (0, False)):
        for isign, cb in ((1, 0), (-1, 1)):
            print(f"case vector/element ft={ft} V={V} S={S} kind={kind} achi={achi} isign={isign}", flush=True)
            clov = epilogue(L, clt, kind, achi, isign)
            g = geom.pack_gauge(U, ft, 12, cb)
            cl = None
            if kind == 1:
                cl = geom.pack_clover(torch.from_numpy(clov).to(DEV), ft, cb)
            if kind == 3:
                cl = geom.pack_full_clover(torch.from_numpy(clov).to(DEV), ft, cb)
            psi_n, chi_n = geom.pack_spinor(psi, ft, 1 - cb), geom.pack_spinor(chi, ft, cb)
            out_n = geom.new_spinor(ft, cb=cb)
            call(qamd, geom, kind, achi, 12, isign, cb, out_n.data, psi_n.data, chi_n.data, g, cl)
            psi_q, chi_q = geom.spinor_to_qphix(psi_n, lay), geom.spinor_to_qphix(chi_n, lay)
            want = geom.spinor_to_qphix(out_n, lay, out=torch.full_like(psi_q, float("nan")))
            mask = accumulate_mask(DIMS, cb, S, V, 5 * kind + cb + 1)
            acc = torch.from_numpy(qphix_mask(lay_q, DIMS, cb, mask)).to(DEV)
            assert acc.numel() == psi_q.numel() // 24
            for a in (None, acc):
                out_v = torch.full_like(psi_q, float("nan"))
                call(qamd, geom, kind, achi, 12, isign, cb, out_v, psi_q, chi_q, g, cl, lay=lay, acc=a)
                out_e = shifted(torch.full_like(psi_q, float("nan")))
                call(qamd, geom, kind, achi, 12, isign, cb, out_e, shifted(psi_q), shifted(chi_q), g, cl, lay=lay,
                     acc=a)
                assert torch.equal(out_v.nan_to_num(7.0), out_e.nan_to_num(7.0)), (kind, achi, isign, a is None)
                if a is None:
                    assert torch.equal(out_v.nan_to_num(7.0), want.nan_to_num(7.0)), (kind, achi, isign)


@pytest.mark.parametrize("ft,V,S", [(1, 8, 8), (1, 16, 4), (0, 8, 4), (2, 16, 8)],
                         ids=["f32_s8", "f32_s4", "f64_s4", "f16_s8"])
def test_inplace_vector_form_without_lds_rows(qamd, ft, V, S):
    """Lx/2 = 24 (the 48^3 lattices): a work-group's rows plus their halo do
    not fit the LDS staging array, so the vector form reads the x-neighbours
    element by element and the y/z/t neighbours by quad.  Same bits as the
    element form and, without a mask, as the native kernels."""
    dims = (48, 4, 4, 2)
    L = qamd.lattice
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    U = L.random_gauge(dims, 0.3, 71).to(DEV)
    psi, chi = L.random_spinor(dims, 72).to(DEV), L.random_spinor(dims, 73).to(DEV)
    clt = L.random_clover(dims, 74)
    lay = qamd.qphix_layout(ft, V, S, True)
    lay_q = qb.Layout(ft, V, S)

    def shifted(t):
        buf = torch.full((t.numel() + 1,), float("nan"), dtype=t.dtype, device=DEV)
        view = buf[1:].view(t.shape)
        view.copy_(t)
        return view

    for kind, achi in ((1, True), (3, False), (0, True)):
        for isign, cb in ((1, 1), (-1, 0)):
            print(f"case no-LDS ft={ft} V={V} S={S} kind={kind} achi={achi} isign={isign}", flush=True)
            clov = epilogue(L, clt, kind, achi, isign)
            g = geom.pack_gauge(U, ft, 12, cb)
            cl = None
            if kind == 1:
                cl = geom.pack_clover(torch.from_numpy(clov).to(DEV), ft, cb)
            if kind == 3:
                cl = geom.pack_full_clover(torch.from_numpy(clov).to(DEV), ft, cb)
            psi_n, chi_n = geom.pack_spinor(psi, ft, 1 - cb), geom.pack_spinor(chi, ft, cb)
            out_n = geom.new_spinor(ft, cb=cb)
            call(qamd, geom, kind, achi, 12, isign, cb, out_n.data, psi_n.data, chi_n.data, g, cl)
            psi_q, chi_q = geom.spinor_to_qphix(psi_n, lay), geom.spinor_to_qphix(chi_n, lay)
            want = geom.spinor_to_qphix(out_n, lay)
            mask = accumulate_mask(dims, cb, S, V, 3 * kind + cb)
            acc = torch.from_numpy(qphix_mask(lay_q, dims, cb, mask)).to(DEV)
            assert acc.numel() == psi_q.numel() // 24
            for a in (None, acc):
                out_v = torch.zeros_like(psi_q)
                call(qamd, geom, kind, achi, 12, isign, cb, out_v, psi_q, chi_q, g, cl, lay=lay, acc=a)
                out_e = shifted(torch.zeros_like(psi_q))
                call(qamd, geom, kind, achi, 12, isign, cb, out_e, shifted(psi_q), shifted(chi_q), g, cl, lay=lay,
                     acc=a)
                assert torch.equal(out_v, out_e), (kind, achi, isign, a is None)
                if a is None:
                    assert torch.equal(out_v, want), (kind, achi, isign)
