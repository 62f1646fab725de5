"""GPU kernels against the reference's own kernels, live, on 32x16x16x16
(the volume of 16^3 x 32, with Lx/2 a multiple of the AVX-512 FP32 SOALEN 16).

The reference's generated AVX-512 bodies (oracle/_ref, built from the
qphix-codegen sources by oracle/ref.mk) and the HIP kernels run on the same
inputs: every family x {dslash, achimbdpsi} x isign x 12/18-real links, on
both checkerboards, with anisotropic coefficients and antiperiodic time.

* FP64 (avx512_f64), FP32 (avx512_f32) and FP16 storage (avx512_f16, FP16
  storage of every field and scalar, FP32 arithmetic): the same bits in
  every element.  The kernels follow the reference's FP32/FP64 rounding
  sequence operation by operation (its generated bodies fix the order of
  every product and fused multiply-add), and FP16 results are rounded to
  FP32 before the conversion to FP16, as the reference stores them.

The same with the reference's twisted-boundary bodies ("tttt") and a complex
phase per dimension.

Skipped where the reference build or AVX-512 is absent.
"""
import numpy as np
import pytest
import torch

from _util import cb_mask, f16_ulps

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DIMS = (32, 16, 16, 16)
# scalars exact in FP16, so the low-precision build (which reads its scalars
# as FP16) sees the same values
P = dict(coeff_s=1.0, coeff_t_f=0.75, coeff_t_b=1.25, t_boundary=-1.0, alpha=3.75, mu=0.03125,
         muinv=0.25, rho=0.125)


def _avx512():
    try:
        return "avx512f" in open("/proc/cpuinfo").read()
    except OSError:
        return False


def _ref(config):
    from oracle import ref
    if not ref.available(config):
        pytest.skip(f"oracle/_ref/libref_{config}.so not built")
    if not _avx512():
        pytest.skip("host CPU has no AVX-512")
    return ref


@pytest.fixture(scope="module")
def lat(qamd):
    L = qamd.lattice
    clov = L.random_clover(DIMS, 1203)
    return dict(U=L.random_gauge(DIMS, 0.3, 1200).numpy(), psi=L.random_spinor(DIMS, 1201).numpy(),
                chi=L.random_spinor(DIMS, 1202).numpy(), clov=clov.numpy(),
                inv=L.invert_clover(clov).numpy(),
                B={s: L.tmclover_blocks(clov, 0.1, s, True).numpy() for s in (1, -1)},
                F={s: L.tmclover_blocks(clov, 0.1, s, False).numpy() for s in (1, -1)})


def _epi(lat, kind, achi, isign):
    if kind == 1:
        return lat["clov"] if achi else lat["inv"]
    if kind == 3:
        return (lat["F"] if achi else lat["B"])[isign]
    return None


def _quantized(lat):
    """Every field rounded to FP16 (exactly representable afterwards, so the
    reference driver's own conversion leaves it unchanged)."""
    h = lambda a: np.asarray(a).astype(np.float16).astype(np.float64)
    out = {k: h(v) for k, v in lat.items() if not isinstance(v, dict)}
    for k in ("B", "F"):
        out[k] = {s: h(v) for s, v in lat[k].items()}
    return out


def _gpu(qamd, geom, prec, kind, achi, isign, cb, recon, lat):
    lex = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    g = geom.pack_gauge(lex(lat["U"]), prec, recon, cb)
    psi = geom.pack_spinor(lex(lat["psi"]), prec, 1 - cb)
    chi = geom.pack_spinor(lex(lat["chi"]), prec, cb)
    res = geom.new_spinor(prec, cb=cb)
    lib = qamd.lib()
    a = (geom.ctx, prec, recon, isign, cb, res.ptr(), psi.ptr())
    cs, ctf, ctb = P["coeff_s"], P["coeff_t_f"], P["coeff_t_b"]
    ep = _epi(lat, kind, achi, isign)
    if kind == 0:
        rc = (lib.qphix_amd_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), P["alpha"], cs, ctf, ctb, None) if achi
              else lib.qphix_amd_dslash(*a, g.ptr(), cs, ctf, ctb, None))
    elif kind == 1:
        cl = geom.pack_clover(lex(ep), prec, cb)
        rc = (lib.qphix_amd_clov_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), cl.ptr(), cs, ctf, ctb, P["rho"], None)
              if achi else lib.qphix_amd_clov_dslash(*a, g.ptr(), cl.ptr(), cs, ctf, ctb, None))
    elif kind == 2:
        rc = (lib.qphix_amd_tm_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), P["alpha"], cs, ctf, ctb, P["mu"], None)
              if achi else lib.qphix_amd_tm_dslash(*a, g.ptr(), cs, ctf, ctb, P["mu"], P["muinv"], None))
    else:
        cl = geom.pack_full_clover(lex(ep), prec, cb)
        rc = (lib.qphix_amd_tm_clov_dslash_achimbdpsi(*a, chi.ptr(), g.ptr(), cl.ptr(), cs, ctf, ctb, P["rho"],
                                                      None)
              if achi else lib.qphix_amd_tm_clov_dslash(*a, g.ptr(), cl.ptr(), cs, ctf, ctb, None))
    assert rc == 0
    return geom.unpack_spinor(res).cpu().numpy()


@pytest.mark.parametrize("config", ["avx512_f64", "avx512_f32", "avx512_f16"])
def test_gpu_equals_reference_kernels(qamd, lat, config):
    ref = _ref(config)
    prec = {"avx512_f64": qamd.F64, "avx512_f32": qamd.F32, "avx512_f16": qamd.F16}[config]
    if prec == qamd.F16:  # both sides read the same FP16-rounded inputs
        lat = _quantized(lat)
    geom = qamd.Geometry(DIMS, t_boundary=P["t_boundary"])
    nthr = 16
    worst = 0.0
    for kind in range(4):
        for achi in (0, 1):
            for isign, cb in ((1, 0), (-1, 1), (1, 1)):
                for recon in (12, 18):
                    ep = _epi(lat, kind, achi, isign)
                    want, _ = ref.apply(config, DIMS, lat["psi"], lat["U"], isign, cb, kind=kind, achi=achi,
                                        chi=lat["chi"], clov=ep, recon=recon, nthreads=nthr, **P)
                    got = _gpu(qamd, geom, prec, kind, achi, isign, cb, recon, lat)
                    m = cb_mask(DIMS, cb)
                    case = (config, kind, achi, isign, cb, recon)
                    if prec == qamd.F16:
                        ulps = f16_ulps(got[m].astype(np.float16), want[m].astype(np.float16))
                        same = float(np.mean(ulps == 0))
                        print(f"  {case}: identical {same:.6f}, max {ulps.max()} ulp", flush=True)
                        assert same == 1.0, (case, ulps.max(), same)
                        worst = max(worst, 1.0 - same)
                    else:
                        err = np.abs(got[m] - want[m]).max() / np.abs(want[m]).max()
                        same = float(np.mean(got[m] == want[m]))
                        print(f"  {case}: identical {same:.6f}, max relative difference {err:.3g}", flush=True)
                        assert same == 1.0, (case, err)
                        worst = max(worst, err)
    print(f"{config}: worst {'fraction of differing FP16 elements' if prec == qamd.F16 else 'relative error'}"
          f" {worst:.3g}", flush=True)


# forward-hop phases per dimension, exactly representable in FP16 (a plain
# complex factor; the kernels do not assume unit modulus)
TBC_PHASES = [0.75 + 0.5j, 0.5 - 0.875j, -0.625 + 0.75j, 0.875 + 0.25j]


@pytest.mark.parametrize("config", ["avx512_f64", "avx512_f32", "avx512_f16"])
def test_gpu_equals_reference_twisted_bc_kernels(qamd, lat, config):
    """The reference's bodies generated with twisted boundary conditions in
    every dimension ("tttt", codegen/dslash.cc:595-630) against the GPU
    kernels with qphix_amd_set_tbc on all four dimensions: every family x
    {dslash, achimbdpsi}, both isigns and checkerboards, 12/18-real links, to
    the same bars as above."""
    ref = _ref(config)
    prec = {"avx512_f64": qamd.F64, "avx512_f32": qamd.F32, "avx512_f16": qamd.F16}[config]
    if prec == qamd.F16:
        lat = _quantized(lat)
    geom = qamd.Geometry(DIMS, t_boundary=P["t_boundary"])
    geom.set_tbc([1, 1, 1, 1], TBC_PHASES)
    for kind in range(4):
        for achi in (0, 1):
            for isign, cb in ((1, 0), (-1, 1)):
                for recon in (12, 18):
                    ep = _epi(lat, kind, achi, isign)
                    want, _ = ref.apply(config, DIMS, lat["psi"], lat["U"], isign, cb, kind=kind, achi=achi,
                                        chi=lat["chi"], clov=ep, recon=recon, nthreads=16,
                                        tbc_phases=TBC_PHASES, **P)
                    got = _gpu(qamd, geom, prec, kind, achi, isign, cb, recon, lat)
                    m = cb_mask(DIMS, cb)
                    case = (config, kind, achi, isign, cb, recon)
                    if prec == qamd.F16:
                        ulps = f16_ulps(got[m].astype(np.float16), want[m].astype(np.float16))
                        same = float(np.mean(ulps == 0))
                        assert same == 1.0, (case, ulps.max(), same)
                    else:
                        err = np.abs(got[m] - want[m]).max() / np.abs(want[m]).max()
                        assert np.array_equal(got[m], want[m]), (case, err)
