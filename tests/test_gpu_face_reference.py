"""The library's face entry points against the reference's own face bodies,
live (oracle/_ref: dslash_face_pack_to_*, *dslash_face_unpack_from_*,
codegen/dslash.cc:433-553,664-744, under the driver's face loops), on the
same inputs, element by element.

* qphix_amd_face_pack against face_proj_dir_{plus,minus}: every face
  (2 mu + fb), both isigns, both checkerboards.
* qphix_amd_{,clov_,tm_,tm_clov_}face_finish against face_finish_dir,
  face_clov_finish_dir, tm_face_finish_dir, tm_clov_face_finish_dir: every
  family, face, isign, 12- and 18-real links, simple and twisted boundary
  bodies, on a random received face and a random partial result.

The buffers map site by site: the reference's per-call half-spinors land at
their lattice sites (the driver scatters the compressed lanes), the
library's face index f enumerates the slice's other coordinates.  The bar
is bit-identity in FP64, FP32 and FP16 storage (FP16 inputs and scalars
exactly representable): the finish kernels follow the reference's rounding
sequence (U h chains, phase, beta reconstruction into the partial result or
into a zeroed spinor, then the clover rows or the twisted term accumulated
into the partial result).  Skipped where the reference build or AVX-512 is
absent.
"""
import numpy as np
import pytest
import torch

from _util import cb_mask, face_from_buffer, face_sites, face_to_buffer, f16_ulps

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
D = (32, 4, 4, 8)
BETA, MU, MUINV = -0.875, 0.03125, 0.25   # exact in FP16
TBC_PHASES = [0.75 + 0.5j, 0.5 - 0.875j, -0.625 + 0.75j, 0.875 + 0.25j]
CONFIGS = ["avx512_f64", "avx512_f32", "avx512_f16"]


def _ref(config):
    from oracle import ref
    if not ref.available(config):
        pytest.skip(f"oracle/_ref/libref_{config}.so not built")
    try:
        if "avx512f" not in open("/proc/cpuinfo").read():
            pytest.skip("host CPU has no AVX-512")
    except OSError:
        pytest.skip("no /proc/cpuinfo")
    return ref


def _prec(qamd, config):
    return {"avx512_f64": qamd.F64, "avx512_f32": qamd.F32, "avx512_f16": qamd.F16}[config]


@pytest.fixture(scope="module")
def lat(qamd):
    L = qamd.lattice
    clov = L.random_clover(D, 3103)
    V = int(np.prod(D))
    rng = np.random.default_rng(3104)
    return dict(U=L.random_gauge(D, 0.3, 3100).numpy(), psi=L.random_spinor(D, 3101).numpy(),
                out=L.random_spinor(D, 3102).numpy(), inv=L.invert_clover(clov).numpy(),
                B={s: L.tmclover_blocks(clov, 0.1, s, True).numpy() for s in (1, -1)},
                hs=rng.standard_normal((V, 12)) * 0.5)


def _quantized(lat):
    h = lambda a: np.asarray(a).astype(np.float16).astype(np.float64)
    out = {k: h(v) for k, v in lat.items() if not isinstance(v, dict)}
    out["B"] = {s: h(v) for s, v in lat["B"].items()}
    return out


def _same(got, want, prec, qamd, case):
    if prec == qamd.F16:
        ulps = f16_ulps(np.asarray(got).astype(np.float16), np.asarray(want).astype(np.float16))
        assert ulps.max() == 0, (case, ulps.max(), float(np.mean(ulps == 0)))
    else:
        assert np.array_equal(got, want), (case, float(np.abs(got - want).max()))


def _lex(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.parametrize("config", CONFIGS)
def test_face_pack_equals_reference(qamd, lat, config):
    ref = _ref(config)
    prec = _prec(qamd, config)
    if prec == qamd.F16:
        lat = _quantized(lat)
    geom = qamd.Geometry(D)
    for isign in (1, -1):
        for cb in (0, 1):
            psi = geom.pack_spinor(_lex(lat["psi"]), prec, 1 - cb)
            for mu in range(4):
                for fb in (0, 1):
                    want = ref.face_pack(config, D, lat["psi"], isign, 2 * mu + fb, 1 - cb)
                    buf = qamd.face_pack(geom, geom.new_face(prec, mu), psi, isign, cb, mu, fb)
                    n = geom.face_sites(mu)
                    got = face_from_buffer(buf.cpu().float().numpy().astype(np.float64) if prec == qamd.F16
                                           else buf.cpu().numpy(), n, buf.element_size())
                    sites = face_sites(D, mu, fb, 1 - cb)
                    _same(got, want[sites], prec, qamd, (config, "pack", isign, cb, mu, fb))


@pytest.mark.parametrize("config", CONFIGS)
@pytest.mark.parametrize("tbc", [False, True])
def test_face_finish_equals_reference(qamd, lat, config, tbc):
    ref = _ref(config)
    prec = _prec(qamd, config)
    if prec == qamd.F16:
        lat = _quantized(lat)
    geom = qamd.Geometry(D)
    if tbc:
        geom.set_tbc([1, 1, 1, 1], TBC_PHASES)
    tdt = {qamd.F64: torch.float64, qamd.F32: torch.float32, qamd.F16: torch.float16}[prec]
    n_cases = 0
    for kind in range(4):
        for isign, cb in ((1, 0), (-1, 1)):
            ep = {1: lat["inv"], 3: lat["B"][isign]}.get(kind)
            for recon in (12, 18):
                u = geom.pack_gauge(_lex(lat["U"]), prec, recon, cb)
                cl = None
                if kind == 1:
                    cl = geom.pack_clover(_lex(ep), prec, cb)
                elif kind == 3:
                    cl = geom.pack_full_clover(_lex(ep), prec, cb)
                for mu in range(4):
                    for fb in (0, 1):
                        if (mu + fb + recon // 6) % 2:  # each face with one of the link forms
                            continue
                        want = ref.face_finish(config, D, kind, isign, recon, 2 * mu + fb, cb, lat["hs"],
                                               lat["U"], lat["out"], BETA, clov=ep, mu=MU, muinv=MUINV,
                                               tbc_phases=TBC_PHASES if tbc else None)
                        sites = face_sites(D, mu, fb, cb)
                        buf = torch.from_numpy(face_to_buffer(lat["hs"][sites], torch.tensor([], dtype=tdt)
                                                              .element_size())).to(DEV).to(tdt)
                        out = geom.pack_spinor(_lex(lat["out"]), prec, cb)
                        qamd.face_finish(geom, kind, out, buf, u, isign, cb, mu, fb, BETA, clov=cl, mu=MU,
                                         muinv=MUINV)
                        got = geom.unpack_spinor(out).cpu().numpy()
                        m = cb_mask(D, cb)
                        _same(got[m], want[m], prec, qamd, (config, tbc, kind, isign, cb, recon, mu, fb))
                        n_cases += 1
    print(f"{config} tbc={tbc}: {n_cases} face finishes identical", flush=True)
