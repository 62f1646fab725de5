"""The oracle against the reference's own kernels.

* committed fixtures written by tests/golden/make_golden.py from the
  reference's generated scalar FP64 and AVX-512 FP32 kernel bodies;
* when oracle/_ref is built (it needs /root/reference at build time), the
  reference kernels of every configuration live, on seeded lattices, for
  every family x {dslash, achimbdpsi} x isign x cb x 12/18 links.

Tolerances: 1e-13 relative for FP64 reference kernels, 1e-6 for FP32
(relative = max-abs difference / max-abs value).
"""
import os

import numpy as np
import pytest

from _util import accumulate_mask, cb_mask, f16_ulps, f16_within, masked_oracle, recon12_links

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def params(f):
    return {str(k): float(v) for k, v in zip(f["param_names"], f["params"])}


def epilogue(qamd, f, kind, achi, isign):
    import torch
    L = qamd.lattice
    clov = torch.from_numpy(np.asarray(f["clov"], dtype=np.float64))
    if kind == 1:
        if achi:
            return clov.numpy()
        return f["inv"].astype(np.float64) if "inv" in f else L.invert_clover(clov).numpy()
    if kind == 3:
        return L.tmclover_blocks(clov, float(f["tm_mu"]), isign, not achi).numpy()
    return None


def test_oracle_matches_reference_scalar_f64_fixture(qamd, orc):
    f = load("ref_scalar_f64_4x4.npz")
    dims = tuple(int(d) for d in f["dims"])
    p = params(f)
    for case, sites, ref_out in zip(f["cases"], f["sites"], f["outs"]):
        kind, achi, isign, cb, recon = (int(c) for c in case)
        out = orc.dslash_sites(dims, sites, f["psi"], f["U"], isign, kind=kind, achi=bool(achi),
                               chi=f["chi"], clov=epilogue(qamd, f, kind, achi, isign), **p)
        err = np.abs(out - ref_out).max() / np.abs(ref_out).max()
        assert err < 1e-13, (case, err)


def test_oracle_matches_reference_avx512_f32_fixture(qamd, orc):
    f = load("ref_avx512_f32_32x2.npz")
    dims = tuple(int(d) for d in f["dims"])
    p = params(f)
    d64 = {k: np.asarray(f[k], dtype=np.float64) for k in ("U", "psi", "chi", "clov", "inv")}
    for case, ref_out in zip(f["cases"], f["outs"]):
        kind, achi, isign, cb, recon = (int(c) for c in case)
        clov = d64["clov"] if achi else d64["inv"]
        out = orc.dslash(dims, d64["psi"], d64["U"], isign, cb, kind=kind, achi=bool(achi),
                         chi=d64["chi"], clov=clov, **p)
        m = cb_mask(dims, cb)
        err = np.abs(out[m] - ref_out).max() / np.abs(ref_out).max()
        assert err < 1e-6, (case, err)


REF_DIMS = {"scalar_f64": (8, 8, 8, 8), "avx512_f64": (16, 4, 4, 8), "avx512_f32": (32, 4, 4, 4),
            "avx2_f64": (8, 8, 8, 8), "avx2_f32": (16, 4, 4, 8),
            "avx512_f64_s4": (16, 4, 4, 8), "avx512_f32_s8": (32, 4, 4, 4)}


def _cpu_has(flag):
    try:
        return flag in open("/proc/cpuinfo").read()
    except OSError:
        return False


@pytest.mark.parametrize("config", list(REF_DIMS))
def test_oracle_matches_reference_kernels_live(qamd, orc, config):
    from oracle import ref
    if not ref.available(config):
        pytest.skip("oracle/_ref not built (needs the reference sources at build time)")
    if "avx512" in config and not _cpu_has("avx512f"):
        pytest.skip("host CPU has no AVX-512")
    if "avx2" in config and not _cpu_has("avx2"):
        pytest.skip("host CPU has no AVX2")
    L = qamd.lattice
    dims = REF_DIMS[config]
    U = L.random_gauge(dims, 0.3, 31).numpy()
    psi = L.random_spinor(dims, 32).numpy()
    chi = L.random_spinor(dims, 33).numpy()
    clt = L.random_clover(dims, 34)
    tol = 1e-13 if config.endswith("f64") else 1e-6
    kw0 = dict(alpha=3.7, mu=0.027, muinv=0.26, rho=0.05, coeff_s=0.9, coeff_t_f=1.1, coeff_t_b=1.2,
               t_boundary=-1.0)
    for kind in range(4):
        for achi in (0, 1):
            for isign in (1, -1):
                for cb in (0, 1):
                    kw = dict(kw0, kind=kind, achi=bool(achi), chi=chi)
                    if kind == 1:
                        kw["clov"] = clt.numpy() if achi else L.invert_clover(clt).numpy()
                    if kind == 3:
                        kw["clov"] = L.tmclover_blocks(clt, 0.1, isign, not achi).numpy()
                    a = orc.dslash(dims, psi, U, isign, cb, **kw)
                    m = cb_mask(dims, cb)
                    for recon in (12, 18):
                        b, _ = ref.apply(config, dims, psi, U, isign, cb, recon=recon, **kw)
                        err = np.abs(a[m] - b[m]).max() / np.abs(a[m]).max()
                        assert err < tol, (kind, achi, isign, cb, recon, err)


@pytest.mark.parametrize("config", ["avx512_f64", "avx512_f32", "avx2_f64"])
def test_oracle_matches_reference_twisted_bc_kernels_live(qamd, orc, config):
    """The reference's bodies generated with twisted boundary conditions in
    every dimension ("tttt", codegen/dslash.cc:595-630): a distinct U(1)
    phase per dimension on every forward hop, its conjugate on backward hops
    (codegen/dslash_common.cc:1491-1524), for every family, both isigns and
    checkerboards, 12/18-real links, against the oracle's tbc."""
    from oracle import ref
    if not ref.available(config):
        pytest.skip("oracle/_ref not built (needs the reference sources at build time)")
    if "avx512" in config and not _cpu_has("avx512f"):
        pytest.skip("host CPU has no AVX-512")
    if "avx2" in config and not _cpu_has("avx2"):
        pytest.skip("host CPU has no AVX2")
    L = qamd.lattice
    dims = REF_DIMS[config]
    U = L.random_gauge(dims, 0.3, 41).numpy()
    psi = L.random_spinor(dims, 42).numpy()
    chi = L.random_spinor(dims, 43).numpy()
    clt = L.random_clover(dims, 44)
    phases = [np.exp(1j * t) for t in (0.3, -0.7, 1.1, 2.0)]
    tol = 1e-13 if config.endswith("f64") else 1e-6
    kw0 = dict(alpha=3.7, mu=0.027, muinv=0.26, rho=0.05, coeff_s=0.9, coeff_t_f=1.1, coeff_t_b=1.2,
               t_boundary=-1.0)
    for kind in range(4):
        for achi in (0, 1):
            for isign, cb in ((1, 0), (-1, 1), (1, 1)):
                kw = dict(kw0, kind=kind, achi=bool(achi), chi=chi)
                if kind == 1:
                    kw["clov"] = clt.numpy() if achi else L.invert_clover(clt).numpy()
                if kind == 3:
                    kw["clov"] = L.tmclover_blocks(clt, 0.1, isign, not achi).numpy()
                a = orc.dslash(dims, psi, U, isign, cb, tbc=[1, 1, 1, 1], tbc_phases=phases, **kw)
                plain = orc.dslash(dims, psi, U, isign, cb, **kw)
                m = cb_mask(dims, cb)
                assert np.abs(a[m] - plain[m]).max() > 0.1 * np.abs(a[m]).max()  # the phases matter
                for recon in (12, 18):
                    b, _ = ref.apply(config, dims, psi, U, isign, cb, recon=recon, tbc_phases=phases, **kw)
                    err = np.abs(a[m] - b[m]).max() / np.abs(a[m]).max()
                    assert err < tol, (kind, achi, isign, cb, recon, err)


def _f16_fixture_case(orc, f, case):
    """The FP64 oracle's result for one case of the FP16 fixture, rounded to
    FP16, with the links the 12-real kernels apply."""
    kind, achi, isign, cb, recon = (int(c) for c in case)
    p = params(f)
    dims = tuple(int(d) for d in f["dims"])
    d = {k: np.asarray(f[k], dtype=np.float64) for k in ("U", "psi", "chi", "clov", "inv", "fB", "fF")}
    U = recon12_links(d["U"]) if recon == 12 else d["U"]
    clov = {1: d["clov"] if achi else d["inv"], 3: d["fF"] if achi else d["fB"]}.get(kind)
    out = orc.dslash(dims, d["psi"], U, isign, cb, kind=kind, achi=bool(achi), chi=d["chi"], clov=clov,
                     **{k: p[k] for k in ("coeff_s", "coeff_t_f", "coeff_t_b", "t_boundary", "alpha",
                                          "mu", "muinv", "rho")})
    return out[cb_mask(dims, cb)]


def test_oracle_matches_reference_avx512_f16_fixture(qamd, orc):
    """The reference's low-precision kernels (FP16 storage of every field and
    scalar, FP32 arithmetic) equal the exact operator on their FP16 inputs,
    rounded to FP16: within 1 unit in the last place (or FP32 accumulation
    error where cancellation leaves a small value, f16_within), and the same
    bits in >= 99% of the elements."""
    f = load("ref_avx512_f16_32x2.npz")
    for case, ref_out in zip(f["cases"], f["outs"]):
        exact = _f16_fixture_case(orc, f, case)
        assert f16_within(exact, ref_out).all(), case
        ulps = f16_ulps(exact.astype(np.float16), ref_out)
        assert np.mean(ulps == 0) >= 0.99, (case, np.mean(ulps == 0))


def test_f16_agreement_criterion_rejects_real_errors():
    """f16_within accepts one FP16 ulp, and FP32-level noise on near-zero
    elements, but not two ulps on an element of the field's scale."""
    want = np.array([19.0, 1.0, -0.25, 1e-4], dtype=np.float16)
    assert f16_within(want, want).all()
    one = want.copy()
    one[0] = np.nextafter(want[0], np.float16(30))
    assert f16_within(one, want).all()
    two = one.copy()
    two[0] = np.nextafter(one[0], np.float16(30))
    assert not f16_within(two, want)[0]
    near0 = want.copy()
    near0[3] = np.float16(1e-4 + 4e-7)  # about 9 FP16 ulps of 1e-4
    assert f16_within(near0, want).all()
    assert not f16_within(np.float16(1e-4 + 2e-4) * np.ones(4, np.float16), want)[3]


@pytest.mark.parametrize("config", ["avx512_f64", "avx512_f64_s4", "avx512_f32_s8", "avx2_f64"])
def test_reference_accumulate_masks_live(qamd, orc, config):
    """The reference's bodies with non-trivial accumulate[] masks (driver
    oracle/ref_driver.cpp) against the oracle's exact direction-by-direction
    decomposition summed under the same per-site mask: every family,
    dslash and achimbdpsi, both isigns and checkerboards, 12/18-real links,
    soalen == veclen and the xy-packed soalen = veclen / 2 (where y is masked
    lane by lane too), hardware-masked (AVX-512) and blended (AVX2) builds."""
    from oracle import ref
    if not ref.available(config):
        pytest.skip("oracle/_ref not built (needs the reference sources at build time)")
    if "avx512" in config and not _cpu_has("avx512f"):
        pytest.skip("host CPU has no AVX-512")
    if "avx2" in config and not _cpu_has("avx2"):
        pytest.skip("host CPU has no AVX2")
    L = qamd.lattice
    dims = REF_DIMS[config]
    S, V = ref.soalen(config), ref.veclen(config)
    U = L.random_gauge(dims, 0.3, 51).numpy()
    psi = L.random_spinor(dims, 52).numpy()
    chi = L.random_spinor(dims, 53).numpy()
    clt = L.random_clover(dims, 54)
    tol = 1e-13 if "f64" in config else 1e-6
    kw0 = dict(alpha=3.7, mu=0.027, muinv=0.26, rho=0.05, coeff_s=0.9, coeff_t_f=1.1, coeff_t_b=1.2,
               t_boundary=-1.0)
    for kind in range(4):
        for achi in (0, 1):
            for isign, cb in ((1, 0), (-1, 1)):
                kw = dict(kw0, kind=kind, achi=bool(achi), chi=chi)
                if kind == 1:
                    kw["clov"] = clt.numpy() if achi else L.invert_clover(clt).numpy()
                if kind == 3:
                    kw["clov"] = L.tmclover_blocks(clt, 0.1, isign, not achi).numpy()
                mask = accumulate_mask(dims, cb, S, V, 100 * kind + 10 * achi + cb)
                want = masked_oracle(orc, dims, psi, U, isign, cb, mask, **kw)
                full = orc.dslash(dims, psi, U, isign, cb, **kw)
                m = cb_mask(dims, cb)
                assert np.abs(want[m] - full[m]).max() > 0.1 * np.abs(full[m]).max()  # the mask matters
                for recon in (12, 18):
                    b, _ = ref.apply(config, dims, psi, U, isign, cb, recon=recon, mask=mask, **kw)
                    err = np.abs(want[m] - b[m]).max() / np.abs(want[m]).max()
                    assert err < tol, (kind, achi, isign, cb, recon, err)
