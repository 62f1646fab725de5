"""Writes the golden fixtures in tests/golden/ from the reference's own
generated kernels (qphix-codegen bodies built by oracle/ref.mk, run through
oracle/ref_driver.cpp).  Run in a container that has /root/reference:

    make -C oracle -f ref.mk && python tests/golden/make_golden.py [f64] [f32] [f16]

ref_scalar_f64_4x4.npz   scalar FP64 kernels on a 4^4 lattice: every operator
                         family x {dslash, achimbdpsi} x two (isign, cb, recon)
                         combinations; outputs at 48 sampled sites per case.
ref_avx512_f32_32x2.npz  AVX-512 FP32 kernels (VECLEN = SOALEN = 16) on
                         32x2x2x2, clover family, outputs on the checkerboard;
                         inputs stored as the FP32 values the kernels consumed.
ref_avx512_f16_32x2.npz  AVX-512 low-precision kernels (FP16 storage of every
                         field and scalar, FP32 arithmetic; ENABLE_LOW_PRECISION)
                         on 32x2x2x2: every family x {dslash, achimbdpsi} x
                         {(+1, cb 0, 18 reals), (-1, cb 1, 12 reals)}; inputs,
                         scalars (exact in FP16) and outputs as FP16.
Each file holds the inputs, the scalar parameters and the outputs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402

PARAMS = dict(alpha=3.7037037037037037, mu=0.1 / 3.7037037037037037,
              muinv=3.7037037037037037 / (3.7037037037037037 ** 2 + 0.01), rho=0.05,
              coeff_s=0.9, coeff_t_f=1.1, coeff_t_b=1.2, t_boundary=-1.0)
TM_MU = 0.1


def lattice(dims, seed):
    import __graft_entry__ as ge
    q = ge.load_package()
    L = q.lattice
    clov = L.random_clover(dims, seed + 3)
    return dict(U=L.random_gauge(dims, 0.3, seed).numpy(), psi=L.random_spinor(dims, seed + 1).numpy(),
                chi=L.random_spinor(dims, seed + 2).numpy(), clov=clov.numpy(),
                inv=L.invert_clover(clov).numpy(),
                Bp=L.tmclover_blocks(clov, TM_MU, 1, True).numpy(),
                Bm=L.tmclover_blocks(clov, TM_MU, -1, True).numpy(),
                Fp=L.tmclover_blocks(clov, TM_MU, 1, False).numpy(),
                Fm=L.tmclover_blocks(clov, TM_MU, -1, False).numpy())


def epilogue(kind, achi, isign, lat):
    if kind == 1:
        return lat["clov"] if achi else lat["inv"]
    if kind == 3:
        s = "p" if isign > 0 else "m"
        return lat[("F" if achi else "B") + s]
    return None


def run(config, dims, lat, kind, achi, isign, cb, recon):
    return ref.apply(config, dims, lat["psi"], lat["U"], isign, cb, kind=kind, achi=achi,
                     chi=lat["chi"], clov=epilogue(kind, achi, isign, lat), recon=recon,
                     **PARAMS)[0]


F16_PARAMS = dict(alpha=3.75, mu=0.03125, muinv=0.25, rho=0.125, coeff_s=1.0, coeff_t_f=0.5,
                  coeff_t_b=0.75, t_boundary=-1.0)


def write_f16():
    """Fixture of the reference's FP16 (low-precision) kernels."""
    import __graft_entry__ as ge
    L = ge.load_package().lattice
    dims = (32, 2, 2, 2)
    h = lambda a: np.asarray(a).astype(np.float16)
    clov = L.random_clover(dims, 803)
    g = np.random.default_rng(804)
    lat = dict(U=h(L.random_gauge(dims, 0.3, 800).numpy()), psi=h(L.random_spinor(dims, 801).numpy()),
               chi=h(L.random_spinor(dims, 802).numpy()), clov=h(clov.numpy()),
               inv=h(L.invert_clover(clov).numpy()),
               fB=h(0.3 * g.normal(size=(int(np.prod(dims)), 2, 6, 6, 2)) + np.eye(6)[None, None, :, :, None]
                    * np.array([1.0, 0.0])),
               fF=h(0.3 * g.normal(size=(int(np.prod(dims)), 2, 6, 6, 2)) + np.eye(6)[None, None, :, :, None]
                    * np.array([1.0, 0.0])))
    par = np.array([(x + y + z + t) & 1 for t in range(2) for z in range(2) for y in range(2)
                    for x in range(32)])
    cases, outs = [], []
    for kind in range(4):
        for achi in (0, 1):
            ep = {1: "clov" if achi else "inv", 3: "fF" if achi else "fB"}.get(kind)
            for isign, cb, recon in ((1, 0, 18), (-1, 1, 12)):
                out = ref.apply("avx512_f16", dims, lat["psi"], lat["U"], isign, cb, kind=kind, achi=achi,
                                chi=lat["chi"], clov=None if ep is None else lat[ep], recon=recon,
                                **F16_PARAMS)[0]
                cases.append((kind, achi, isign, cb, recon))
                outs.append(out[par == cb].astype(np.float16))
    np.savez_compressed(os.path.join(HERE, "ref_avx512_f16_32x2.npz"), dims=np.array(dims),
                        cases=np.array(cases), outs=np.array(outs),
                        params=np.array([F16_PARAMS[k] for k in sorted(F16_PARAMS)]),
                        param_names=np.array(sorted(F16_PARAMS)), **lat)


def main(which=("f64", "f32", "f16")):
    if "f16" in which:
        write_f16()
    if "f64" not in which and "f32" not in which:
        return
    rng = np.random.default_rng(2024)
    dims = (4, 4, 4, 4)
    lat = lattice(dims, 500)
    par = np.array([(x + y + z + t) & 1 for t in range(4) for z in range(4) for y in range(4)
                    for x in range(4)])
    cases, sites, outs = [], [], []
    for kind in range(4):
        for achi in (0, 1):
            for isign, cb, recon in ((1, 0, 12), (-1, 1, 18)):
                out = run("scalar_f64", dims, lat, kind, achi, isign, cb, recon)
                s = np.sort(rng.choice(np.nonzero(par == cb)[0], 48, replace=False))
                cases.append((kind, achi, isign, cb, recon))
                sites.append(s)
                outs.append(out[s])
    np.savez(os.path.join(HERE, "ref_scalar_f64_4x4.npz"), dims=np.array(dims),
             cases=np.array(cases), sites=np.array(sites), outs=np.array(outs),
             U=lat["U"], psi=lat["psi"], chi=lat["chi"], clov=lat["clov"],
             params=np.array([PARAMS[k] for k in sorted(PARAMS)]), param_names=np.array(sorted(PARAMS)),
             tm_mu=TM_MU)

    dims = (32, 2, 2, 2)
    lat = lattice(dims, 600)
    for k in ("U", "psi", "chi", "clov", "inv"):
        lat[k] = lat[k].astype(np.float32).astype(np.float64)
    par = np.array([(x + y + z + t) & 1 for t in range(2) for z in range(2) for y in range(2)
                    for x in range(32)])
    cases, outs = [], []
    for achi in (0, 1):
        for isign, cb in ((1, 0), (-1, 1)):
            out = run("avx512_f32", dims, lat, 1, achi, isign, cb, 12)
            cases.append((1, achi, isign, cb, 12))
            outs.append(out[par == cb].astype(np.float32))
    f32 = {k: lat[k].astype(np.float32) for k in ("U", "psi", "chi", "clov", "inv")}
    np.savez(os.path.join(HERE, "ref_avx512_f32_32x2.npz"), dims=np.array(dims), cases=np.array(cases),
             outs=np.array(outs), U=f32["U"], psi=f32["psi"], chi=f32["chi"], clov=f32["clov"],
             inv=f32["inv"],
             params=np.array([PARAMS[k] for k in sorted(PARAMS)]), param_names=np.array(sorted(PARAMS)),
             tm_mu=TM_MU)


if __name__ == "__main__":
    main(tuple(sys.argv[1:]) or ("f64", "f32", "f16"))
