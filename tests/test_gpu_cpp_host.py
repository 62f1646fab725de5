"""GPU test of the C++ host classes (include/qphix_amd.hpp) through the C++
host program tests/cpp/host_api_check, built with the library.

16x16x8x8: every family (Dslash, ClovDslash, TMDslash, TMClovDslash) x
{dslash, dslashAChiMinusBDPsi} x isign +1 (cb 0) / -1 (cb 1), FP64 and FP32
storage with 12-real links, anisotropic coefficients and antiperiodic time,
against the FP64 oracle (1e-13 / 1e-6 relative), and each call again in the
classes' in-place mode on QPhiX FourSpinorBlock arrays (use_qphix_arrays),
which the host program requires to give the same bits; then the even-odd
clover
operator's FP64 CG and FP16 mixed CG to a true residual of 1e-10 and their
agreement; its force and force_clover at the FP64 solution against the same
calls through the Python binding."""
import os
import re
import subprocess

import numpy as np
import pytest
import torch

from _util import cb_mask

pytestmark = pytest.mark.gpu

DEV = "cuda:0"

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "host_api_check")
DIMS = (16, 16, 8, 8)
# the scalars the host program uses
ANISO_S, ANISO_T, ALPHA, BETA, RHO, MASS, TMU = 0.9, 1.1, 4.2, 0.25, 0.05, 0.2, 0.1


def test_cpp_classes_against_oracle(qamd, orc, tmp_path):
    assert os.path.exists(EXE), "tests/cpp/host_api_check is not built (build() builds it)"
    L = qamd.lattice
    U = L.random_gauge(DIMS, 0.3, 3100).numpy()
    psi = L.random_spinor(DIMS, 3101).numpy()
    chi = L.random_spinor(DIMS, 3102).numpy()
    clov_t = L.random_clover(DIMS, 3103)
    clov, inv = clov_t.numpy(), L.invert_clover(clov_t).numpy()
    full = {(s, invert): L.tmclover_blocks(clov_t, TMU, s, invert).numpy() for s in (1, -1) for invert in (True, False)}
    inp = tmp_path / "in.bin"
    with open(inp, "wb") as f:
        np.asarray(DIMS, dtype=np.int32).tofile(f)
        for a in (U, psi, chi, clov, inv, full[1, True], full[1, False], full[-1, True], full[-1, False]):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)
    out = tmp_path / "out.bin"
    print("running host_api_check gpu", DIMS, flush=True)
    r = subprocess.run([EXE, "gpu", str(inp), str(out)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    V = int(np.prod(DIMS))
    res = np.fromfile(out, dtype=np.float64).reshape(-1, V, 4, 3, 2)
    assert res.shape[0] == 2 * 2 * 4 * 2
    k = 0
    for prec, tol in (("f64", 1e-13), ("f32", 1e-6)):
        for isign in (1, -1):
            cb = 0 if isign > 0 else 1
            m = cb_mask(DIMS, cb)
            for kind in range(4):
                for achi in (0, 1):
                    c = BETA if achi else 1.0
                    kw = dict(kind=kind, achi=bool(achi), coeff_s=c * ANISO_S, coeff_t_f=c * ANISO_T,
                              coeff_t_b=c * ANISO_T, t_boundary=-1.0, alpha=ALPHA, rho=RHO, chi=chi)
                    if kind == 1:
                        kw["clov"] = clov if achi else inv
                    elif kind == 2:
                        a_op = 4.0 + MASS
                        kw["mu"] = TMU / ALPHA if achi else TMU / a_op
                        kw["muinv"] = a_op / (a_op ** 2 + TMU ** 2)
                    elif kind == 3:
                        kw["clov"] = full[isign, not achi]
                    want = orc.dslash(DIMS, psi, U, isign, cb, **kw)
                    got = res[k]
                    k += 1
                    err = np.abs(got[m] - want[m]).max() / np.abs(want[m]).max()
                    assert err < tol, (prec, isign, kind, achi, err)
                    assert np.all(got[~m] == 0)
    line = re.search(r"cg f64_iters (\d+) f64_rsd (\S+) f16_iters (\d+) f16_rsd (\S+) f16_reliable (\d+) "
                     r"rel_diff (\S+)", r.stdout)
    assert line, r.stdout
    f64_rsd, f16_rsd, diff = float(line.group(2)), float(line.group(4)), float(line.group(6))
    print(r.stdout.strip(), flush=True)
    assert f64_rsd < 1e-10 and f16_rsd < 1e-10
    assert diff < 1e-8
    # EvenOddOperator::force / force_clover against the same calls through
    # the Python binding, at the same (deterministic) FP64 CG solution
    fr = np.fromfile(str(out) + ".force", dtype=np.float64).reshape(2, V, 4, 3, 3, 2)
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    Ud = torch.from_numpy(U).to(DEV)
    A = qamd.clover_from_links(geom, Ud, 1.0 / (2 * 0.135), 1.0)
    op = qamd.EvenOddCloverOperator(geom, Ud, A, prec_sloppy=qamd.F16, recon_sloppy=12)
    x = geom.new_spinor(qamd.F64, cb=1)
    b = geom.pack_spinor(torch.from_numpy(psi).to(DEV), qamd.F64, 1)
    qamd.InvCG(op, max_iters=2000)(x, b, 1e-10, mixed=False)
    for k, csw in ((0, None), (1, 1.0)):
        F = op.force(torch.zeros_like(Ud), x, Ud, scale=0.5, project=True, csw=csw).cpu().numpy()
        err = np.abs(fr[k] - F).max() / np.abs(F).max()
        assert err < 1e-13, (k, err)
    op.close()


def test_cpp_face_wrappers_split_t(qamd, orc, tmp_path):
    """The C++ face wrappers (Geometry::face_pack, Dslash / ClovDslash /
    TMDslash / TMClovDslash::face_finish_dslash and face_finish_achimbdpsi)
    on a 16x8x8x8 lattice split in two along t (host_api_check split): each
    half runs the interior operator, packs its t faces, and completes its
    boundary sites from the other half's faces; the halves assembled must
    equal the whole-lattice FP64 oracle (1e-13), for every family, dslash and
    achimbdpsi (isign +1, output checkerboard 0, antiperiodic time).  The host
    program then repeats the split with the spinors in QPhiX FourSpinorBlock
    arrays (classes in their in-place mode, Geometry::face_pack from the
    arrays, finishes into them) and requires the same bits."""
    import qphix_blocks as qb
    assert os.path.exists(EXE), "tests/cpp/host_api_check is not built (build() builds it)"
    G = (16, 8, 8, 8)
    Lloc = (16, 8, 8, 4)
    L = qamd.lattice
    U = L.random_gauge(G, 0.3, 3200).numpy()
    psi = L.random_spinor(G, 3201).numpy()
    chi = L.random_spinor(G, 3202).numpy()
    clov_t = L.random_clover(G, 3203)
    clov, inv = clov_t.numpy(), L.invert_clover(clov_t).numpy()
    B = L.tmclover_blocks(clov_t, TMU, 1, True).numpy()
    F = L.tmclover_blocks(clov_t, TMU, 1, False).numpy()
    lay = qb.Layout(0, 1, 1, compress12=False)
    halves = []
    for k in range(2):
        origin = (0, 0, 0, k * Lloc[3])
        t, z, y, x = np.meshgrid(*(np.arange(n) for n in Lloc[::-1]), indexing="ij")
        gx, gy, gz, gt = (c.ravel() + o for c, o in zip((x, y, z, t), origin))
        gi = ((gt * G[2] + gz) * G[1] + gy) * G[0] + gx
        garr = qb.pack_gauge(U, lay, Lloc, 0, origin=origin, gdims=G).astype(np.float64).ravel()
        halves.append((gi, garr))
    inp = tmp_path / "split.bin"
    with open(inp, "wb") as f:
        np.asarray(Lloc, dtype=np.int32).tofile(f)
        np.asarray([halves[0][1].size], dtype=np.int32).tofile(f)
        for gi, garr in halves:
            garr.tofile(f)
            for a in (psi, chi, inv, clov, B, F):
                np.ascontiguousarray(a[gi], dtype=np.float64).tofile(f)
    out = tmp_path / "split_out.bin"
    print("running host_api_check split", G, "in two along t", flush=True)
    r = subprocess.run([EXE, "split", str(inp), str(out)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    Vl = int(np.prod(Lloc))
    res = np.fromfile(out, dtype=np.float64).reshape(4, 2, 2, Vl, 4, 3, 2)
    m = cb_mask(G, 0)
    ml = cb_mask(Lloc, 0)  # the halves start at even t, so local parity = global parity
    for kind in range(4):
        for achi in (0, 1):
            c = BETA if achi else 1.0
            kw = dict(kind=kind, achi=bool(achi), coeff_s=c * ANISO_S, coeff_t_f=c * ANISO_T,
                      coeff_t_b=c * ANISO_T, t_boundary=-1.0, alpha=ALPHA, rho=RHO, chi=chi)
            if kind == 1:
                kw["clov"] = clov if achi else inv
            elif kind == 2:
                a_op = 4.0 + MASS
                kw["mu"] = TMU / ALPHA if achi else TMU / a_op
                kw["muinv"] = a_op / (a_op ** 2 + TMU ** 2)
            elif kind == 3:
                kw["clov"] = F if achi else B
            want = orc.dslash(G, psi, U, 1, 0, **kw)
            got = np.zeros_like(want)
            for k, (gi, _) in enumerate(halves):
                got[gi[ml]] = res[kind, achi, k][ml]
            err = np.abs(got[m] - want[m]).max() / np.abs(want[m]).max()
            print(f"  split t: kind {kind} achi {achi}: {err:.3g}", flush=True)
            assert err < 1e-13, (kind, achi, err)
