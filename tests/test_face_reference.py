"""The reference's face bodies (oracle/_ref: dslash_face_pack_to_*,
*dslash_face_unpack_from_*, codegen/dslash.cc:433-553,664-744) against the
FP64 oracle, on the CPU.

A lattice cut along one face: the reference's body kernel with the hop
across the cut switched off at the face sites (its accumulate[] masks), the
neighbour's face packed by the reference's face pack body and completed by
its face finish body, must give the oracle's operator on the whole lattice.
This pins what the face bodies compute and the conventions a host passes to
them (and to the library's face entry points, which follow them):
  * the pack of face 2 mu + fb holds (1 -+ isign gamma_mu) psi of the sites
    at x_mu = 0 (fb 0) or L_mu - 1 (fb 1), the projector of the receiving
    side's hop;
  * the finish of face 2 mu + fb completes the sites at x_mu = 0 (fb 0) or
    L_mu - 1 (fb 1) through U_mu(x - mu)^dagger or U_mu(x);
  * beta: for the dslash, -coeff for the Wilson and twisted-mass finish (they
    subtract beta times the hop), +coeff for the clover and twisted-mass
    clover finish (A^-1 or B applied to beta times the hop); for the
    achimbdpsi operators the Wilson finish with +coeff; coeff the hop's
    coefficient with the time boundary factor on the first / last slice.
Every family, both isigns and checkerboards, every face, 12- and 18-real
links, the xy-packed build (SOALEN < VECLEN: lane-masked x and y faces) and
twisted boundary phases.  Skipped where the reference build or AVX-512 is
absent.
"""
import numpy as np
import pytest

from _util import cb_mask, parity, coords

D = (16, 4, 4, 4)
CS, CTF, CTB, TB = 0.9, 1.15, 1.05, -1.0
ALPHA, MU_PHYS, RHO = 4.2, 0.13, 0.07
MU, MUINV = MU_PHYS / ALPHA, ALPHA / (ALPHA ** 2 + MU_PHYS ** 2)
TBC_PHASES = [0.75 + 0.5j, 0.5 - 0.875j, -0.625 + 0.75j, 0.875 + 0.25j]


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


@pytest.fixture(scope="module")
def lat(qamd):
    L = qamd.lattice
    clov = L.random_clover(D, 2103)
    return dict(U=L.random_gauge(D, 0.3, 2100).numpy(), psi=L.random_spinor(D, 2101).numpy(),
                chi=L.random_spinor(D, 2102).numpy(), clov=clov.numpy(), inv=L.invert_clover(clov).numpy(),
                B={s: L.tmclover_blocks(clov, MU_PHYS, s, True).numpy() for s in (1, -1)},
                F={s: L.tmclover_blocks(clov, MU_PHYS, s, False).numpy() for s in (1, -1)})


def _epi(lat, kind, achi, isign):
    if kind == 1:
        return lat["clov"] if achi else lat["inv"]
    if kind == 3:
        return (lat["F"] if achi else lat["B"])[isign]
    return None


def _coeff(mu, fb):
    if mu < 3:
        return CS
    return CTB * TB if fb == 0 else CTF * TB


def face_beta(kind, achi, mu, fb):
    """beta of the face finish completing the dslash (achi 0) or the
    achimbdpsi operator (achi 1; always the Wilson finish)."""
    c = _coeff(mu, fb)
    if achi:
        return c
    return c if kind in (1, 3) else -c


def split_case(ref, config, lat, kind, achi, isign, cb, mu, fb, recon, tbc):
    """interior (reference body, hop 2 mu + fb off at the face sites) + the
    reference's face pack / finish of that hop."""
    ep = _epi(lat, kind, achi, isign)
    kw = dict(kind=kind, achi=achi, chi=lat["chi"], clov=ep, recon=recon, coeff_s=CS, coeff_t_f=CTF,
              coeff_t_b=CTB, t_boundary=TB, alpha=ALPHA, mu=MU, muinv=MUINV, rho=RHO,
              tbc_phases=TBC_PHASES if tbc else None)
    c = coords(D)
    on = (parity(D) == cb) & (c[:, mu] == (0 if fb == 0 else D[mu] - 1))
    mask = np.full(c.shape[0], 255, np.uint8)
    mask[on] &= np.uint8(~(1 << (2 * mu + fb)) & 255)
    interior, _ = ref.apply(config, D, lat["psi"], lat["U"], isign, cb, mask=mask, **kw)
    # the neighbour across the cut packs its opposite face (checkerboard 1 - cb)
    packed = ref.face_pack(config, D, lat["psi"], isign, 2 * mu + (1 - fb), 1 - cb)
    src = c.copy()
    src[:, mu] = (src[:, mu] + (-1 if fb == 0 else 1)) % D[mu]
    lexsrc = ((src[:, 3] * D[2] + src[:, 2]) * D[1] + src[:, 1]) * D[0] + src[:, 0]
    hs = np.zeros((c.shape[0], 12))
    hs[on] = packed[lexsrc[on]]
    fam = 0 if achi else kind
    fclov = None if achi else ep
    return ref.face_finish(config, D, fam, isign, recon, 2 * mu + fb, cb, hs, lat["U"], interior,
                           face_beta(kind, achi, mu, fb), clov=fclov, mu=MU, muinv=MUINV,
                           tbc_phases=TBC_PHASES if tbc else None)


@pytest.mark.parametrize("config", ["avx512_f64", "avx512_f64_s4"])
@pytest.mark.parametrize("tbc", [False, True])
def test_reference_faces_complete_the_operator(qamd, orc, lat, config, tbc):
    ref = _ref(config)
    worst = 0.0
    for kind in range(4):
        for achi in (0, 1):
            for isign, cb in ((1, 0), (-1, 1)):
                ep = _epi(lat, kind, achi, isign)
                want = orc.dslash(D, lat["psi"], lat["U"], isign, cb, kind=kind, achi=achi, chi=lat["chi"],
                                  clov=ep, coeff_s=CS, coeff_t_f=CTF, coeff_t_b=CTB, t_boundary=TB,
                                  tbc=[1, 1, 1, 1] if tbc else None, tbc_phases=TBC_PHASES if tbc else None,
                                  alpha=ALPHA, mu=MU, muinv=MUINV, rho=RHO)
                m = cb_mask(D, cb)
                for mu in range(4):
                    for fb in (0, 1):
                        if achi and (mu + fb) % 2:  # achimbdpsi: every other face suffices
                            continue
                        recon = 12 if (mu + fb + kind) % 2 == 0 else 18
                        got = split_case(ref, config, lat, kind, achi, isign, cb, mu, fb, recon, tbc)
                        err = np.abs(got[m] - want[m]).max() / np.abs(want[m]).max()
                        assert err < 1e-13, (config, tbc, kind, achi, isign, cb, mu, fb, recon, err)
                        worst = max(worst, err)
    print(f"{config} tbc={tbc}: worst relative difference {worst:.3g}")


def test_reference_face_pack_is_the_projection(qamd, orc, lat):
    """The packed half-spinors are the upper two spin components of
    (1 + s gamma_mu) psi, s = -isign for the back face and +isign for the
    forward face (DeGrand-Rossi gamma matrices, not the reference's tables)."""
    from _util import GAMMA, cplx
    ref = _ref("avx512_f64_s4")
    psi = cplx(lat["psi"])  # (V, 4, 3)
    c = coords(D)
    for isign in (1, -1):
        for pcb in (0, 1):
            for mu in range(4):
                for fb in (0, 1):
                    hs = ref.face_pack("avx512_f64_s4", D, lat["psi"], isign, 2 * mu + fb, pcb)
                    on = (parity(D) == pcb) & (c[:, mu] == (0 if fb == 0 else D[mu] - 1))
                    s = -isign if fb == 0 else isign
                    proj = np.einsum("ab,vbc->vac", np.eye(4) + s * GAMMA[mu], psi[on])[:, :2]
                    want = np.stack([proj.real, proj.imag], axis=-1).reshape(-1, 12)
                    assert np.abs(hs[on] - want).max() < 1e-15, (isign, pcb, mu, fb)
                    assert not hs[~on].any()
