"""Shared helpers for the tests: DeGrand-Rossi gamma matrices, checkerboard
masks and complex views of the lexicographic FP64 arrays."""
import numpy as np

# DeGrand-Rossi basis (the basis the reference's projector tables encode).
GAMMA = np.zeros((4, 4, 4), dtype=complex)
GAMMA[0][0, 3], GAMMA[0][1, 2], GAMMA[0][2, 1], GAMMA[0][3, 0] = 1j, 1j, -1j, -1j
GAMMA[1][0, 3], GAMMA[1][1, 2], GAMMA[1][2, 1], GAMMA[1][3, 0] = -1, 1, 1, -1
GAMMA[2][0, 2], GAMMA[2][1, 3], GAMMA[2][2, 0], GAMMA[2][3, 1] = 1j, -1j, -1j, 1j
GAMMA[3][0, 2], GAMMA[3][1, 3], GAMMA[3][2, 0], GAMMA[3][3, 1] = 1, 1, 1, 1
GAMMA5 = GAMMA[0] @ GAMMA[1] @ GAMMA[2] @ GAMMA[3]


def cplx(a):
    """(..., 2) float -> (...) complex."""
    a = np.asarray(a)
    return a[..., 0] + 1j * a[..., 1]


def real(z):
    return np.stack([z.real, z.imag], axis=-1)


def coords(dims):
    Lx, Ly, Lz, Lt = dims
    t, z, y, x = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx), indexing="ij")
    return np.stack([x.ravel(), y.ravel(), z.ravel(), t.ravel()], axis=1)


def parity(dims):
    return coords(dims).sum(axis=1) & 1


def cb_mask(dims, cb):
    return parity(dims) == cb


def g5(psi):
    """gamma_5 on a (V,4,3,2) real array."""
    out = psi.copy()
    out[:, 2:] *= -1
    return out


def inner(a, b):
    """<a, b> = sum conj(a) b over (V,4,3,2) real arrays."""
    return np.sum(np.conj(cplx(a)) * cplx(b))


def on_cb(a, dims, cb):
    out = np.zeros_like(a)
    m = cb_mask(dims, cb)
    out[m] = a[m]
    return out


def rel_err(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))


def f16_ulps(a, b):
    """Distance in FP16 units in the last place between two float16 arrays
    (sign-magnitude bit patterns mapped to a monotone integer line)."""
    def line(x):
        i = np.asarray(x, dtype=np.float16).view(np.int16).astype(np.int32)
        return np.where(i < 0, -(i & 0x7FFF), i)
    return np.abs(line(a) - line(b))


def recon12_links(U):
    """Links as 12-real kernels apply them: the reference keeps columns 0, 1
    of every link (rows of its transposed storage) and rebuilds column 2 =
    conj(col0 x col1) (codegen/dslash_common.cc:1664-1674)."""
    z = U[..., 0] + 1j * U[..., 1]
    c0, c1 = z[..., :, 0], z[..., :, 1]
    z = z.copy()
    z[..., :, 2] = np.conj(np.cross(c0, c1))
    return np.stack([z.real, z.imag], axis=-1)


def f16_within(got, want):
    """Elementwise agreement of two FP16 results of the same FP32-arithmetic
    operator: within one FP16 unit in the last place of `want`, or within
    2^-20 of the largest |want| where cancellation leaves a small value (two
    summation orders in FP32 differ by about 2^-24 of the summands there,
    more than one FP16 ulp of a near-zero result)."""
    w = np.asarray(want, dtype=np.float16)
    g = np.asarray(got, dtype=np.float16)
    diff = np.abs(g.astype(np.float64) - w.astype(np.float64))
    ulp = np.spacing(np.abs(w)).astype(np.float64)
    return diff <= np.maximum(ulp, 2.0 ** -20 * np.abs(w.astype(np.float64)).max())


def accumulate_mask(dims, cb, soalen, veclen, seed):
    """A per-site direction mask (one byte per lexicographic site, bit d =
    2 mu + fb) that the reference's bodies can express: random per site in
    x, and in y when soalen < veclen (the directions they mask lane by lane,
    codegen/dslash.cc:109,285-290,570-578); the same for every site of a
    body call's vector (ngy rows x soalen lanes) otherwise, since the bodies
    gate those directions per call with `if (accumulate[d])`."""
    Lx, Ly, Lz, Lt = dims
    ngy, nvecs = veclen // soalen, Lx // 2 // soalen
    rng = np.random.default_rng(seed)
    t, z, y, xh = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx // 2),
                              indexing="ij")
    t, z, y, xh = t.ravel(), z.ravel(), y.ravel(), xh.ravel()
    lex = ((t * Lz + z) * Ly + y) * Lx + 2 * xh + ((cb + y + z + t) & 1)
    vec = ((t * Lz + z) * (Ly // ngy) + y // ngy) * nvecs + xh // soalen
    nvec = vec.max() + 1
    mask = np.zeros(Lx * Ly * Lz * Lt, dtype=np.uint8)
    for d in range(8):
        per_site = d < 2 or (d < 4 and soalen < veclen)
        bits = rng.integers(0, 2, len(lex)) if per_site else rng.integers(0, 2, nvec)[vec]
        mask[lex] |= (bits << d).astype(np.uint8)
    return mask


def masked_oracle(orc, dims, psi, U, isign, cb, mask, **kw):
    """The operator with direction d dropped at the sites whose mask bit d is
    clear: the oracle's result split exactly into its 8 direction terms
    (links of one dimension kept, the others zeroed; forward and backward
    hops separated by a twisted phase i in that dimension, forward e^{i pi/2},
    backward its conjugate), then summed under the mask."""
    c = lambda a: a[..., 0] + 1j * a[..., 1]
    base = c(orc.dslash(dims, psi, np.zeros_like(U), isign, cb, **kw))  # the a chi part
    total = base.copy()
    bits = (mask.astype(np.int64)[:, None] >> np.arange(8)[None, :]) & 1  # (V, 8)
    for mu in range(4):
        Um = np.zeros_like(U)
        Um[:, mu] = U[:, mu]
        tbc = [int(m == mu) for m in range(4)]
        h0 = c(orc.dslash(dims, psi, Um, isign, cb, **kw)) - base
        h1 = c(orc.dslash(dims, psi, Um, isign, cb, tbc=tbc, tbc_phases=[1j if m == mu else 1 for m in range(4)],
                          **kw)) - base
        fwd, bwd = (h0 - 1j * h1) / 2, (h0 + 1j * h1) / 2
        total += bits[:, 2 * mu, None, None] * bwd + bits[:, 2 * mu + 1, None, None] * fwd
    return np.stack([total.real, total.imag], axis=-1)


def face_sites(dims, mu, side, cb):
    """Lexicographic sites of checkerboard cb on the slice x_mu = 0 (side 0)
    or x_mu = L_mu - 1 (side 1), in the order of the face index f of the
    library's face buffers: the slice's other coordinates, the first of them
    halved on the checkerboard (x, or y when mu = x), fastest first."""
    Lx, Ly, Lz, Lt = dims
    if mu == 0:
        t, z, yh = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly // 2), indexing="ij")
        t, z, yh = t.ravel(), z.ravel(), yh.ravel()
        x = np.full_like(t, Lx - 1 if side else 0)
        y = 2 * yh + ((cb + x + z + t) & 1)
    else:
        if mu == 1:
            t, z, xh = (a.ravel() for a in np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Lx // 2),
                                                       indexing="ij"))
            y = np.full_like(t, Ly - 1 if side else 0)
        elif mu == 2:
            t, y, xh = (a.ravel() for a in np.meshgrid(np.arange(Lt), np.arange(Ly), np.arange(Lx // 2),
                                                       indexing="ij"))
            z = np.full_like(t, Lz - 1 if side else 0)
        else:
            z, y, xh = (a.ravel() for a in np.meshgrid(np.arange(Lz), np.arange(Ly), np.arange(Lx // 2),
                                                       indexing="ij"))
            t = np.full_like(z, Lt - 1 if side else 0)
        x = 2 * xh + ((cb + y + z + t) & 1)
    return ((t * Lz + z) * Ly + y) * Lx + x


def face_width(itemsize):
    """Reals per vector unit of a face buffer (12 reals per face site)."""
    w = 16 // itemsize
    return w if 12 % w == 0 else 8 // itemsize


def face_from_buffer(buf, nface, itemsize):
    """Face buffer (flat: unit-major, then face site, then the unit's reals)
    -> (nface, 12) in [spin 0|1][colour][re|im] order."""
    w = face_width(itemsize)
    return np.asarray(buf).reshape(12 // w, nface, w).transpose(1, 0, 2).reshape(nface, 12)


def face_to_buffer(hs, itemsize):
    """(nface, 12) -> the flat face buffer layout."""
    w = face_width(itemsize)
    n = hs.shape[0]
    return np.ascontiguousarray(np.asarray(hs).reshape(n, 12 // w, w).transpose(1, 0, 2)).ravel()
