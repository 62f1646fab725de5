"""Numpy restatement of the QPhiX block arrays a host hands to the
reference's kernels (test infrastructure; the library converts them on the
device with qphix_amd_*_from_qphix / qphix_amd_spinor_to_qphix).

Block types (codegen/include/decl_common.h:9-62), per checkerboard:
  FourSpinorBlock  FT[3 colour][4 spin][2][soalen]
  SU3MatrixBlock   FT[8][nrow][3][2][veclen], entry [d][i][j] = M_d[j][i],
                   M_{2mu} = U_mu(x - mu), M_{2mu+1} = U_mu(x)
                   (codegen/dslash_common.cc:1678-1727), nrow 2 with compress12
  CloverBlock      {diag1[6][V], off_diag1[15][2][V], diag2, off_diag2}
  FullCloverBlock  {block1[6][6][2][V], block2}
Block order of QPhiX's checkerboarded geometry (spinor blocks
t Pxyz + z Pxy + y nvecs + xh/soalen; gauge/clover blocks
(t Pxyz + z Pxy)/ngy + (y/ngy) nvecs + xh/soalen, lane (y%ngy) soalen + xh%soalen).
At soalen == veclen without padding this is the block order of the
reference driver oracle/ref_driver.cpp (checked in tests/test_qphix_blocks.py).
"""
import numpy as np

NP_FT = {0: np.float64, 1: np.float32, 2: np.float16}


class Layout:
    def __init__(self, ft, veclen, soalen, compress12=True, pad_xy=0, pad_xyz=0):
        self.ft, self.veclen, self.soalen = ft, veclen, soalen
        self.compress12, self.pad_xy, self.pad_xyz = bool(compress12), pad_xy, pad_xyz

    def ctype(self, qamd):
        return qamd.qphix_layout(self.ft, self.veclen, self.soalen, self.compress12, self.pad_xy,
                                 self.pad_xyz)

    def geometry(self, dims):
        Lx, Ly, Lz, Lt = dims
        nvecs = (Lx // 2) // self.soalen
        ngy = self.veclen // self.soalen
        pxy = Ly * nvecs + self.pad_xy
        pxyz = Lz * pxy + self.pad_xyz
        return nvecs, ngy, pxy, pxyz

    def n_spinor_blocks(self, dims):
        return dims[3] * self.geometry(dims)[3]

    def n_matrix_blocks(self, dims):
        return self.n_spinor_blocks(dims) // self.geometry(dims)[1]


def cb_sites(dims, cb):
    """(xh, y, z, t, lex) of the sites of checkerboard cb."""
    Lx, Ly, Lz, Lt = dims
    t, z, y, xh = np.meshgrid(np.arange(Lt), np.arange(Lz), np.arange(Ly), np.arange(Lx // 2),
                              indexing="ij")
    t, z, y, xh = t.ravel(), z.ravel(), y.ravel(), xh.ravel()
    x = 2 * xh + ((cb + y + z + t) & 1)
    lex = ((t * Lz + z) * Ly + y) * Lx + x
    return xh, y, z, t, lex, x


def spinor_slots(lay, dims, cb):
    nvecs, ngy, pxy, pxyz = lay.geometry(dims)
    xh, y, z, t, lex, _ = cb_sites(dims, cb)
    return t * pxyz + z * pxy + y * nvecs + xh // lay.soalen, xh % lay.soalen, lex


def matrix_slots(lay, dims, cb):
    nvecs, ngy, pxy, pxyz = lay.geometry(dims)
    xh, y, z, t, lex, _ = cb_sites(dims, cb)
    blk = (t * pxyz + z * pxy) // ngy + (y // ngy) * nvecs + xh // lay.soalen
    return blk, (y % ngy) * lay.soalen + xh % lay.soalen, lex


def pack_spinor(lex_field, lay, dims, cb, fill=0.0):
    blk, lane, lex = spinor_slots(lay, dims, cb)
    arr = np.full((lay.n_spinor_blocks(dims), 3, 4, 2, lay.soalen), fill, dtype=NP_FT[lay.ft])
    # lex [V][spin][colour][2] -> [colour][spin][2]
    arr[blk, :, :, :, lane] = np.transpose(lex_field[lex], (0, 2, 1, 3))
    return arr.ravel()


def unpack_spinor(arr, lay, dims, cb):
    blk, lane, lex = spinor_slots(lay, dims, cb)
    a = np.asarray(arr).reshape(lay.n_spinor_blocks(dims), 3, 4, 2, lay.soalen)
    out = np.zeros((int(np.prod(dims)), 4, 3, 2))
    out[lex] = np.transpose(a[blk, :, :, :, lane].astype(np.float64), (0, 2, 1, 3))
    return out


def pack_gauge(U, lay, dims, cb, origin=(0, 0, 0, 0), gdims=None):
    """U: lex (V,4,3,3,2) FP64, U_mu(x) row-major, on the lattice gdims
    (default dims).  The blocks cover the sub-lattice dims at `origin` (even
    coordinate sum); back links of its boundary sites come from across the
    cut, as a host with a gauge halo provides them."""
    gdims = tuple(dims) if gdims is None else tuple(gdims)
    GX, GY, GZ, GT = gdims
    blk, lane, lex = matrix_slots(lay, dims, cb)
    _, y, z, t, _, x = cb_sites(dims, cb)
    nrow = 2 if lay.compress12 else 3
    arr = np.zeros((lay.n_matrix_blocks(dims), 8, nrow, 3, 2, lay.veclen), dtype=NP_FT[lay.ft])
    c = np.stack([x, y, z, t], axis=1) + np.asarray(origin)[None, :]
    G = np.array(gdims)
    for d in range(8):
        mu, fb = d // 2, d % 2
        n = c.copy()
        if not fb:
            n[:, mu] = (n[:, mu] - 1) % G[mu]
        s = ((n[:, 3] * GZ + n[:, 2]) * GY + n[:, 1]) * GX + n[:, 0]
        M = U[s, mu]  # [site][row][col][2]
        ST = np.transpose(M, (0, 2, 1, 3))  # S[i][j] = M[j][i]
        arr[blk, d, :, :, :, lane] = ST[:, :nrow]
    return arr.ravel()


def pack_clover(clov, lay, dims, cb):
    """clov: lex (V,2,36) packed Hermitian blocks -> CloverBlock array."""
    blk, lane, lex = matrix_slots(lay, dims, cb)
    arr = np.zeros((lay.n_matrix_blocks(dims), 72, lay.veclen), dtype=NP_FT[lay.ft])
    arr[blk, :, lane] = clov.reshape(-1, 72)[lex]
    return arr.ravel()


def pack_full_clover(fclov, lay, dims, cb):
    """fclov: lex (V,2,6,6,2) full blocks -> FullCloverBlock array."""
    blk, lane, lex = matrix_slots(lay, dims, cb)
    arr = np.zeros((lay.n_matrix_blocks(dims), 144, lay.veclen), dtype=NP_FT[lay.ft])
    arr[blk, :, lane] = fclov.reshape(-1, 144)[lex]
    return arr.ravel()
