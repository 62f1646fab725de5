"""CPU checks of the QPhiX block-layout restatement (tests/qphix_blocks.py)
and of the converter entry points' host-side validation.

* The restated block arrays equal, element for element, the arrays the
  reference driver hands to the reference's own kernels (oracle/_ref,
  ref.pack), for spinor, gauge (12 and 18 rows), clover and full clover on
  both checkerboards, in every built configuration: soalen == veclen, and
  the xy-packed soalen = veclen / 2 builds (one gauge vector over two rows
  of y).  Those
  kernels reproduce the FP64 oracle to rounding (tests/test_golden.py), so
  this pins the layout the converters must read.
* qphix_amd_qphix_field_bytes matches the restatement's block counts and
  rejects layouts that do not tile the lattice (no GPU work is launched).
"""
import ctypes

import numpy as np
import pytest

import qphix_blocks as qb

CONFIGS = ("scalar_f64", "avx2_f64", "avx2_f32", "avx512_f64", "avx512_f32", "avx512_f64_s4", "avx512_f32_s8")


def _ref():
    from oracle import ref
    return ref


@pytest.mark.parametrize("config", CONFIGS)
def test_restated_blocks_equal_reference_driver(qamd, config):
    ref = _ref()
    if not ref.available(config):
        pytest.skip(f"oracle/_ref/libref_{config}.so not built")
    S, V = ref.soalen(config), ref.veclen(config)
    ft = 1 if ref.lib(config).ref_precision() == 1 else 0
    dims = (2 * S * 2, 4, 2, 4)
    L = qamd.lattice
    U = L.random_gauge(dims, 0.4, 71).numpy()
    psi = L.random_spinor(dims, 72).numpy()
    clov = L.random_clover(dims, 73).numpy()
    fcl = L.tmclover_blocks(L.random_clover(dims, 73), 0.1, 1, False).numpy()
    for cb in (0, 1):
        for recon in (12, 18):
            lay = qb.Layout(ft, V, S, compress12=recon == 12)
            got = qb.pack_gauge(U, lay, dims, cb)
            np.testing.assert_array_equal(got, ref.pack(config, 1, dims, U, cb, recon))
        lay = qb.Layout(ft, V, S)
        np.testing.assert_array_equal(qb.pack_spinor(psi, lay, dims, cb), ref.pack(config, 0, dims, psi, cb))
        np.testing.assert_array_equal(qb.pack_clover(clov, lay, dims, cb), ref.pack(config, 2, dims, clov, cb))
        np.testing.assert_array_equal(qb.pack_full_clover(fcl, lay, dims, cb),
                                      ref.pack(config, 3, dims, fcl, cb))


def test_unpack_inverts_pack():
    dims = (16, 4, 2, 4)
    rng = np.random.default_rng(3)
    psi = rng.normal(size=(int(np.prod(dims)), 4, 3, 2))
    for lay in (qb.Layout(0, 8, 4, pad_xy=2, pad_xyz=6), qb.Layout(1, 4, 4, pad_xy=1, pad_xyz=3)):
        for cb in (0, 1):
            back = qb.unpack_spinor(qb.pack_spinor(psi, lay, dims, cb, fill=np.nan), lay, dims, cb)
            m = qb.cb_sites(dims, cb)[4]
            np.testing.assert_allclose(back[m], psi[m], rtol=1e-7 if lay.ft else 0)


def test_qphix_field_bytes_and_validation(qamd):
    lib = qamd.lib()
    ctx = ctypes.c_void_p()
    assert lib.qphix_amd_ctx_create(ctypes.byref(ctx), 16, 4, 2, 4, -1.0) == 0
    try:
        dims = (16, 4, 2, 4)
        for lay in (qb.Layout(0, 8, 8), qb.Layout(1, 16, 4, True, 4, 8), qb.Layout(2, 8, 2, False, 4, 4)):
            c = lay.ctype(qamd)
            s = {0: 8, 1: 4, 2: 2}[lay.ft]
            nrow = 2 if lay.compress12 else 3
            assert lib.qphix_amd_qphix_field_bytes(ctx, ctypes.byref(c), 0) == \
                lay.n_spinor_blocks(dims) * 24 * lay.soalen * s
            assert lib.qphix_amd_qphix_field_bytes(ctx, ctypes.byref(c), 1) == \
                lay.n_matrix_blocks(dims) * 8 * nrow * 6 * lay.veclen * s
            assert lib.qphix_amd_qphix_field_bytes(ctx, ctypes.byref(c), 2) == \
                lay.n_matrix_blocks(dims) * 72 * lay.veclen * s
            assert lib.qphix_amd_qphix_field_bytes(ctx, ctypes.byref(c), 3) == \
                lay.n_matrix_blocks(dims) * 144 * lay.veclen * s
        bad = [qamd.qphix_layout(0, 8, 3),          # soalen does not divide veclen
               qamd.qphix_layout(0, 32, 16),        # soalen 16 does not divide Lx/2 = 8
               qamd.qphix_layout(0, 32, 4),         # ngy = 8 does not divide Ly = 4
               qamd.qphix_layout(0, 8, 4, True, 1, 0),   # pad_xy not a multiple of ngy
               qamd.qphix_layout(3, 8, 8),          # BF16 is not a QPhiX storage type
               qamd.qphix_layout(1, 8, 8, True, -2, 0)]
        buf = ctypes.create_string_buffer(256)
        p = ctypes.c_void_p((ctypes.addressof(buf) + 15) & ~15)
        for c in bad:
            assert lib.qphix_amd_qphix_field_bytes(ctx, ctypes.byref(c), 0) == 0
            assert lib.qphix_amd_spinor_from_qphix(ctx, ctypes.byref(c), p, 1, p, 0, None) == -1
        good = qamd.qphix_layout(1, 8, 8)
        assert lib.qphix_amd_gauge_from_qphix(ctx, ctypes.byref(good), p, 1, 13, p, 0, None) == -1
        assert lib.qphix_amd_spinor_to_qphix(ctx, ctypes.byref(good), p, p, 1, 2, None) == -1
        assert lib.qphix_amd_clover_from_qphix(ctx, None, p, 1, p, 0, None) == -1
        assert lib.qphix_amd_full_clover_from_qphix(ctx, ctypes.byref(good), None, 1, p, 0, None) == -1
        odd = ctypes.c_void_p(p.value + 4)
        assert lib.qphix_amd_spinor_from_qphix(ctx, ctypes.byref(good), odd, 1, p, 0, None) == -6
    finally:
        lib.qphix_amd_ctx_destroy(ctx)
