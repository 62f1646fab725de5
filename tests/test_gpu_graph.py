"""Operator calls captured in a HIP graph and replayed.

The operator, in-place QPhiX, converter and scratch-free force entry points
only enqueue kernels on the stream they are given (no allocation, no
synchronisation, no memset), so a host can capture them once and replay
them.  Each test captures a sequence with torch.cuda.graph (the calls pass
torch's current stream, the capture stream), then replays it three times;
before every replay new inputs are copied into the captured buffers with
eager work and a device synchronisation, and after it the outputs must equal,
bit for bit, the same calls made eagerly on the same inputs.
"""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
DIMS = (8, 8, 8, 16)
REPLAYS = 3


def _capture(fn):
    """Warm fn up on a side stream, then capture it into a graph."""
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        fn()
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn()
    torch.cuda.synchronize()
    return g


def _replay_vs_eager(g, fn, load_inputs, outputs):
    for k in range(REPLAYS):
        print(f"replay {k}", flush=True)
        load_inputs(k)
        for o in outputs:
            o.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = [o.clone() for o in outputs]
        for o in outputs:
            o.fill_(float("nan"))
        fn()
        torch.cuda.synchronize()
        for a, b in zip(got, outputs):
            assert not torch.isnan(b.float()).any()
            assert torch.equal(a, b), f"replay {k} differs from the eager call"


@pytest.mark.parametrize("prec_name", ["F32", "F16", "BF16"])
def test_clover_operator_pair_in_graph(qamd, prec_name):
    """clov_dslash on cb 0 then clov_dslash_achimbdpsi on cb 1 (the even-odd
    clover operator's two kernels)."""
    print(f"case graph clover pair {prec_name} {DIMS}", flush=True)
    prec = getattr(qamd, prec_name)
    L = qamd.lattice
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    U = L.random_gauge(DIMS, 0.3, 700, device=DEV)
    clov = L.random_clover(DIMS, 701, device=DEV)
    g0, g1 = geom.pack_gauge(U, prec, 12, 0), geom.pack_gauge(U, prec, 12, 1)
    inv0 = geom.pack_clover(L.invert_clover(clov), prec, 0)
    clov1 = geom.pack_clover(clov, prec, 1)
    psi = geom.pack_spinor(L.random_spinor(DIMS, 702, device=DEV), prec, 1)
    t0, out = geom.new_spinor(prec, cb=0), geom.new_spinor(prec, cb=1)
    op = qamd.ClovDslash(geom)

    def fn():
        op.dslash(t0, psi, g0, inv0, 1, 0)
        op.dslashAChiMinusBDPsi(out, t0, psi, g1, clov1, 0.25, 1, 1)

    def load(k):
        psi.data.copy_(geom.pack_spinor(L.random_spinor(DIMS, 710 + k, device=DEV), prec, 1).data)

    _replay_vs_eager(_capture(fn), fn, load, [t0.data, out.data])


def test_tm_clover_and_wilson_in_graph(qamd):
    """FP16 twisted-mass clover dslash / achimbdpsi with both isigns and a
    Wilson dslash, in one graph."""
    print(f"case graph tm-clover + wilson F16 {DIMS}", flush=True)
    prec = qamd.F16
    L = qamd.lattice
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    U = L.random_gauge(DIMS, 0.3, 720, device=DEV)
    clov = L.random_clover(DIMS, 721, device=DEV)
    g0, g1 = geom.pack_gauge(U, prec, 12, 0), geom.pack_gauge(U, prec, 12, 1)
    binv = {s: geom.pack_full_clover(L.tmclover_blocks(clov, 0.05, s, True), prec, 0) for s in (1, -1)}
    fwd = {s: geom.pack_full_clover(L.tmclover_blocks(clov, 0.05, s, False), prec, 1) for s in (1, -1)}
    psi = geom.pack_spinor(L.random_spinor(DIMS, 722, device=DEV), prec, 1)
    chi = geom.pack_spinor(L.random_spinor(DIMS, 723, device=DEV), prec, 1)
    outs = {s: (geom.new_spinor(prec, cb=0), geom.new_spinor(prec, cb=1)) for s in (1, -1)}
    w = geom.new_spinor(prec, cb=0)
    tm = qamd.TMClovDslash(geom)
    wil = qamd.Dslash(geom)

    def fn():
        for s in (1, -1):
            tm.dslash(outs[s][0], psi, g0, binv[s], s, 0)
            tm.dslashAChiMinusBDPsi(outs[s][1], outs[s][0], chi, g1, fwd[s], 0.25, s, 1)
        wil.dslash(w, psi, g0, 1, 0)

    def load(k):
        psi.data.copy_(geom.pack_spinor(L.random_spinor(DIMS, 730 + k, device=DEV), prec, 1).data)
        chi.data.copy_(geom.pack_spinor(L.random_spinor(DIMS, 740 + k, device=DEV), prec, 1).data)

    targets = [o.data for s in (1, -1) for o in outs[s]] + [w.data]
    _replay_vs_eager(_capture(fn), fn, load, targets)


def test_in_place_qphix_arrays_in_graph(qamd):
    """Conversion of a native spinor into a QPhiX array, then the clover
    dslash on QPhiX arrays in place (FP32, veclen 16, soalen 8)."""
    dims = (16, 8, 8, 8)  # Lx / 2 a multiple of soalen 8, Ly of veclen / soalen
    print(f"case graph in-place QPhiX F32 {dims}", flush=True)
    prec = qamd.F32
    L = qamd.lattice
    geom = qamd.Geometry(dims, t_boundary=-1.0)
    U = L.random_gauge(dims, 0.3, 750, device=DEV)
    clov = L.random_clover(dims, 751, device=DEV)
    g0 = geom.pack_gauge(U, prec, 12, 0)
    inv0 = geom.pack_clover(L.invert_clover(clov), prec, 0)
    lay = qamd.qphix_layout(prec, 16, 8, True)
    psi = geom.pack_spinor(L.random_spinor(dims, 752, device=DEV), prec, 1)
    psi_q = qamd.QPhiXSpinor(geom.spinor_to_qphix(psi, lay), lay)
    res_q = qamd.QPhiXSpinor(torch.empty_like(psi_q.data), lay)
    op = qamd.ClovDslash(geom)

    def fn():
        geom.spinor_to_qphix(psi, lay, out=psi_q.data)
        op.dslash(res_q, psi_q, g0, inv0, 1, 0)

    def load(k):
        psi.data.copy_(geom.pack_spinor(L.random_spinor(dims, 760 + k, device=DEV), prec, 1).data)

    _replay_vs_eager(_capture(fn), fn, load, [psi_q.data, res_q.data])


def test_split_lattice_faces_in_graph(qamd):
    """One part of a lattice split in t (FP32): the interior clover dslash,
    then the packing of a face and the completion of the other from a
    buffer, in one graph."""
    print(f"case graph faces F32 {DIMS} split t", flush=True)
    prec = qamd.F32
    L = qamd.lattice
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    geom.set_comm([0, 0, 0, 1])
    U = L.random_gauge(DIMS, 0.3, 800, device=DEV)
    clov = L.random_clover(DIMS, 801, device=DEV)
    g0 = geom.pack_gauge(U, prec, 12, 0)
    inv0 = geom.pack_clover(L.invert_clover(clov), prec, 0)
    psi = geom.pack_spinor(L.random_spinor(DIMS, 802, device=DEV), prec, 1)
    res = geom.new_spinor(prec, cb=0)
    send = geom.new_face(prec, 3)
    recv = geom.new_face(prec, 3)
    op = qamd.ClovDslash(geom)

    def fn():
        op.dslash(res, psi, g0, inv0, 1, 0)
        qamd.face_pack(geom, send, psi, 1, 0, 3, 1)
        qamd.face_finish(geom, 1, res, recv, g0, 1, 0, 3, 0, 1.0, clov=inv0)

    def load(k):
        psi.data.copy_(geom.pack_spinor(L.random_spinor(DIMS, 810 + k, device=DEV), prec, 1).data)
        recv.copy_(torch.randn(recv.shape, dtype=recv.dtype, device=DEV, generator=None))

    _replay_vs_eager(_capture(fn), fn, load, [res.data, send])


def test_noalloc_clover_force_in_graph(qamd):
    """qphix_amd_clover_force_noalloc (workspace in the library's static
    array) accumulating into a zeroed force field."""
    print(f"case graph noalloc clover force {DIMS}", flush=True)
    lib = qamd.lib()
    L = qamd.lattice
    geom = qamd.Geometry(DIMS, t_boundary=-1.0)
    U = L.random_gauge(DIMS, 0.3, 770, device=DEV)
    chi = L.random_spinor(DIMS, 771, device=DEV)
    psi = L.random_spinor(DIMS, 772, device=DEV)
    force = torch.zeros_like(U)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731

    def fn():
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert lib.qphix_amd_clover_force_noalloc(geom.ctx, p(force), p(chi), p(psi), p(U), 1.5, -0.5, -2.0, 1,
                                                  st) == 0

    def load(k):
        chi.copy_(L.random_spinor(DIMS, 780 + k, device=DEV))
        psi.copy_(L.random_spinor(DIMS, 790 + k, device=DEV))

    g = _capture(fn)
    for k in range(REPLAYS):
        print(f"replay {k}", flush=True)
        load(k)
        force.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = force.clone()
        force.zero_()
        fn()
        torch.cuda.synchronize()
        assert torch.count_nonzero(got) > 0
        assert torch.equal(got, force), f"replay {k} differs from the eager call"
