"""CPU checks that the Python host layer rejects bad arguments before any
pointer reaches the C ABI (which would read or write device memory out of
bounds for a field of the wrong type, size or checkerboard).  No GPU is
needed: every case raises QPhiXAMDError before a launch."""
import pytest
import torch


@pytest.fixture(scope="module")
def geom(qamd):
    g = qamd.Geometry((4, 4, 4, 4), t_boundary=-1.0)
    yield g
    g.close()


def _spinor(qamd, geom, prec=None, cb=1, n=None):
    from qphix_codegen_amd.dslash import Field
    prec = qamd.F64 if prec is None else prec
    dt = {qamd.F64: torch.float64, qamd.F32: torch.float32, qamd.F16: torch.float16}[prec]
    return Field(torch.zeros(24 * geom.vcb if n is None else n, dtype=dt), prec, qamd.FIELD_SPINOR, cb=cb)


def test_lexicographic_fields_checked(qamd, geom):
    from qphix_codegen_amd import solver
    V = 2 * geom.vcb
    good_u = torch.zeros(V, 4, 3, 3, 2, dtype=torch.float64)
    bad = {
        "float32 links": torch.zeros(V, 4, 3, 3, 2, dtype=torch.float32),
        "short links": torch.zeros(V - 1, 4, 3, 3, 2, dtype=torch.float64),
        "links of a spinor's size": torch.zeros(V, 4, 3, 2, dtype=torch.float64),
    }
    for what, u in bad.items():
        with pytest.raises(qamd.QPhiXAMDError):
            solver.clover_from_links(geom, u, 4.0, 1.0)
        with pytest.raises(qamd.QPhiXAMDError):
            solver.reunitarize(geom, u)
    # the right shape on the host is still refused: the ABI takes device pointers
    with pytest.raises(qamd.QPhiXAMDError, match="device"):
        solver.clover_from_links(geom, good_u, 4.0, 1.0)
    clov = torch.zeros(V, 2, 36, dtype=torch.float64)
    with pytest.raises(qamd.QPhiXAMDError):
        solver.clover_invert(geom, clov[:-1])
    with pytest.raises(qamd.QPhiXAMDError):
        solver.tmclover_blocks(geom, clov.float(), 0.1, 1, True)
    with pytest.raises(qamd.QPhiXAMDError):
        solver.tmclover_blocks(geom, clov, 0.1, 2, True)


def test_clover_force_arguments(qamd, geom):
    from qphix_codegen_amd import solver
    V = 2 * geom.vcb
    u = torch.zeros(V, 4, 3, 3, 2, dtype=torch.float64)
    sp = torch.zeros(V, 4, 3, 2, dtype=torch.float64)
    # a float32 or non-contiguous force would be written out of bounds
    with pytest.raises(qamd.QPhiXAMDError, match="force"):
        solver.clover_force(geom, u.float(), sp, sp, u, 1.0)
    with pytest.raises(qamd.QPhiXAMDError, match="contiguous"):
        solver.clover_force(geom, u.transpose(0, 1), sp, sp, u, 1.0)
    with pytest.raises(qamd.QPhiXAMDError, match="chi_lex"):
        solver.clover_force(geom, u, sp[:-1], sp, u, 1.0)
    with pytest.raises(qamd.QPhiXAMDError, match="u_lex"):
        solver.clover_force(geom, u, sp, sp, u.float(), 1.0)


def test_operator_arguments(qamd, geom):
    from qphix_codegen_amd import solver
    V = 2 * geom.vcb
    u = torch.zeros(V, 4, 3, 3, 2, dtype=torch.float64)
    clov = torch.zeros(V, 2, 36, dtype=torch.float64)
    with pytest.raises(qamd.QPhiXAMDError):
        solver.EvenOddCloverOperator(geom, u, clov, prec_sloppy=7)
    with pytest.raises(qamd.QPhiXAMDError):
        solver.EvenOddCloverOperator(geom, u, clov, recon_sloppy=13)
    with pytest.raises(qamd.QPhiXAMDError, match="clov_lex"):
        solver.EvenOddCloverOperator(geom, u, clov[:, 0])
    # an operator object without a device operator behind it: every check
    # below must fire before the C call would be reached
    op = object.__new__(solver.EvenOddCloverOperator)
    op._op, op.geom, op.prec_sloppy = None, geom, qamd.F16
    x = _spinor(qamd, geom)
    with pytest.raises(qamd.QPhiXAMDError, match="FP64"):
        op.force(u, _spinor(qamd, geom, prec=qamd.F32), u)
    with pytest.raises(qamd.QPhiXAMDError, match="checkerboard"):
        op.force(u, _spinor(qamd, geom, cb=0), u)
    with pytest.raises(qamd.QPhiXAMDError, match="geometry"):
        op.force(u, _spinor(qamd, geom, n=24 * geom.vcb - 8), u)
    with pytest.raises(qamd.QPhiXAMDError, match="force"):
        op.force(u.float(), x, u)
    with pytest.raises(qamd.QPhiXAMDError, match="u_lex"):
        op.force(u, x, u[:-1])
    with pytest.raises(qamd.QPhiXAMDError, match="precision"):
        op.apply(_spinor(qamd, geom, prec=qamd.F32), _spinor(qamd, geom, prec=qamd.F32))
    cg = solver.InvCG(op)
    with pytest.raises(qamd.QPhiXAMDError, match="FP64"):
        cg(_spinor(qamd, geom, prec=qamd.F16), x, 1e-10)
    with pytest.raises(qamd.QPhiXAMDError, match="checkerboard"):
        cg(x, _spinor(qamd, geom, cb=0), 1e-10)


def test_dslash_force_arguments(qamd, geom):
    from qphix_codegen_amd import autograd
    V = 2 * geom.vcb
    u = torch.zeros(V, 4, 3, 3, 2, dtype=torch.float64)
    chi64 = _spinor(qamd, geom, cb=0)
    psi64 = _spinor(qamd, geom, cb=1)
    with pytest.raises(qamd.QPhiXAMDError, match="share a precision"):
        autograd.dslash_force(geom, u, chi64, _spinor(qamd, geom, prec=qamd.F32, cb=1), u, 1, 0)
    with pytest.raises(qamd.QPhiXAMDError, match="share a precision"):
        autograd.dslash_force(geom, u, _spinor(qamd, geom, prec=qamd.F16, cb=0),
                              _spinor(qamd, geom, prec=qamd.F16, cb=1), u, 1, 0)
    with pytest.raises(qamd.QPhiXAMDError, match="cb"):
        autograd.dslash_force(geom, u, chi64, psi64, u, 1, 1)
    with pytest.raises(qamd.QPhiXAMDError, match="force"):
        autograd.dslash_force(geom, u[:, :2], chi64, psi64, u, 1, 0)


def test_face_buffers_sized_for_both_faces(qamd, geom):
    """fb = 2 (both faces in one launch) needs a buffer of two faces; a
    one-face buffer is refused before the call (host tensors here: the size
    check comes first)."""
    psi = _spinor(qamd, geom, qamd.F32, cb=1)
    one = torch.zeros(geom.face_sites(3) * 12, dtype=torch.float32)
    with pytest.raises(qamd.QPhiXAMDError, match="size"):
        qamd.face_pack(geom, one, psi, 1, 0, 3, 2)
    res = _spinor(qamd, geom, qamd.F32, cb=0)
    g = qamd.Field(torch.zeros(8), qamd.F32, qamd.FIELD_GAUGE, recon=12, cb=0)
    with pytest.raises(qamd.QPhiXAMDError, match="size"):
        qamd.face_finish(geom, 0, res, one, g, 1, 0, 3, 2, 1.0)
    two = torch.zeros(2 * geom.face_sites(3) * 12, dtype=torch.float32)
    assert two.numel() == geom.new_face(qamd.F32, 3, device="cpu", both=True).numel()
