"""GPU test of qphix_amd_clover_force_noalloc, the scratch-free variant of
qphix_amd_clover_force, called through the C ABI: on two calls in a row (an
8^4 lattice, then a smaller 4x4x4x6 one, so the second call runs over what
the first left in the static array) it must give exactly the results of the
allocating entry point (the same kernels on the same inputs); after the
first call on the device (which looks up the static array's device address,
and with it the runtime places the array) it must leave the device's free
memory unchanged; and a lattice whose workspace exceeds the static array
must be refused with QPHIX_AMD_ESCRATCH before any launch."""
import ctypes

import pytest
import torch

pytestmark = pytest.mark.gpu

DEV = "cuda:0"
ESCRATCH = -7


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _fields(qamd, dims, seed):
    L = qamd.lattice
    U = L.random_gauge(dims, 0.3, seed, device=DEV)
    chi = L.random_spinor(dims, seed + 1, device=DEV)
    psi = L.random_spinor(dims, seed + 2, device=DEV)
    return U, chi, psi


def test_clover_force_noalloc_matches_allocating_entry(qamd):
    lib = qamd.lib()
    for case, (dims, seed, project, csw, we, wo) in enumerate((((8, 8, 8, 8), 500, 1, 1.5, -0.5, -2.0),
                                                              ((4, 4, 4, 6), 600, 0, 1.0, 0.7, 1.3))):
        print(f"case noalloc clover force dims={dims} project={project}", flush=True)
        geom = qamd.Geometry(dims, t_boundary=-1.0)
        U, chi, psi = _fields(qamd, dims, seed)
        ref = torch.zeros_like(U)
        got = torch.zeros_like(U)
        assert lib.qphix_amd_clover_force(geom.ctx, _p(ref), _p(chi), _p(psi), _p(U), csw, we, wo, project,
                                          _stream()) == 0
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        assert lib.qphix_amd_clover_force_noalloc(geom.ctx, _p(got), _p(chi), _p(psi), _p(U), csw, we, wo,
                                                  project, _stream()) == 0
        torch.cuda.synchronize()
        if case > 0:
            assert torch.cuda.mem_get_info()[0] == free0
        assert torch.count_nonzero(ref) > 0
        assert torch.equal(got, ref)
        # a second call on the same lattice accumulates the same increment
        free0 = torch.cuda.mem_get_info()[0]
        assert lib.qphix_amd_clover_force_noalloc(geom.ctx, _p(got), _p(chi), _p(psi), _p(U), csw, we, wo,
                                                  project, _stream()) == 0
        torch.cuda.synchronize()
        assert torch.cuda.mem_get_info()[0] == free0
        assert torch.equal(got, 2 * ref)


def test_clover_force_noalloc_refuses_large_lattice(qamd):
    """16^3x48 needs 196,608 sites x 384 B = 75.5 MB of weights > 64 MiB."""
    lib = qamd.lib()
    geom = qamd.Geometry((16, 16, 16, 48), t_boundary=-1.0)
    dummy = torch.zeros(16, dtype=torch.float64, device=DEV)
    assert lib.qphix_amd_clover_force_noalloc(geom.ctx, _p(dummy), _p(dummy), _p(dummy), _p(dummy), 1.0, 1.0,
                                              1.0, 1, _stream()) == ESCRATCH
    torch.cuda.synchronize()
    assert torch.count_nonzero(dummy) == 0
    assert b"static scratch" in lib.qphix_amd_error_string(ESCRATCH)
