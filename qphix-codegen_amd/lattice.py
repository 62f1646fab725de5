"""Seeded synthetic lattices in the lexicographic FP64 layout.

site = ((t*Lz + z)*Ly + y)*Lx + x.  Shapes (float64):
  spinor       (V, 4, 3, 2)           spin, colour, re/im
  gauge        (V, 4, 3, 3, 2)        U_mu(x), row-major
  clover       (V, 2, 36)             per block: 6 real diagonals, then 15
                                      complex lower-triangle entries,
                                      od(r, c) = r(r-1)/2 + c for r > c
  full clover  (V, 2, 6, 6, 2)

The packed clover order is the reference's CloverBlock
(codegen/include/decl_common.h:15-32, codegen/dslash_common.cc:1150-1166);
the full block is its FullCloverBlock (decl_common.h:56-61).

Everything is generated with torch so that the same code builds 8^4 test
lattices on the CPU and 32^3x64 benchmark lattices on the GPU.
"""
import torch


def volume(dims):
    Lx, Ly, Lz, Lt = dims
    return Lx * Ly * Lz * Lt


def parity(dims, device="cpu"):
    """(V,) int tensor of (x+y+z+t) & 1 in lexicographic order."""
    Lx, Ly, Lz, Lt = dims
    t, z, y, x = torch.meshgrid(torch.arange(Lt, device=device), torch.arange(Lz, device=device),
                                torch.arange(Ly, device=device), torch.arange(Lx, device=device),
                                indexing="ij")
    return ((x + y + z + t) & 1).reshape(-1)


def _gen(seed, device):
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    return g


def gellmann(device="cpu"):
    """The 8 Gell-Mann matrices, (8, 3, 3) complex128."""
    l = torch.zeros(8, 3, 3, dtype=torch.complex128, device=device)
    l[0, 0, 1] = l[0, 1, 0] = 1
    l[1, 0, 1], l[1, 1, 0] = -1j, 1j
    l[2, 0, 0], l[2, 1, 1] = 1, -1
    l[3, 0, 2] = l[3, 2, 0] = 1
    l[4, 0, 2], l[4, 2, 0] = -1j, 1j
    l[5, 1, 2] = l[5, 2, 1] = 1
    l[6, 1, 2], l[6, 2, 1] = -1j, 1j
    l[7, 0, 0] = l[7, 1, 1] = 1 / 3 ** 0.5
    l[7, 2, 2] = -2 / 3 ** 0.5
    return l


def reunitarize(u):
    """Project (..., 3, 3) complex matrices onto SU(3): Gram-Schmidt on the
    first two rows, third row = conj(row0 x row1)."""
    r0 = u[..., 0, :]
    r0 = r0 / torch.linalg.vector_norm(r0, dim=-1, keepdim=True)
    r1 = u[..., 1, :]
    r1 = r1 - (r0.conj() * r1).sum(-1, keepdim=True) * r0
    r1 = r1 / torch.linalg.vector_norm(r1, dim=-1, keepdim=True)
    r2 = torch.linalg.cross(r0, r1, dim=-1).conj()
    return torch.stack([r0, r1, r2], dim=-2)


def su3_from_algebra(w):
    """exp(i sum_a w_a lambda_a / 2), reunitarised; w is (..., 8) float64."""
    lam = gellmann(w.device)
    h = torch.einsum("...a,aij->...ij", w.to(torch.complex128), lam) * 0.5
    return reunitarize(torch.linalg.matrix_exp(1j * h))


def cplx_to_real(z):
    return torch.view_as_real(z.contiguous()).contiguous()


def real_to_cplx(r):
    return torch.view_as_complex(r.contiguous())


def random_gauge(dims, sigma, seed, device="cpu"):
    """Random SU(3) links U = exp(i H), H = sum_a w_a lambda_a/2,
    w ~ N(0, sigma^2), reunitarised.  Returns (V, 4, 3, 3, 2) float64."""
    V = volume(dims)
    w = torch.randn(V, 4, 8, generator=_gen(seed, device), dtype=torch.float64, device=device) * sigma
    return cplx_to_real(su3_from_algebra(w))


def unit_gauge(dims, device="cpu"):
    V = volume(dims)
    u = torch.eye(3, dtype=torch.complex128, device=device).expand(V, 4, 3, 3)
    return cplx_to_real(u)


def random_spinor(dims, seed, device="cpu"):
    """i.i.d. unit Gaussian re/im per component, (V, 4, 3, 2) float64."""
    V = volume(dims)
    return torch.randn(V, 4, 3, 2, generator=_gen(seed, device), dtype=torch.float64, device=device)


def random_clover(dims, seed, device="cpu", diag_sigma=0.1, off_sigma=0.05):
    """Packed Hermitian blocks: diagonals 1 + diag_sigma*N(0,1), off-diagonals
    off_sigma*(N + iN).  Returns (V, 2, 36) float64."""
    V = volume(dims)
    g = _gen(seed, device)
    diag = 1.0 + diag_sigma * torch.randn(V, 2, 6, generator=g, dtype=torch.float64, device=device)
    off = off_sigma * torch.randn(V, 2, 30, generator=g, dtype=torch.float64, device=device)
    return torch.cat([diag, off], dim=-1)


_OD = [(r, c) for r in range(6) for c in range(r)]  # od index order: r(r-1)/2 + c


def clover_dense(clov):
    """(..., 36) packed block -> (..., 6, 6) complex128 Hermitian."""
    shp = clov.shape[:-1]
    a = torch.zeros(*shp, 6, 6, dtype=torch.complex128, device=clov.device)
    idx = torch.arange(6, device=clov.device)
    a[..., idx, idx] = clov[..., :6].to(torch.complex128)
    off = torch.complex(clov[..., 6::2], clov[..., 7::2])
    rows = torch.tensor([r for r, _ in _OD], device=clov.device)
    cols = torch.tensor([c for _, c in _OD], device=clov.device)
    a[..., rows, cols] = off
    a[..., cols, rows] = off.conj()
    return a


def clover_pack(a):
    """(..., 6, 6) Hermitian complex -> (..., 36) packed float64."""
    idx = torch.arange(6, device=a.device)
    diag = a[..., idx, idx].real
    rows = torch.tensor([r for r, _ in _OD], device=a.device)
    cols = torch.tensor([c for _, c in _OD], device=a.device)
    off = a[..., rows, cols]
    offr = torch.stack([off.real, off.imag], dim=-1).reshape(*a.shape[:-2], 30)
    return torch.cat([diag, offr], dim=-1)


def invert_clover(clov):
    """Packed Hermitian blocks -> packed inverse blocks (FP64)."""
    a = clover_dense(clov)
    inv = torch.linalg.inv(a)
    inv = 0.5 * (inv + inv.conj().transpose(-1, -2))
    return clover_pack(inv)


def tmclover_blocks(clov, mu, sign, invert):
    """Full complex blocks of A + i*sign*mu*g5 (g5 = +1 on block 0, -1 on
    block 1), or of its inverse.  Returns (V, 2, 6, 6, 2) float64."""
    a = clover_dense(clov)
    g5 = torch.tensor([1.0, -1.0], dtype=torch.float64, device=clov.device)
    eye = torch.eye(6, dtype=torch.complex128, device=clov.device)
    a = a + 1j * sign * mu * g5.view(1, 2, 1, 1) * eye
    if invert:
        a = torch.linalg.inv(a)
    return cplx_to_real(a)
