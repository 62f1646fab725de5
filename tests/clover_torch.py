"""Differentiable torch restatement of the clover term built from the links
(test infrastructure: torch autograd of it is the reference gradient for the
clover-term force kernel).

A(x) = alpha + (csw/2) sum_{mu<nu} sigma_{mu nu} (x) F_{mu nu}(x),
sigma_{mu nu} = (i/2)[g_mu, g_nu] in the DeGrand-Rossi basis, F the traceless
part of (Q - Q^dag)/(8i), Q the four plaquette leaves starting and ending at
x, counter-clockwise in the (mu, nu) plane.  A is returned as a (V, 12, 12)
complex matrix in the (spin, colour) index 3*spin + colour of the
lexicographic spinor layout.  tests/test_force_oracle.py checks it against
the oracle's builder.
"""
import torch

from _util import GAMMA


def sigma():
    g = torch.from_numpy(GAMMA).to(torch.complex128)
    out = []
    for mu in range(4):
        for nu in range(mu + 1, 4):
            out.append(0.5j * (g[mu] @ g[nu] - g[nu] @ g[mu]))
    return torch.stack(out)  # (6, 4, 4)


def _shift(dims, mu, d):
    """Index array: site x -> site x + d mu (periodic)."""
    Lx, Ly, Lz, Lt = dims
    idx = torch.arange(Lx * Ly * Lz * Lt).reshape(Lt, Lz, Ly, Lx)
    return torch.roll(idx, shifts=-d, dims=3 - mu).reshape(-1)


def clover_matrix(U, dims, alpha, csw):
    """U: (V, 4, 3, 3) complex128 -> A: (V, 12, 12) complex128."""
    V = U.shape[0]
    sg = sigma()
    dag = lambda m: m.conj().transpose(-1, -2)
    A = alpha * torch.eye(12, dtype=torch.complex128).expand(V, 12, 12).clone()
    plane = 0
    for mu in range(4):
        for nu in range(mu + 1, 4):
            pm, pn, mm, mn = _shift(dims, mu, 1), _shift(dims, nu, 1), _shift(dims, mu, -1), _shift(dims, nu, -1)
            mm_pn = pn[mm]   # x - mu + nu
            mm_mn = mn[mm]   # x - mu - nu
            mn_pm = pm[mn]   # x - nu + mu
            Um, Un = U[:, mu], U[:, nu]
            P1 = Um @ Un[pm] @ dag(Um[pn]) @ dag(Un)
            P2 = Un @ dag(Um[mm_pn]) @ dag(Un[mm]) @ Um[mm]
            P3 = dag(Um[mm]) @ dag(Un[mm_mn]) @ Um[mm_mn] @ Un[mn]
            P4 = dag(Un[mn]) @ Um[mn] @ Un[mn_pm] @ dag(Um)
            Q = P1 + P2 + P3 + P4
            F = (Q - dag(Q)) / 8j
            tr = torch.diagonal(F, dim1=-2, dim2=-1).sum(-1) / 3
            F = F - tr[:, None, None] * torch.eye(3, dtype=torch.complex128)
            # (sigma (x) F)[(s1,i),(s2,j)] = sigma[s1,s2] F[i,j]
            A = A + 0.5 * csw * torch.einsum("ab,vij->vaibj", sg[plane], F).reshape(V, 12, 12)
            plane += 1
    return A


def pack_blocks(A):
    """(V, 12, 12) -> (V, 2, 36) packed Hermitian blocks (the library's
    clover layout: 6 real diagonals, 15 complex lower-triangle entries)."""
    V = A.shape[0]
    out = torch.zeros(V, 2, 36, dtype=torch.float64)
    for b in range(2):
        blk = A[:, 6 * b:6 * b + 6, 6 * b:6 * b + 6]
        for r in range(6):
            out[:, b, r] = blk[:, r, r].real
            for c in range(r):
                od = r * (r - 1) // 2 + c
                out[:, b, 6 + 2 * od] = blk[:, r, c].real
                out[:, b, 6 + 2 * od + 1] = blk[:, r, c].imag
    return out
