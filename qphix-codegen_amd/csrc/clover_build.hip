// clover_build.hip -- the clover term from the gauge links, its inverse and
// the twisted-mass clover blocks, all in FP64 on the lexicographic layout.
//
// No counterpart in the reference: its kernels take A or A^-1 as input
// (codegen/dslash_common.cc:1098-1232).  Convention:
//   A(x) = alpha + (csw/2) sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x),
//   sigma_{mu nu} = (i/2)[g_mu, g_nu] (DeGrand-Rossi), F_{mu nu} the
//   traceless part of (Q - Q^dagger)/(8i), Q the four plaquette leaves at x.
// sigma_{mu nu} is block diagonal in spin (blocks {0,1} and {2,3}), so A is
// stored as two packed Hermitian 6x6 blocks (6 real diagonals + 15 complex
// lower-triangle entries, od(r,c) = r(r-1)/2 + c).
#include "blas.h"

namespace qamd {

namespace {

struct M3 {
  cx<double> m[3][3];
};

// Link U_mu(s) of the lexicographic [V][4][18] link field; DAG: its adjoint.
template <bool DAG = false>
__device__ __forceinline__ M3 ldl(const double *U, long s, int mu)
{
  M3 r;
  const double *p = U + (s * 4 + mu) * 18;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const double re = p[(i * 3 + j) * 2], im = p[(i * 3 + j) * 2 + 1];
      if (DAG)
        r.m[j][i] = {re, -im};
      else
        r.m[i][j] = {re, im};
    }
  return r;
}

__device__ __forceinline__ M3 mul(const M3 &a, const M3 &b)
{
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      cx<double> v = a.m[i][0] * b.m[0][j];
      v = v + a.m[i][1] * b.m[1][j];
      v = v + a.m[i][2] * b.m[2][j];
      r.m[i][j] = v;
    }
  return r;
}

// a * b^dagger
__device__ __forceinline__ M3 mul_nd(const M3 &a, const M3 &b)
{
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      cx<double> v = a.m[i][0] * conj(b.m[j][0]);
      v = v + a.m[i][1] * conj(b.m[j][1]);
      v = v + a.m[i][2] * conj(b.m[j][2]);
      r.m[i][j] = v;
    }
  return r;
}

// a^dagger * b
__device__ __forceinline__ M3 mul_dn(const M3 &a, const M3 &b)
{
  M3 r;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      cx<double> v = cmul_conj(a.m[0][i], b.m[0][j]);
      v = v + cmul_conj(a.m[1][i], b.m[1][j]);
      v = v + cmul_conj(a.m[2][i], b.m[2][j]);
      r.m[i][j] = v;
    }
  return r;
}

__device__ __forceinline__ long lex_of(const Geom &g, const int c[4])
{
  return (((long)c[3] * g.Lz + c[2]) * g.Ly + c[1]) * g.Lx + c[0];
}

__device__ __forceinline__ long shifted(const Geom &g, const int c0[4], int mu, int dmu, int nu, int dnu)
{
  const int L[4] = {g.Lx, g.Ly, g.Lz, g.Lt};
  int c[4] = {c0[0], c0[1], c0[2], c0[3]};
  c[mu] = (c[mu] + dmu + L[mu]) % L[mu];
  c[nu] = (c[nu] + dnu + L[nu]) % L[nu];
  return lex_of(g, c);
}

// sigma_{mu nu}[a][b] for the DeGrand-Rossi basis, computed once on the
// host and passed by value (16 complex entries per plane, 6 planes).
struct SigmaTable {
  double re[6][4][4], im[6][4][4];
};

__global__ __launch_bounds__(128) void clover_from_links_k(double *clov, const double *U, Geom g,
                                                           double alpha, double csw, SigmaTable sig)
{
  const long V = (long)g.vcb * 2;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  int c[4];
  {
    long r = s;
    c[0] = (int)(r % g.Lx);
    r /= g.Lx;
    c[1] = (int)(r % g.Ly);
    r /= g.Ly;
    c[2] = (int)(r % g.Lz);
    c[3] = (int)(r / g.Lz);
  }
  cx<double> A[2][6][6];
#pragma unroll
  for (int b = 0; b < 2; b++)
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
      for (int j = 0; j < 6; j++) A[b][i][j] = {i == j ? alpha : 0.0, 0.0};
  int plane = 0;
  for (int mu = 0; mu < 4; mu++) {
    for (int nu = mu + 1; nu < 4; nu++, plane++) {
      const long xpm = shifted(g, c, mu, 1, nu, 0), xpn = shifted(g, c, mu, 0, nu, 1);
      const long xmm = shifted(g, c, mu, -1, nu, 0), xmn = shifted(g, c, mu, 0, nu, -1);
      const long xmmpn = shifted(g, c, mu, -1, nu, 1), xmmmn = shifted(g, c, mu, -1, nu, -1);
      const long xpmmn = shifted(g, c, mu, 1, nu, -1);
      // four leaves, each starting and ending at x
      M3 P1 = mul_nd(mul(ldl(U, s, mu), ldl(U, xpm, nu)), mul(ldl(U, s, nu), ldl(U, xpn, mu)));
      M3 P2 = mul(mul_nd(ldl(U, s, nu), ldl(U, xmmpn, mu)),
                  mul_dn(ldl(U, xmm, nu), ldl(U, xmm, mu)));
      M3 P3 = mul(mul_dn(mul(ldl(U, xmmmn, nu), ldl(U, xmm, mu)), ldl(U, xmmmn, mu)),
                  ldl(U, xmn, nu));
      M3 P4 = mul(mul_dn(ldl(U, xmn, nu), ldl(U, xmn, mu)),
                  mul_nd(ldl(U, xpmmn, nu), ldl(U, s, mu)));
      cx<double> F[3][3];
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          cx<double> q = P1.m[i][j] + P2.m[i][j] + P3.m[i][j] + P4.m[i][j];
          cx<double> qd = P1.m[j][i] + P2.m[j][i] + P3.m[j][i] + P4.m[j][i];
          // (q - conj(qd)) / (8 i)
          cx<double> d = {q.re - qd.re, q.im + qd.im};
          F[i][j] = {d.im / 8.0, -d.re / 8.0};
        }
      cx<double> tr = {(F[0][0].re + F[1][1].re + F[2][2].re) / 3.0,
                       (F[0][0].im + F[1][1].im + F[2][2].im) / 3.0};
#pragma unroll
      for (int i = 0; i < 3; i++) F[i][i] = F[i][i] - tr;
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int r = 0; r < 6; r++)
#pragma unroll
          for (int q = 0; q < 6; q++) {
            const int s1 = 2 * b + r / 3, s2 = 2 * b + q / 3;
            cx<double> sg = {sig.re[plane][s1][s2], sig.im[plane][s1][s2]};
            A[b][r][q] = A[b][r][q] + (0.5 * csw) * (sg * F[r % 3][q % 3]);
          }
    }
  }
#pragma unroll
  for (int b = 0; b < 2; b++) {
    double *blk = clov + s * 72 + b * 36;
#pragma unroll
    for (int r = 0; r < 6; r++) {
      blk[r] = A[b][r][r].re;
#pragma unroll
      for (int q = 0; q < r; q++) {
        const int od = r * (r - 1) / 2 + q;
        blk[6 + 2 * od] = A[b][r][q].re;
        blk[6 + 2 * od + 1] = A[b][r][q].im;
      }
    }
  }
}

// Inverse of a Hermitian positive-definite 6x6 block from its lower triangle
// (FP64, every index compile-time so the block stays in registers):
// Cholesky B = L L^dagger, M = L^-1, B^-1 = M^dagger M; the lower triangle of
// the (exactly Hermitian) inverse is returned in Bi.
__device__ __forceinline__ void hpd_inverse6(const cx<double> B[6][6], cx<double> Bi[6][6])
{
  cx<double> L[6][6];
  double dinv[6];
#pragma unroll
  for (int j = 0; j < 6; j++) {
    double d = B[j][j].re;
#pragma unroll
    for (int k = 0; k < j; k++) d -= L[j][k].re * L[j][k].re + L[j][k].im * L[j][k].im;
    const double lj = sqrt(d);
    dinv[j] = 1.0 / lj;
    L[j][j] = {lj, 0.0};
#pragma unroll
    for (int i = j + 1; i < 6; i++) {
      cx<double> t = B[i][j];
#pragma unroll
      for (int k = 0; k < j; k++) t = t - L[i][k] * conj(L[j][k]);
      L[i][j] = {t.re * dinv[j], t.im * dinv[j]};
    }
  }
  cx<double> M[6][6];  // lower triangular L^-1
#pragma unroll
  for (int j = 0; j < 6; j++) {
    M[j][j] = {dinv[j], 0.0};
#pragma unroll
    for (int i = j + 1; i < 6; i++) {
      cx<double> t = {0.0, 0.0};
#pragma unroll
      for (int k = j; k < i; k++) t = cmac(t, L[i][k], M[k][j]);
      M[i][j] = {-t.re * dinv[i], -t.im * dinv[i]};
    }
  }
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      cx<double> t = {0.0, 0.0};
#pragma unroll
      for (int k = i; k < 6; k++) t = cmac_conj(t, M[k][i], M[k][j]);
      Bi[i][j] = t;
    }
}

__device__ __forceinline__ void load_packed(const double *blk, cx<double> A[6][6])
{
#pragma unroll
  for (int r = 0; r < 6; r++) {
    A[r][r] = {blk[r], 0.0};
#pragma unroll
    for (int q = 0; q < r; q++) {
      const int od = r * (r - 1) / 2 + q;
      A[r][q] = {blk[6 + 2 * od], blk[6 + 2 * od + 1]};
      A[q][r] = {blk[6 + 2 * od], -blk[6 + 2 * od + 1]};
    }
  }
}

// One Hermitian block A per thread (k even: spins 0,1, g5 = +1; k odd: spins
// 2,3, g5 = -1).  mode 0: packed A^-1; mode 1: full A + i m, m = sign mu g5;
// mode 2: full (A + i m)^-1.  Both inverses are (A - i m)(A^2 + m^2)^-1 (m = 0
// for mode 0): the bracket is Hermitian positive definite for any invertible
// Hermitian A, so no pivoting is needed and nothing is assumed about the
// sign of A's eigenvalues.  The rounding error grows like cond(A)^2 eps,
// about 1e-15 for clover blocks near alpha 1 (cond 1.5-3).
__global__ __launch_bounds__(64) void clover_blocks_k(double *out, const double *clov, long nblk,
                                                      int mode, double mu, int sign)
{
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nblk) return;
  const double m = mode == 0 ? 0.0 : sign * mu * ((k & 1) == 0 ? 1.0 : -1.0);
  cx<double> A[6][6];
  load_packed(clov + k * 36, A);
  if (mode == 1) {
    double *o = out + k * 72;
#pragma unroll
    for (int r = 0; r < 6; r++)
#pragma unroll
      for (int q = 0; q < 6; q++) {
        o[(r * 6 + q) * 2] = A[r][q].re;
        o[(r * 6 + q) * 2 + 1] = A[r][q].im + (r == q ? m : 0.0);
      }
    return;
  }
  cx<double> B[6][6], Bi[6][6];  // lower triangles of A^2 + m^2 and its inverse
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      cx<double> t = {i == j ? m * m : 0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 6; q++) t = cmac(t, A[i][q], A[q][j]);
      B[i][j] = t;
    }
  hpd_inverse6(B, Bi);
  // entries are stored as they are formed (mode 0: the packed lower triangle)
  double *o = out + k * (mode == 0 ? 36 : 72);
#pragma unroll
  for (int r = 0; r < 6; r++)
#pragma unroll
    for (int c = 0; c < 6; c++) {
      if (mode == 0 && c > r) continue;  // Hermitian: the lower triangle suffices
      cx<double> t = {0.0, 0.0};
#pragma unroll
      for (int q = 0; q < 6; q++) {
        const cx<double> a = {A[r][q].re, A[r][q].im - (r == q ? m : 0.0)};
        const cx<double> b = q >= c ? Bi[q][c] : conj(Bi[c][q]);
        t = cmac(t, a, b);
      }
      if (mode != 0) {
        o[(r * 6 + c) * 2] = t.re;
        o[(r * 6 + c) * 2 + 1] = t.im;
      } else if (r == c) {
        o[r] = t.re;
      } else {
        const int od = r * (r - 1) / 2 + c;
        o[6 + 2 * od] = t.re;
        o[6 + 2 * od + 1] = t.im;
      }
    }
}

SigmaTable make_sigma()
{
  // DeGrand-Rossi gamma matrices as (re, im) tables
  double gr[4][4][4] = {}, gi[4][4][4] = {};
  gi[0][0][3] = 1; gi[0][1][2] = 1; gi[0][2][1] = -1; gi[0][3][0] = -1;
  gr[1][0][3] = -1; gr[1][1][2] = 1; gr[1][2][1] = 1; gr[1][3][0] = -1;
  gi[2][0][2] = 1; gi[2][1][3] = -1; gi[2][2][0] = -1; gi[2][3][1] = 1;
  gr[3][0][2] = 1; gr[3][1][3] = 1; gr[3][2][0] = 1; gr[3][3][1] = 1;
  SigmaTable t{};
  int plane = 0;
  for (int mu = 0; mu < 4; mu++)
    for (int nu = mu + 1; nu < 4; nu++, plane++)
      for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) {
          double cr = 0, ci = 0;  // [g_mu, g_nu]_{ab}
          for (int k = 0; k < 4; k++) {
            cr += gr[mu][a][k] * gr[nu][k][b] - gi[mu][a][k] * gi[nu][k][b];
            ci += gr[mu][a][k] * gi[nu][k][b] + gi[mu][a][k] * gr[nu][k][b];
            cr -= gr[nu][a][k] * gr[mu][k][b] - gi[nu][a][k] * gi[mu][k][b];
            ci -= gr[nu][a][k] * gi[mu][k][b] + gi[nu][a][k] * gr[mu][k][b];
          }
          // (i/2) * commutator
          t.re[plane][a][b] = -0.5 * ci;
          t.im[plane][a][b] = 0.5 * cr;
        }
  return t;
}

// ---- link derivative of the clover term ------------------------------------
// For L = Re sum_x w(x) chi(x)^dag A(x) psi(x), w = w_even / w_odd by site
// parity, dA = (csw/2) sum_{mu<nu} sigma_{mu nu} (x) dF_{mu nu}, and
//   Re tr[dF L_mn] = Re tr[dQ_mn K_mn],  K_mn = (csw/2)(L' + L'^dag)/(8i),
// with L_mn[j][i] = w sum_{a,b} sigma_mn[a][b] psi[b][j] conj(chi[a][i]) the
// spin-traced colour matrix and L' its traceless part (F is the traceless
// part of (Q - Q^dag)/(8i)).  Each link then collects, from the 24 leaves it
// lies on (2 squares in each of its 3 planes, every square a leaf of its 4
// corners), G with dL = Re tr(dU G).
// Spinor sources of the weights: lexicographic FP64 fields [V][24], or (the
// operator's force) native FP64 spinors per checkerboard.
struct WeightSrc {
  const double *chi, *psi;          // lexicographic, or null
  const double *chi_cb[2], *psi_cb[2];  // native, by checkerboard
};

__global__ __launch_bounds__(128) void clover_weights_k(double *K, WeightSrc src, Geom g, double csw, double w_even,
                                                        double w_odd, SigmaTable sig)
{
  const long V = (long)g.vcb * 2;
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  long r = s;
  const int x = (int)(r % g.Lx);
  r /= g.Lx;
  const int y = (int)(r % g.Ly);
  r /= g.Ly;
  const int z = (int)(r % g.Lz);
  const int t = (int)(r / g.Lz);
  const int par = (x + y + z + t) & 1;
  const double w = par ? w_odd : w_even;
  double ar[24], br[24];
  if (src.chi) {
#pragma unroll
    for (int k = 0; k < 24; k++) {
      ar[k] = src.chi[s * 24 + k];
      br[k] = src.psi[s * 24 + k];
    }
  } else {
    const int cbi = (int)(s >> 1);  // (((t Lz + z) Ly + y) Lx + x) / 2, Lx even
    load_reals<double, spinor_w<double>(), 0, 24>(src.chi_cb[par], g.vcb, cbi, ar);
    load_reals<double, spinor_w<double>(), 0, 24>(src.psi_cb[par], g.vcb, cbi, br);
  }
  cx<double> a[4][3], b[4][3];  // chi, psi
#pragma unroll
  for (int sp = 0; sp < 4; sp++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      a[sp][c] = {ar[(sp * 3 + c) * 2], ar[(sp * 3 + c) * 2 + 1]};
      b[sp][c] = {br[(sp * 3 + c) * 2], br[(sp * 3 + c) * 2 + 1]};
    }
  for (int p = 0; p < 6; p++) {
    cx<double> L[3][3];
#pragma unroll
    for (int j = 0; j < 3; j++)
#pragma unroll
      for (int i = 0; i < 3; i++) L[j][i] = {0.0, 0.0};
#pragma unroll
    for (int sa = 0; sa < 4; sa++) {
      cx<double> u[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};  // (sigma psi)[sa][j]
#pragma unroll
      for (int sb = 0; sb < 4; sb++) {
        const cx<double> sg = {sig.re[p][sa][sb], sig.im[p][sa][sb]};
#pragma unroll
        for (int j = 0; j < 3; j++) u[j] = cmac(u[j], sg, b[sb][j]);
      }
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) L[j][i] = cmac_conj(L[j][i], a[sa][i], u[j]);
    }
    // H = L' + L'^dag, K = f H / i = -i f H: anti-Hermitian and traceless,
    // stored as its 8 independent reals (see ld_k)
    const double trr = (L[0][0].re + L[1][1].re + L[2][2].re) / 3.0;
    double *o = K + (long)p * kKReals * V + s;  // component e at o[e * V]
    const double f = 0.5 * csw * w / 8.0;
    o[0] = -f * (2.0 * L[0][0].re - 2.0 * trr);
    o[V] = -f * (2.0 * L[1][1].re - 2.0 * trr);
    const int rr[3] = {0, 0, 1}, cc[3] = {1, 2, 2};
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const cx<double> h = L[rr[k]][cc[k]] + conj(L[cc[k]][rr[k]]);
      o[(2 + 2 * k) * V] = f * h.im;
      o[(3 + 2 * k) * V] = -f * h.re;
    }
  }
}

// K is stored component-major, K[(p * 8 + e) * V + s], so that a wave's
// loads of one component are contiguous; the 8 reals of the anti-Hermitian
// traceless K of plane p at site s: Im K00, Im K11, K01, K02, K12 (re, im);
// K22 = -i (Im K00 + Im K11), K_ji = -conj(K_ij).
__device__ __forceinline__ M3 ld_k(const double *K, long V, long s, int p)
{
  const double *q = K + (long)p * kKReals * V + s;
  const double d0 = q[0], d1 = q[V];
  const cx<double> k01 = {q[2 * V], q[3 * V]}, k02 = {q[4 * V], q[5 * V]}, k12 = {q[6 * V], q[7 * V]};
  M3 r;
  r.m[0][0] = {0.0, d0};
  r.m[1][1] = {0.0, d1};
  r.m[2][2] = {0.0, -(d0 + d1)};
  r.m[0][1] = k01;
  r.m[1][0] = {-k01.re, k01.im};
  r.m[0][2] = k02;
  r.m[2][0] = {-k02.re, k02.im};
  r.m[1][2] = k12;
  r.m[2][1] = {-k12.re, k12.im};
  return r;
}

// acc += a b
__device__ __forceinline__ void mac(M3 &acc, const M3 &a, const M3 &b)
{
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      cx<double> v = acc.m[i][j];
#pragma unroll
      for (int k = 0; k < 3; k++) v = cmac(v, a.m[i][k], b.m[k][j]);
      acc.m[i][j] = v;
    }
}

__device__ __forceinline__ void acc_dag(M3 &G, const M3 &a)
{
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) G.m[i][j] = G.m[i][j] + conj(a.m[j][i]);
}

__device__ __forceinline__ int plane_of(int a, int b)  // a < b
{
  return a == 0 ? b - 1 : (a == 1 ? b + 1 : 5);
}

// The link U_mu(y) (coordinates c) collects G = the derivative of
// sum_x sum_{planes} Re tr(K_p(x) Q_p(x)) with respect to it, summed over
// the two squares of each plane (mu, nu) that contain the link, each square
// with its four corner insertions; squares are U_a(z) U_b(z+a) U_a(z+b)^dag
// U_b(z)^dag, a < b, with K of plane (a, b) inserted before the factor that
// starts at each corner.  square_g adds square q (0..5: nu the (q/2)-th
// direction other than mu, q odd the square on the far side) to G.  Each
// square's four terms are evaluated as nested products (8 matrix products
// per square) so that few 3x3 matrices are live at once.
__device__ __forceinline__ void square_g(M3 &G, const double *K, const double *U, const Geom &g, long V,
                                         const int c[4], long y, int mu, int q)
{
  int nu = q >> 1;
  if (nu >= mu) nu++;
  const int a = mu < nu ? mu : nu, b = mu < nu ? nu : mu;
  const int p = plane_of(a, b);
  const int kind = (mu == a ? 0 : 2) + (q & 1);
  if (kind == 0) {  // square based at y, the link its first factor:
    // G += U2 (U3 (U4 K0 + K3 U4) + K2 U3 U4) + K1 U2 U3 U4
    const long ya = shifted(g, c, a, 1, b, 0), yab = shifted(g, c, a, 1, b, 1);
    const long yb = shifted(g, c, a, 0, b, 1);
    const M3 U4 = ldl<true>(U, y, b);
    M3 T = mul(U4, ld_k(K, V, y, p));
    mac(T, ld_k(K, V, yb, p), U4);
    const M3 U3 = ldl<true>(U, yb, a);
    const M3 U34 = mul(U3, U4);
    M3 S = mul(U3, T);
    mac(S, ld_k(K, V, yab, p), U34);
    const M3 U2 = ldl(U, ya, b);
    mac(G, U2, S);
    mac(G, ld_k(K, V, ya, p), mul(U2, U34));
  } else if (kind == 1) {  // square based at y - b, the link its third factor (daggered):
    // M = V4 ((K0 V1 + V1 K1) V2 + V1 V2 K2) + K3 V4 V1 V2, G += M^dag
    const long ymb = shifted(g, c, a, 0, b, -1), ymba = shifted(g, c, a, 1, b, -1);
    const long ya = shifted(g, c, a, 1, b, 0);
    const M3 V1 = ldl(U, ymb, a);
    M3 T = mul(ld_k(K, V, ymb, p), V1);
    mac(T, V1, ld_k(K, V, ymba, p));
    const M3 V2 = ldl(U, ymba, b);
    const M3 V12 = mul(V1, V2);
    M3 S = mul(T, V2);
    mac(S, V12, ld_k(K, V, ya, p));
    const M3 V4 = ldl<true>(U, ymb, b);
    M3 M = mul(V4, S);
    mac(M, ld_k(K, V, y, p), mul(V4, V12));
    acc_dag(G, M);
  } else if (kind == 2) {  // square based at y - a, the link its second factor:
    // G += W3 (W4 (K0 W1 + W1 K1) + K3 W4 W1) + K2 W3 W4 W1
    const long yma = shifted(g, c, a, -1, b, 0), ymab = shifted(g, c, a, -1, b, 1);
    const long yb = shifted(g, c, a, 0, b, 1);
    const M3 W1 = ldl(U, yma, a);
    M3 X = mul(ld_k(K, V, yma, p), W1);
    mac(X, W1, ld_k(K, V, y, p));
    const M3 W4 = ldl<true>(U, yma, b);
    M3 S = mul(W4, X);
    const M3 W41 = mul(W4, W1);
    mac(S, ld_k(K, V, ymab, p), W41);
    const M3 W3 = ldl<true>(U, ymab, a);
    mac(G, W3, S);
    mac(G, ld_k(K, V, yb, p), mul(W3, W41));
  } else {  // square based at y, the link its fourth factor (daggered):
    // M = U1 (U2 (K2 U3 + U3 K3) + K1 U2 U3) + K0 U1 U2 U3, G += M^dag
    const long ya = shifted(g, c, a, 1, b, 0), yab = shifted(g, c, a, 1, b, 1);
    const long yb = shifted(g, c, a, 0, b, 1);
    const M3 U3 = ldl<true>(U, yb, a);
    M3 X = mul(ld_k(K, V, yab, p), U3);
    mac(X, U3, ld_k(K, V, yb, p));
    const M3 U2 = ldl(U, ya, b);
    M3 S = mul(U2, X);
    const M3 U23 = mul(U2, U3);
    mac(S, ld_k(K, V, ya, p), U23);
    const M3 U1 = ldl(U, y, a);
    M3 M = mul(U1, S);
    mac(M, ld_k(K, V, y, p), mul(U1, U23));
    acc_dag(G, M);
  }
}

// Hopping terms fused into the gather (their G added before the
// projection).
struct LinkHop {
  HopTerm term[2];
  int n;
  HopCoef k;
};

// The gather's work-groups: one chunk of kLinkChunk sites of a cross-section
// (cross_sweep) and one direction mu.  At 48^3 x 96 (960 B of K and links
// per site) a cross-section is 8 rows, 369 KB, and a t-neighbour was read
// 2.2 MB of data earlier, against 13 MB when each XCD swept whole t-slabs.
constexpr int kLinkChunk = 64;

// One thread per link (y, mu), its six squares one per loop iteration (no
// loads hoisted across squares), then the hopping terms.  Output:
// scale * TA(U G) (project) or scale * G^dagger, added to the force field of
// the same layout as U.
template <int HOP>
__global__ __launch_bounds__(kLinkChunk) void clover_links_k(double *force, const double *K, const double *U, Geom g,
                                                             CrossSweep w, double scale, int project, LinkHop hop)
{
  const long V = (long)g.vcb * 2;
  int mu, c[4];
  const long y = cross_sweep_site(g, w, (int)threadIdx.x, mu, c);
  if (y < 0) return;  // the last chunk of a cross-section may be partial
  M3 G;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) G.m[i][k] = {0.0, 0.0};
#pragma unroll 1
  for (int q = 0; q < 6; q++) square_g(G, K, U, g, V, c, y, mu, q);
  if constexpr (HOP) {
    const int cbi_y = (((c[3] * g.Lz + c[2]) * g.Ly + c[1]) * g.Lxh) + (c[0] >> 1);
#pragma unroll 1
    for (int h = 0; h < hop.n; h++) hop_g_add<double>(G.m, hop.term[h], hop.k, g, c, cbi_y, mu);
  }
  if (!project) {  // G^dagger
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const int e = (i * 3 + k) * 2;
        double *op = force + (y * 4 + mu) * 18 + e;
        double *oi = force + (y * 4 + mu) * 18 + e + 1;
        *op += scale * G.m[k][i].re;
        *oi -= scale * G.m[k][i].im;
      }
    return;
  }
  // F = TA(U G)
  const M3 W = mul(ldl(U, y, mu), G);
  const double trim = (W.m[0][0].im + W.m[1][1].im + W.m[2][2].im) / 3.0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const int e = (i * 3 + k) * 2;
      double *op = force + (y * 4 + mu) * 18 + e;
      double *oi = force + (y * 4 + mu) * 18 + e + 1;
      *op += scale * 0.5 * (W.m[i][k].re - W.m[k][i].re);
      *oi += scale * (0.5 * (W.m[i][k].im + W.m[k][i].im) - (i == k ? trim : 0.0));
    }
}

template <int HOP>
hipError_t launch_clover_links(const qphix_amd_ctx *ctx, double *force, const double *K, const double *U,
                               double scale, int project, const LinkHop &lh, hipStream_t st)
{
  const CrossSweep w = cross_sweep(ctx->g, kLinkChunk, 4);
  const long nb = cross_sweep_blocks(ctx->g, w);
  if (nb > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL((clover_links_k<HOP>), dim3((unsigned)nb), dim3(kLinkChunk), 0, st, force, K, U, ctx->g, w,
                     scale, project, lh);
  return hipGetLastError();
}

}  // namespace

hipError_t run_clover_from_links(const qphix_amd_ctx *ctx, double *clov, const double *U,
                                 double alpha, double csw, hipStream_t st)
{
  const long V = (long)ctx->g.vcb * 2;
  const int bs = 128;
  hipLaunchKernelGGL(clover_from_links_k, dim3(blocks_for(V, bs)), dim3(bs), 0, st, clov, U, ctx->g,
                     alpha, csw, make_sigma());
  return hipGetLastError();
}

hipError_t run_clover_force(const qphix_amd_ctx *ctx, double *force, double *K, const double *chi,
                            const double *psi, const double *U, double csw, double w_even, double w_odd,
                            double scale, int project, hipStream_t st)
{
  const long V = (long)ctx->g.vcb * 2;
  const int bs = 128;
  WeightSrc src{};
  src.chi = chi;
  src.psi = psi;
  hipLaunchKernelGGL(clover_weights_k, dim3(blocks_for(V, bs)), dim3(bs), 0, st, K, src, ctx->g, csw, w_even,
                     w_odd, make_sigma());
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_clover_links<0>(ctx, force, K, U, scale, project, LinkHop{}, st);
}

hipError_t run_clover_hop_force(const qphix_amd_ctx *ctx, double *force, double *K, const double *const chi_cb[2],
                                const double *const psi_cb[2], const double *U, double csw, double w_even,
                                double w_odd, double scale, int project, const HopTerm *hop, int nhop,
                                hipStream_t st)
{
  if (nhop < 0 || nhop > 2 || (ctx->g.Lx & 1)) return hipErrorInvalidValue;
  const long V = (long)ctx->g.vcb * 2;
  const int bs = 128;
  WeightSrc src{};
  for (int c = 0; c < 2; c++) {
    src.chi_cb[c] = chi_cb[c];
    src.psi_cb[c] = psi_cb[c];
  }
  hipLaunchKernelGGL(clover_weights_k, dim3(blocks_for(V, bs)), dim3(bs), 0, st, K, src, ctx->g, csw, w_even,
                     w_odd, make_sigma());
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  LinkHop lh{};
  for (int i = 0; i < nhop; i++) lh.term[i] = hop[i];
  lh.n = nhop;
  lh.k = hop_unit_coef(ctx);
  return launch_clover_links<1>(ctx, force, K, U, scale, project, lh, st);
}

hipError_t run_clover_blocks(const qphix_amd_ctx *ctx, double *out, const double *clov, int mode,
                             double mu, int sign, hipStream_t st)
{
  const long nblk = (long)ctx->g.vcb * 2 * 2;
  const int bs = 64;
  hipLaunchKernelGGL(clover_blocks_k, dim3(blocks_for(nblk, bs)), dim3(bs), 0, st, out, clov, nblk,
                     mode, mu, sign);
  return hipGetLastError();
}

}  // namespace qamd
