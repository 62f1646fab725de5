// force.hip -- gradient of L = Re<chi, D psi> with respect to the gauge
// links (the outer-product fermion-force stencil).  No counterpart in the
// reference, whose kernels only apply D; D is the hop of the reference's
// body (codegen/dslash_common.cc:1729-1826) with chi on checkerboard cb and
// psi on 1-cb.
//
// Every link U_mu(y) enters exactly one hop:
//   y on cb:    forward hop at x = y,      chi(y)^+ (1 - s g_mu) ph U psi(y+mu)
//   y on 1-cb:  backward hop at x = y+mu,  chi(y+mu)^+ (1 + s g_mu) ph^* U^+ psi(y)
// so L contains Re tr(U G) with the 3x3 colour matrix
//   G = beta ph sum_spin psi(y+mu) (x) [(1 - s g)chi(y)]^+          (forward)
//   G = beta ph sum_spin [(1 + s g)chi(y+mu)] (x) psi(y)^+          (backward)
// (beta the hop coefficient with the time-boundary factor, ph the twisted-BC
// phase).  Output, accumulated with a scale into an FP64 lexicographic field
// [V][4][3][3][2]:
//   project = 1: F = TA(U G), the traceless anti-Hermitian su(3) force:
//                d/de L(exp(i e T) U) = Re tr(i T F) for Hermitian T;
//   project = 0: G^dagger, the gradient with respect to the real and
//                imaginary parts of the link entries.
#include "internal.h"

namespace qamd {

namespace {

constexpr int kForceBlock = 128;

template <typename R>
__device__ __forceinline__ void ld_spinor(cx<R> v[4][3], const R *r)
{
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int c = 0; c < 3; c++) v[s][c] = {r[(s * 3 + c) * 2], r[(s * 3 + c) * 2 + 1]};
}

// (1 + S g_MU) v, the full 4-spinor, DeGrand-Rossi basis
template <int MU, int S, typename R>
__device__ __forceinline__ void proj_full(const cx<R> v[4][3], cx<R> o[4][3])
{
#pragma unroll
  for (int c = 0; c < 3; c++) {
    cx<R> h0, h1;
    if (MU == 0) {
      h0 = S > 0 ? v[0][c] + timesI(v[3][c]) : v[0][c] - timesI(v[3][c]);
      h1 = S > 0 ? v[1][c] + timesI(v[2][c]) : v[1][c] - timesI(v[2][c]);
      o[2][c] = S > 0 ? cx<R>{R(0), R(0)} - timesI(h1) : timesI(h1);
      o[3][c] = S > 0 ? cx<R>{R(0), R(0)} - timesI(h0) : timesI(h0);
    } else if (MU == 1) {
      h0 = S > 0 ? v[0][c] - v[3][c] : v[0][c] + v[3][c];
      h1 = S > 0 ? v[1][c] + v[2][c] : v[1][c] - v[2][c];
      o[2][c] = S > 0 ? h1 : cx<R>{R(0), R(0)} - h1;
      o[3][c] = S > 0 ? cx<R>{R(0), R(0)} - h0 : h0;
    } else if (MU == 2) {
      h0 = S > 0 ? v[0][c] + timesI(v[2][c]) : v[0][c] - timesI(v[2][c]);
      h1 = S > 0 ? v[1][c] - timesI(v[3][c]) : v[1][c] + timesI(v[3][c]);
      o[2][c] = S > 0 ? cx<R>{R(0), R(0)} - timesI(h0) : timesI(h0);
      o[3][c] = S > 0 ? timesI(h1) : cx<R>{R(0), R(0)} - timesI(h1);
    } else {
      h0 = S > 0 ? v[0][c] + v[2][c] : v[0][c] - v[2][c];
      h1 = S > 0 ? v[1][c] + v[3][c] : v[1][c] - v[3][c];
      o[2][c] = S > 0 ? h0 : cx<R>{R(0), R(0)} - h0;
      o[3][c] = S > 0 ? h1 : cx<R>{R(0), R(0)} - h1;
    }
    o[0][c] = h0;
    o[1][c] = h1;
  }
}

struct ForceParams {
  double *force;
  const void *chi, *psi;
  const double *U;
  Geom g;
  int cb;  // checkerboard of chi (the output of D)
  int isign;
  double coeff_s, coeff_tf, coeff_tb, tbound, scale;
  int tbc_mask;
  double tbc_re[4], tbc_im[4];
  int project;
};

template <typename S, int MU, int ISIGN>
__device__ __forceinline__ void force_dir(const ForceParams &p, long y, const int c[4], int par)
{
  using R = typename Traits<S>::R;
  const Geom &g = p.g;
  const int L[4] = {g.Lx, g.Ly, g.Lz, g.Lt};
  int n[4] = {c[0], c[1], c[2], c[3]};
  n[MU] = n[MU] + 1 == L[MU] ? 0 : n[MU] + 1;
  const int cbi_y = (((c[3] * g.Lz + c[2]) * g.Ly + c[1]) * g.Lxh) + (c[0] >> 1);
  const int cbi_n = (((n[3] * g.Lz + n[2]) * g.Ly + n[1]) * g.Lxh) + (n[0] >> 1);
  const bool fwd = par == p.cb;  // y on chi's checkerboard: forward hop at y
  const S *chi = static_cast<const S *>(p.chi);
  const S *psi = static_cast<const S *>(p.psi);
  constexpr int WS = spinor_w<S>();
  R a[24], b[24];
  cx<R> va[4][3], vb[4][3], pr[4][3];
  double beta;
  if (MU < 3)
    beta = p.coeff_s;
  else if (fwd)
    beta = c[3] == g.Lt - 1 ? p.coeff_tf * p.tbound : p.coeff_tf;
  else
    beta = n[3] == 0 ? p.coeff_tb * p.tbound : p.coeff_tb;
  cx<double> G[3][3];
  if (fwd) {
    // G = sum_s psi(y+mu)_s (x) [(1 - isign g) chi(y)]_s^+
    load_reals<S, WS, 0, 24>(chi, g.vcb, cbi_y, a);
    load_reals<S, WS, 0, 24>(psi, g.vcb, cbi_n, b);
    ld_spinor(va, a);
    ld_spinor(vb, b);
    proj_full<MU, -ISIGN>(va, pr);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        cx<R> v = {R(0), R(0)};
#pragma unroll
        for (int s = 0; s < 4; s++) v = v + cmul_conj(pr[s][j], vb[s][i]);
        G[i][j] = {(double)v.re, (double)v.im};
      }
  } else {
    // G = sum_s [(1 + isign g) chi(y+mu)]_s (x) psi(y)_s^+
    load_reals<S, WS, 0, 24>(chi, g.vcb, cbi_n, a);
    load_reals<S, WS, 0, 24>(psi, g.vcb, cbi_y, b);
    ld_spinor(va, a);
    ld_spinor(vb, b);
    proj_full<MU, ISIGN>(va, pr);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        cx<R> v = {R(0), R(0)};
#pragma unroll
        for (int s = 0; s < 4; s++) v = v + cmul_conj(vb[s][j], pr[s][i]);
        G[i][j] = {(double)v.re, (double)v.im};
      }
  }
  cx<double> f = {beta, 0.0};
  if (p.tbc_mask & (1 << MU)) f = {beta * p.tbc_re[MU], beta * p.tbc_im[MU]};
  double *out = p.force + (y * 4 + MU) * 18;
  if (!p.project) {
    // G^dagger
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const cx<double> v = conj(f * G[j][i]);
        out[(i * 3 + j) * 2] += p.scale * v.re;
        out[(i * 3 + j) * 2 + 1] += p.scale * v.im;
      }
    return;
  }
  // W = U G, F = (W - W^+)/2 - tr(W - W^+)/6
  const double *u = p.U + (y * 4 + MU) * 18;
  cx<double> W[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      cx<double> v = {0.0, 0.0};
#pragma unroll
      for (int k = 0; k < 3; k++) v = v + cx<double>{u[(i * 3 + k) * 2], u[(i * 3 + k) * 2 + 1]} * G[k][j];
      W[i][j] = f * v;
    }
  const double trim = (W[0][0].im + W[1][1].im + W[2][2].im) / 3.0;
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      double re = 0.5 * (W[i][j].re - W[j][i].re);
      double im = 0.5 * (W[i][j].im + W[j][i].im) - (i == j ? trim : 0.0);
      out[(i * 3 + j) * 2] += p.scale * re;
      out[(i * 3 + j) * 2 + 1] += p.scale * im;
    }
}

template <typename S, int ISIGN>
__global__ __launch_bounds__(kForceBlock) void force_k(ForceParams p)
{
  const long V = (long)p.g.vcb * 2;
  const long y = (long)blockIdx.x * kForceBlock + threadIdx.x;
  if (y >= V) return;
  int c[4];
  long r = y;
  c[0] = (int)(r % p.g.Lx);
  r /= p.g.Lx;
  c[1] = (int)(r % p.g.Ly);
  r /= p.g.Ly;
  c[2] = (int)(r % p.g.Lz);
  c[3] = (int)(r / p.g.Lz);
  const int par = (c[0] + c[1] + c[2] + c[3]) & 1;
  force_dir<S, 0, ISIGN>(p, y, c, par);
  force_dir<S, 1, ISIGN>(p, y, c, par);
  force_dir<S, 2, ISIGN>(p, y, c, par);
  force_dir<S, 3, ISIGN>(p, y, c, par);
}

template <typename S>
hipError_t launch_force(const ForceParams &p, hipStream_t st)
{
  const long V = (long)p.g.vcb * 2;
  const int grid = blocks_for(V, kForceBlock);
  if (p.isign > 0)
    hipLaunchKernelGGL((force_k<S, 1>), dim3(grid), dim3(kForceBlock), 0, st, p);
  else
    hipLaunchKernelGGL((force_k<S, -1>), dim3(grid), dim3(kForceBlock), 0, st, p);
  return hipGetLastError();
}

}  // namespace

}  // namespace qamd

using namespace qamd;

extern "C" int qphix_amd_dslash_force(qphix_amd_ctx *ctx, int prec, int isign, int cb,
                                      double *force_lex, const void *chi, const void *psi,
                                      const double *u_lex, double coeff_s, double coeff_t_f,
                                      double coeff_t_b, double scale, int project, void *stream)
{
  if (!ctx || !force_lex || !chi || !psi || !u_lex) return QPHIX_AMD_EINVAL;
  if (ctx->comm_mask) return QPHIX_AMD_EINVAL;  // needs the whole lattice
  if ((isign != 1 && isign != -1) || (cb != 0 && cb != 1)) return QPHIX_AMD_EINVAL;
  if (prec != QPHIX_AMD_F64 && prec != QPHIX_AMD_F32) return QPHIX_AMD_EINVAL;
  ForceParams p{};
  p.force = force_lex;
  p.chi = chi;
  p.psi = psi;
  p.U = u_lex;
  p.g = ctx->g;
  p.cb = cb;
  p.isign = isign;
  p.coeff_s = coeff_s;
  p.coeff_tf = coeff_t_f;
  p.coeff_tb = coeff_t_b;
  p.tbound = ctx->tbound;
  p.scale = scale;
  p.tbc_mask = ctx->tbc_mask;
  for (int m = 0; m < 4; m++) {
    p.tbc_re[m] = ctx->tbc_re[m];
    p.tbc_im[m] = ctx->tbc_im[m];
  }
  p.project = project ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = prec == QPHIX_AMD_F64 ? launch_force<double>(p, st) : launch_force<float>(p, st);
  return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP;
}
