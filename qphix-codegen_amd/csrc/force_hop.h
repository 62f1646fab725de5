// force_hop.h -- the hopping term's contribution to the gradient of
// L = Re<chi, D psi> with respect to one link (used by the force kernels
// and, fused with the clover term, by the clover-force gather).  D is the
// hop of the reference's body (codegen/dslash_common.cc:1729-1826).
//
// Every link U_mu(y) enters exactly one hop:
//   y on cb:    forward hop at x = y,      chi(y)^+ (1 - s g_mu) ph U psi(y+mu)
//   y on 1-cb:  backward hop at x = y+mu,  chi(y+mu)^+ (1 + s g_mu) ph^* U^+ psi(y)
// so L contains Re tr(U G) with the 3x3 colour matrix
//   G = beta ph sum_spin psi(y+mu) (x) [(1 - s g)chi(y)]^+          (forward)
//   G = beta ph sum_spin [(1 + s g)chi(y+mu)] (x) psi(y)^+          (backward)
// (beta the hop coefficient with the time-boundary factor, ph the twisted-BC
// phase).
#pragma once

#include "internal.h"

namespace qamd {

// The upper two spin components of (1 + s g_mu) v for a run-time direction
// mu and sign s = +-1, DeGrand-Rossi basis:
//   h0 = v0 + c0 v[k0], h1 = v1 + c1 v[k1],
//   mu 0: k = (3, 2), c = (s i, s i)    mu 1: k = (3, 2), c = (-s, s)
//   mu 2: k = (2, 3), c = (s i, -s i)   mu 3: k = (2, 3), c = (s, s)
// (each part of c is 0 or +-1, so the products are exact).  The lower two
// components are unit multiples of these, so for P = 1 + s g (Hermitian,
// P^2 = 2P):
//   sum_spin conj((P v)_s,j) q_s,i = v^+ P q = sum_{k=0,1} conj(hv_k,j) hq_k,i
// with hv, hq the upper halves of P v and P q.
template <typename R>
__device__ __forceinline__ void proj_half(const R *v, int mu, R s, cx<R> h[2][3])
{
  const bool imag = mu == 0 || mu == 2;
  const bool k3 = mu < 2;  // k0 = 3, k1 = 2 (else k0 = 2, k1 = 3)
  const R c0 = mu == 1 ? -s : s, c1 = mu == 2 ? -s : s;
  auto at = [&](int sp, int c) { return cx<R>{v[(sp * 3 + c) * 2], v[(sp * 3 + c) * 2 + 1]}; };
  auto mulc = [&](R m, cx<R> a) { return imag ? cx<R>{m * -a.im, m * a.re} : cx<R>{m * a.re, m * a.im}; };
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cx<R> v2 = at(2, c), v3 = at(3, c);
    h[0][c] = at(0, c) + mulc(c0, k3 ? v3 : v2);
    h[1][c] = at(1, c) + mulc(c1, k3 ? v2 : v3);
  }
}

// Hop coefficients and twisted-boundary phases of the differentiated D.
struct HopCoef {
  double coeff_s, coeff_tf, coeff_tb, tbound;
  int tbc_mask;
  double tbc_re[4], tbc_im[4];
};

// One term w Re<chi, D psi>: chi (checkerboard cb) and psi (1 - cb) native
// spinors of storage S.
struct HopTerm {
  const void *chi, *psi;
  int cb, isign;
  double w;
};

// G += w beta ph G_hop for link U_mu(y); c: y's coordinates, cbi_y its
// checkerboard index.  Forward (y on chi's checkerboard): P = 1 - isign g,
// v = chi(y), q = psi(y+mu), G_hop = H; backward: P = 1 + isign g,
// v = chi(y+mu), q = psi(y), G_hop = H^dagger; H_ij = sum_spin
// conj((P v)_s,j) q_s,i.  Lanes of one wave may differ in checkerboard and
// in mu and still take the same path.
template <typename S>
__device__ __forceinline__ void hop_g_add(cx<double> G[3][3], const HopTerm &t, const HopCoef &k,
                                          const Geom &g, const int c[4], int cbi_y, int mu)
{
  using R = typename Traits<S>::R;
  const int L[4] = {g.Lx, g.Ly, g.Lz, g.Lt};
  int n[4];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    n[i] = c[i] + (mu == i);
    if (n[i] == L[i]) n[i] = 0;
  }
  const int cbi_n = (((n[3] * g.Lz + n[2]) * g.Ly + n[1]) * g.Lxh) + (n[0] >> 1);
  const bool fwd = ((c[0] + c[1] + c[2] + c[3]) & 1) == t.cb;
  const S *chi = static_cast<const S *>(t.chi);
  const S *psi = static_cast<const S *>(t.psi);
  double beta;
  if (mu < 3)
    beta = k.coeff_s;
  else if (fwd)
    beta = c[3] == g.Lt - 1 ? k.coeff_tf * k.tbound : k.coeff_tf;
  else
    beta = n[3] == 0 ? k.coeff_tb * k.tbound : k.coeff_tb;
  const R sg = fwd ? R(-t.isign) : R(t.isign);
  cx<R> hv[2][3], hq[2][3];
  {
    R v[24];
    load_reals<S, spinor_w<S>(), 0, 24>(chi, g.vcb, fwd ? cbi_y : cbi_n, v);
    proj_half(v, mu, sg, hv);
  }
  {
    R q[24];
    load_reals<S, spinor_w<S>(), 0, 24>(psi, g.vcb, fwd ? cbi_n : cbi_y, q);
    proj_half(q, mu, sg, hq);
  }
  cx<double> f = {t.w * beta, 0.0};
  if ((k.tbc_mask >> mu) & 1) {
    const double pr = mu == 0 ? k.tbc_re[0] : mu == 1 ? k.tbc_re[1] : mu == 2 ? k.tbc_re[2] : k.tbc_re[3];
    const double pi = mu == 0 ? k.tbc_im[0] : mu == 1 ? k.tbc_im[1] : mu == 2 ? k.tbc_im[2] : k.tbc_im[3];
    f = {t.w * beta * pr, t.w * beta * pi};
  }
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      // forward G_ij = H_ij, backward G_ij = conj(H_ji)
      const int a = fwd ? i : j, b = fwd ? j : i;
      const cx<R> v = cmul_conj(hv[0][b], hq[0][a]) + cmul_conj(hv[1][b], hq[1][a]);
      const cx<double> h = {(double)v.re, fwd ? (double)v.im : -(double)v.im};
      G[i][j] = G[i][j] + f * h;
    }
}

}  // namespace qamd
