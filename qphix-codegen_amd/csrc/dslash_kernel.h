// dslash_kernel.h -- the Wilson-type dslash family as one templated HIP
// kernel: one output site per lane, the 8 neighbour spinors gathered with
// coalesced site-fastest loads, SU(3) links rebuilt from 12 or 18 reals in
// registers, and the clover / twisted-mass / a*chi - b*D*psi epilogues
// fused before the single store.
//
// The arithmetic follows the reference's generated body
// (codegen/dslash_common.cc:1729-1932, direction loop codegen/dslash.cc:200-323):
//   per direction d = 2*mu + fb: project the neighbour 4-spinor to a
//   half-spinor with (1 + s*gamma_mu) (s = isign for backward hops, -isign for
//   forward hops; codegen/dslash_common.cc:748-766), multiply by the hop
//   matrix (U^dagger backward, U forward; codegen/dslash_common.cc:1678-1727),
//   apply the twisted-BC phase (codegen/dslash_common.cc:1491-1524) and the
//   direction coefficient, then reconstruct the lower two spins
//   (codegen/dslash_common.cc:784-890).
#pragma once

#include <type_traits>

#include "common.h"

namespace qamd {

template <typename R> struct DslashParams {
  void *out;
  const void *in;
  const void *chi;
  const void *gauge;
  const void *clov;
  Geom g;
  int cb;
  int nblocks;  // real (unpadded) work-groups
  int order;    // work-group traversal: 0 t-slabs per XCD (natural order),
                // 1 z-slabs per XCD visited t-major, 2 t-slabs visited z-major,
                // 3 t-slabs visited one z-part at a time
  int zs, ts, bpp;  // z-planes / t-slices per XCD slab, work-groups per (z,t) plane
  FastDiv f_per, f_zb, f_bpp;  // order 1: division by zs*Lt*bpp, zs*bpp and bpp
  int lds;          // stage the x/y neighbour rows in LDS (whole rows per work-group)
  R coeff_s, coeff_tf, coeff_tb, tbound;
  int comm;  // bit mu: dimension mu is split; hops across its wrap-around are dropped
  int tbc_mask;
  R tbc_re[4], tbc_im[4];
  R alpha, mu, muinv, rho;
  R one;  // 1, opaque to the compiler: the FP16 projection's FMA multiplier
  // when set: per work-group sums of |out|^2 over the stored values, at
  // norm[first site / kDslashBlock] (every work-group full)
  double *norm;
  // when set (with norm): out holds a CG residual r, and the launch stores
  // r - (*step) * result in its place; norm then sums |r|^2
  const double *step;
  // when set and *stop != 0 the launch does nothing (CG iterations queued
  // past a reliable update)
  const int *stop;
  // QPhiX FourSpinorBlock arrays (kernel template SPL = 1): soalen 1 << q_sh,
  // block strides q_nvecs (y), q_pxy (z), q_pxyz (t)
  int q_sh, q_nvecs, q_pxy, q_pxyz;
  // per-site direction mask (bit d = 2 mu + fb), QPhiX site order, or null
  const unsigned char *acc;
  // EPI_TRUE_RESIDUAL: the right-hand side b of out = b - (*step) * result
  const void *rin;
};

// What a launch does with its result at a site: store it; store it and sum
// |out|^2 per work-group; subtract it, times the device-resident CG step
// length, from the residual held in out, and sum |r|^2 per work-group; or
// (FP64 only, the CG's true residual b - M^dag M x) store b - step * result
// with b read from rin, and sum its |.|^2.
enum { EPI_STORE = 0, EPI_NORM = 1, EPI_RESIDUAL = 2, EPI_TRUE_RESIDUAL = 3 };

#ifndef QAMD_XCD_REMAP
#define QAMD_XCD_REMAP 1
#endif
// default work-group traversal (see DslashParams::order); -1 picks it from
// the geometry
#ifndef QAMD_BLOCK_ORDER
#define QAMD_BLOCK_ORDER -1
#endif
// z-parts per t-slab for order 3
#ifndef QAMD_ZPARTS
#define QAMD_ZPARTS 2
#endif
#ifndef QAMD_DSLASH_BLOCK
#define QAMD_DSLASH_BLOCK 128
#endif
constexpr int kDslashBlock = QAMD_DSLASH_BLOCK;

// Work-group traversal order (DslashParams::order) for this geometry: the
// requested one (qphix_amd_set_traversal, else QAMD_BLOCK_ORDER) when the
// geometry admits it, else 0.  Automatic choice: 1 when one t-slice holds
// more than 32768 sites per checkerboard or Lt <= Lz, else 3 (0 where the
// geometry does not admit 3).
// * With order 0 every XCD sweeps its own t-slab.  Between the first and the
//   last read of an input t-slice the eight XCDs then stream two slices each
//   of every field: ~230 MB at 32^3 slices (FP32 clover), within the 256 MB
//   Infinity Cache, but ~760 MB at 48^3 slices, where the input re-reads go
//   to HBM.  Thin slabs (Lt/8 <= Lz/8 slices) also put most slices next to
//   another XCD's slab.
// * Order 1 has the XCDs sweep t together, each owning a z-range.
// * Order 3 is order 0 with each XCD's slab swept one z-part at a time, so
//   the three input t-slices that meet in its L2 are half as large.
// Measured FP32 clov_dslash, order 0 -> 1: 48^3x96 915 -> 855 us, 32^3x32
// 82.7 -> 82.0, 32^3x16 46.6 -> 42.7, 32^3x8 24.3 -> 22.6, 32x32x64x64
// 340 -> 334; not taken: 32^3x64 152 -> 166, 32^3x128 331 -> 334,
// 16^3x32 and 64x32x32x64 equal.  At 32^3x64, order 0 -> 3 (rocprofv3):
// clov_dslash 155.0 -> 152.8 us, its achimbdpsi 166.4 -> 165.1, Wilson
// 108.6 -> 107.9, with 3-5% fewer L2 misses (2, 4 or 8 z-parts equal).
inline int dslash_traversal_for(const Geom &g, int requested)
{
  if (g.sz % kDslashBlock != 0) return 0;
  int ord = requested;
  if (ord < 0) ord = ((long)g.sz * g.Lz > 32768 || g.Lt <= g.Lz) ? 1 : 3;
  if (ord == 1 && g.Lz % kNumXcd == 0) return 1;
  if (ord == 2 && g.Lt % kNumXcd == 0) return 2;
  if (ord == 3 && g.Lt % kNumXcd == 0 && g.Lz % QAMD_ZPARTS == 0) return 3;
  return 0;
}

// Minimum waves per SIMD requested from the register allocator (0 = none).
#ifndef QAMD_DSLASH_MIN_WAVES
#define QAMD_DSLASH_MIN_WAVES 0
#endif

// Half-spinor projection (1 + SGN*gamma_MU) psi -> (h0, h1), DeGrand-Rossi.
template <int MU, int SGN, typename R>
__device__ __forceinline__ void project(const R *psi, cx<R> h[2][3])
{
#pragma unroll
  for (int c = 0; c < 3; c++) {
    cx<R> p0 = {psi[(0 * 3 + c) * 2], psi[(0 * 3 + c) * 2 + 1]};
    cx<R> p1 = {psi[(1 * 3 + c) * 2], psi[(1 * 3 + c) * 2 + 1]};
    cx<R> p2 = {psi[(2 * 3 + c) * 2], psi[(2 * 3 + c) * 2 + 1]};
    cx<R> p3 = {psi[(3 * 3 + c) * 2], psi[(3 * 3 + c) * 2 + 1]};
    if (MU == 0) {  // h0 = p0 + s i p3, h1 = p1 + s i p2
      h[0][c] = SGN > 0 ? p0 + timesI(p3) : p0 - timesI(p3);
      h[1][c] = SGN > 0 ? p1 + timesI(p2) : p1 - timesI(p2);
    } else if (MU == 1) {  // h0 = p0 - s p3, h1 = p1 + s p2
      h[0][c] = SGN > 0 ? p0 - p3 : p0 + p3;
      h[1][c] = SGN > 0 ? p1 + p2 : p1 - p2;
    } else if (MU == 2) {  // h0 = p0 + s i p2, h1 = p1 - s i p3
      h[0][c] = SGN > 0 ? p0 + timesI(p2) : p0 - timesI(p2);
      h[1][c] = SGN > 0 ? p1 - timesI(p3) : p1 + timesI(p3);
    } else {  // h0 = p0 + s p2, h1 = p1 + s p3
      h[0][c] = SGN > 0 ? p0 + p2 : p0 - p2;
      h[1][c] = SGN > 0 ? p1 + p3 : p1 - p3;
    }
  }
}

// Accumulate b times the reconstructed 4-spinor of (1 + SGN*gamma_MU) from
// w = U h, b folded into fused multiply-adds.
template <int MU, int SGN, typename R>
__device__ __forceinline__ void reconstruct(cx<R> acc[4][3], const cx<R> w[2][3], R b)
{
  const R sb = SGN > 0 ? b : -b;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    acc[0][c] = {fma(b, w[0][c].re, acc[0][c].re), fma(b, w[0][c].im, acc[0][c].im)};
    acc[1][c] = {fma(b, w[1][c].re, acc[1][c].re), fma(b, w[1][c].im, acc[1][c].im)};
    if (MU == 0) {  // r2 = -s i w1, r3 = -s i w0
      acc[2][c] = {fma(sb, w[1][c].im, acc[2][c].re), fma(-sb, w[1][c].re, acc[2][c].im)};
      acc[3][c] = {fma(sb, w[0][c].im, acc[3][c].re), fma(-sb, w[0][c].re, acc[3][c].im)};
    } else if (MU == 1) {  // r2 = s w1, r3 = -s w0
      acc[2][c] = {fma(sb, w[1][c].re, acc[2][c].re), fma(sb, w[1][c].im, acc[2][c].im)};
      acc[3][c] = {fma(-sb, w[0][c].re, acc[3][c].re), fma(-sb, w[0][c].im, acc[3][c].im)};
    } else if (MU == 2) {  // r2 = -s i w0, r3 = s i w1
      acc[2][c] = {fma(sb, w[0][c].im, acc[2][c].re), fma(-sb, w[0][c].re, acc[2][c].im)};
      acc[3][c] = {fma(-sb, w[1][c].im, acc[3][c].re), fma(sb, w[1][c].re, acc[3][c].im)};
    } else {  // r2 = s w0, r3 = s w1
      acc[2][c] = {fma(sb, w[0][c].re, acc[2][c].re), fma(sb, w[0][c].im, acc[2][c].im)};
      acc[3][c] = {fma(sb, w[1][c].re, acc[3][c].re), fma(sb, w[1][c].im, acc[3][c].im)};
    }
  }
}

// Load the two hop matrices of direction MU at a site (d = 2 MU back,
// 2 MU + 1 forward).  Each is stored as the rows of X, X = U_mu(x - mu)^dagger
// for the back hop and X = U_mu(x)^T for the forward hop, so that the kept
// rows are the reference's stored rows (it stores the transpose of each link,
// codegen/dslash_common.cc:1678-1727): rows 0,1 stored, row 2 stored (18) or
// rebuilt as conj(row0 x row1) (12; codegen/dslash_common.cc:1664-1674).
// Returns the applied matrices: ub = X_back, uf = X_fwd^T.  FP16 links: nz
// (an opaque -0.0) starts the product chains of the rebuilt row.
template <typename S, int RECON, int MU, typename R>
__device__ __forceinline__ void load_link_pair(const S *gauge, int vcb, int site, cx<R> ub[3][3],
                                               cx<R> uf[3][3], R nz = R(-0.0))
{
  constexpr int WL = gauge_w<S, RECON>();
  R lk[2 * RECON];
  load_reals<S, WL, 2 * MU * RECON, 2 * RECON, R, true>(gauge, vcb, site, lk);
#pragma unroll
  for (int fb = 0; fb < 2; fb++) {
    cx<R> x[3][3];
    const R *l = lk + fb * RECON;
#pragma unroll
    for (int r = 0; r < RECON / 6; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) x[r][c] = {l[(r * 3 + c) * 2], l[(r * 3 + c) * 2 + 1]};
    if (RECON == 12) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        if constexpr (std::is_same<S, __half>::value)
          x[2][c] = conj_diff(x[0][c1], x[1][c2], x[0][c2], x[1][c1], nz);
        else
          x[2][c] = conj_diff(x[0][c1], x[1][c2], x[0][c2], x[1][c1]);
      }
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (fb)
          uf[r][c] = x[c][r];
        else
          ub[r][c] = x[r][c];
      }
  }
}

// Input spinors staged in LDS (QAMD_LDS_XY): [16-byte chunk][staged site].
// A work-group (128 consecutive sites) stages the x rows it touches and the
// row on either side of them once; the x and y hops (4 of the 8) then read
// LDS instead of issuing global loads that mostly hit L2.  The barrier also splits the
// kernel into phases, which lowers register pressure (FP32 166 -> 128 VGPRs,
// occupancy 3 -> 4 waves/SIMD): clov_dslash FP32 155 -> 152 us, FP16
// 81 -> 77 us at 32^3x64.
#ifndef QAMD_LDS_XY
#define QAMD_LDS_XY 1
#endif
// staged sites per work-group: 8 rows of 16 plus 2, 4 of 32 plus 2, up to 6
// partial rows of 24 plus 2, or up to 8 of 20 plus 2 (the host checks the
// geometry's worst case).  FP32: 19,968 B, so 8 work-groups (4 waves per
// SIMD) still fit a CU's 160 KB.
constexpr int kLdsSites = kDslashBlock + 80;
using Vec16 = typename RawVec<16>::T;

// A neighbour spinor in registers.  FP16 storage keeps the raw halves: its
// projection reads them directly with mixed-precision FMAs (one v_fma_mix
// per projected real instead of two conversions and an add; the FP16
// operator kernels are close to the VALU issue limit).  Every other storage
// type is converted to the arithmetic type on load.
template <typename S, typename R> struct NbrSpinor {
  R v[24];
};
template <> struct NbrSpinor<__half, float> {
  __half v[24];
};

template <typename S, typename R>
__device__ __forceinline__ void load_nbr(const S *__restrict__ in, int vcb, int site, NbrSpinor<S, R> &n)
{
  if constexpr (std::is_same<S, __half>::value) {
    constexpr int W = spinor_w<S>();
    using V = typename RawVec<W * sizeof(S)>::T;
    const V *vb = reinterpret_cast<const V *>(in);
#pragma unroll
    for (int c = 0; c < 24 / W; c++) {
      V v = vb[(unsigned)(c * vcb + site)];
      __builtin_memcpy(n.v + c * W, &v, sizeof(V));
    }
  } else {
    load_reals<S, spinor_w<S>(), 0, 24>(in, vcb, site, n.v);
  }
}

template <typename S, typename R>
__device__ __forceinline__ void load_nbr_lds(const Vec16 *lds, int j, NbrSpinor<S, R> &n)
{
  constexpr int W = 16 / (int)sizeof(S);
#pragma unroll
  for (int c = 0; c < 24 / W; c++) {
    Vec16 v = lds[c * kLdsSites + j];
    if constexpr (std::is_same<S, __half>::value) {
      __builtin_memcpy(n.v + c * W, &v, 16);
    } else {
      S tmp[W];
      __builtin_memcpy(tmp, &v, 16);
#pragma unroll
      for (int e = 0; e < W; e++) n.v[c * W + e] = (R)to_real(tmp[e]);
    }
  }
}

// ---- QPhiX FourSpinorBlock arrays, FT[3 colour][4 spin][2][soalen] -------
// (codegen/include/decl_common.h:11; block t pxyz + z pxy + y nvecs +
// xh / soalen, lane xh % soalen).  Element (spin s, colour c, re/im r) of the
// site whose element 0 sits at q is at q + ((c * 4 + s) * 2 + r) * soalen: a
// wave's load of one element reads soalen consecutive values per block.
template <typename R>
__device__ __forceinline__ int qphix_site(const DslashParams<R> &p, int xh, int y, int z, int t)
{
  const int blk = t * p.q_pxyz + z * p.q_pxy + y * p.q_nvecs + (xh >> p.q_sh);
  return ((blk * 24) << p.q_sh) + (xh & ((1 << p.q_sh) - 1));
}

template <typename S, typename R>
__device__ __forceinline__ void load_nbr_q(const S *__restrict__ in, int q, int sh, NbrSpinor<S, R> &n)
{
#pragma unroll
  for (int sp = 0; sp < 4; sp++)
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int r = 0; r < 2; r++) {
        const S v = in[q + (((c * 4 + sp) * 2 + r) << sh)];
        if constexpr (std::is_same<S, __half>::value)
          n.v[(sp * 3 + c) * 2 + r] = v;
        else
          n.v[(sp * 3 + c) * 2 + r] = (R)to_real(v);
      }
}

// Non-temporal scalar access; 16-bit storage types go through their bits.
template <typename S> __device__ __forceinline__ S nt_load(const S *a)
{
  if constexpr (sizeof(S) == 2) {
    const unsigned short u = __builtin_nontemporal_load(reinterpret_cast<const unsigned short *>(a));
    S v;
    __builtin_memcpy(&v, &u, 2);
    return v;
  } else {
    return __builtin_nontemporal_load(a);
  }
}

template <typename S> __device__ __forceinline__ void nt_store(S v, S *a)
{
  if constexpr (sizeof(S) == 2) {
    unsigned short u;
    __builtin_memcpy(&u, &v, 2);
    __builtin_nontemporal_store(u, reinterpret_cast<unsigned short *>(a));
  } else {
    __builtin_nontemporal_store(v, a);
  }
}

template <typename S, typename R, bool NT = false>
__device__ __forceinline__ void load_q_reals(const S *__restrict__ base, int q, int sh, R *out)
{
#pragma unroll
  for (int sp = 0; sp < 4; sp++)
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int r = 0; r < 2; r++) {
        const S *a = base + q + (((c * 4 + sp) * 2 + r) << sh);
        out[(sp * 3 + c) * 2 + r] = (R)to_real((NT && QAMD_NT_STREAM) ? nt_load(a) : *a);
      }
}

template <typename S, typename R>
__device__ __forceinline__ void store_q_reals(S *__restrict__ base, int q, int sh, const R *in)
{
#pragma unroll
  for (int sp = 0; sp < 4; sp++)
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int r = 0; r < 2; r++) {
        S *a = base + q + (((c * 4 + sp) * 2 + r) << sh);
        const S v = to_storage<S>(in[(sp * 3 + c) * 2 + r]);
        if (QAMD_NT_STREAM)
          nt_store(v, a);
        else
          *a = v;
      }
}

// ---- QPhiX arrays read and written as vectors (SPL = 2) -------------------
// FP32 or FP16 with soalen >= 4: the four lanes of a quad hold consecutive
// sites of one block, whose element e sits in one 16-byte (FP16: 8-byte)
// vector at element-0 offset qf + e * soalen (qf: the quad's first site).
// (FP64 reads its spinors element by element, 8 bytes per lane, and uses
// 16-byte vectors of lane pairs only to stage the x/y rows in LDS: the
// 64-bit transposes cost more than they save at one wave per SIMD.)  Lane j of the group loads the vectors of elements e = j (mod 4)
// (2 for FP64), so that it holds element e of all the group's sites, and a
// transpose inside the group by DPP lane permutes gives every lane the
// elements of its own site: 6 (12) vector loads instead of 24 scalar ones.
// The host picks SPL = 2 only when every quad (pair) is one block's
// consecutive sites (Lx/2 a multiple of 4 (2)) and the arrays are 16-byte
// aligned; every lane of a quad then runs the same code.
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
constexpr int kDppSwapPairs = 0xB1;  // quad_perm [1, 0, 3, 2]
constexpr int kDppSwapHalves = 0x4E;  // quad_perm [2, 3, 0, 1]

// 4 x 4 transpose across the lanes of a quad: lane j's a[k] becomes lane
// k's a[j]
__device__ __forceinline__ void quad_transpose(float a[4])
{
  const int j = (int)threadIdx.x & 3;
  const bool odd = j & 1, hi = j & 2;
  const float r0 = dpp_f<kDppSwapPairs>(odd ? a[0] : a[1]);
  if (odd) a[0] = r0; else a[1] = r0;
  const float r1 = dpp_f<kDppSwapPairs>(odd ? a[2] : a[3]);
  if (odd) a[2] = r1; else a[3] = r1;
  const float q0 = dpp_f<kDppSwapHalves>(hi ? a[0] : a[2]);
  if (hi) a[0] = q0; else a[2] = q0;
  const float q1 = dpp_f<kDppSwapHalves>(hi ? a[1] : a[3]);
  if (hi) a[1] = q1; else a[3] = q1;
}

// 4 x 4 transpose of 16-bit entries across a quad: lane j's words
// w[0] = [a_j0 | a_j1], w[1] = [a_j2 | a_j3] (low half first) become lane
// k's [a_0k | a_1k], [a_2k | a_3k].  Within lane pairs, even lanes keep the
// low halves and odd lanes the high halves of their own and their partner's
// word, one v_perm_b32 with a per-lane byte select (bytes 0-3 of its second
// operand p, 4-7 of its first w: even [w.0 w.1 p.0 p.1], odd
// [p.2 p.3 w.2 w.3]); then the off-diagonal words are swapped between pairs.
template <int CTRL>
__device__ __forceinline__ unsigned dpp_u(unsigned v)
{
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ void quad_transpose_h(unsigned w[2])
{
  const int j = (int)threadIdx.x & 3;
  const bool odd = j & 1, hi = j & 2;
  const unsigned sel = odd ? 0x07060302u : 0x01000504u;
#pragma unroll
  for (int k = 0; k < 2; k++) w[k] = __builtin_amdgcn_perm(w[k], dpp_u<kDppSwapPairs>(w[k]), sel);
  const unsigned r = dpp_u<kDppSwapHalves>(hi ? w[0] : w[1]);
  if (hi) w[0] = r; else w[1] = r;
}

// The site's 24 elements in QPhiX order E = (c * 4 + s) * 2 + r, from the
// arrays at the group's first-site offset qf (NT: non-temporal loads)
template <typename S, bool NT>
__device__ __forceinline__ void load_qv_elems(const S *__restrict__ base, int qf, int sh, S out[24])
{
  if constexpr (std::is_same<S, float>::value) {
    const int j = (int)threadIdx.x & 3;
#pragma unroll
    for (int m = 0; m < 6; m++) {
      const u32x4 *a = reinterpret_cast<const u32x4 *>(base + qf + ((4 * m + j) << sh));
      const u32x4 v = (NT && QAMD_NT_STREAM) ? __builtin_nontemporal_load(a) : *a;
      float t[4];
      __builtin_memcpy(t, &v, 16);
      quad_transpose(t);
#pragma unroll
      for (int k = 0; k < 4; k++) out[4 * m + k] = t[k];
    }
  } else if constexpr (std::is_same<S, __half>::value) {
    // 8-byte vectors: element e of the quad's four sites
    const int j = (int)threadIdx.x & 3;
#pragma unroll
    for (int m = 0; m < 6; m++) {
      const u32x2 *a = reinterpret_cast<const u32x2 *>(base + qf + ((4 * m + j) << sh));
      const u32x2 v = (NT && QAMD_NT_STREAM) ? __builtin_nontemporal_load(a) : *a;
      unsigned w[2] = {v.x, v.y};
      quad_transpose_h(w);
      __builtin_memcpy(out + 4 * m, w, 8);
    }
  } else {
    static_assert(std::is_same<S, __half>::value, "vector QPhiX access: FP32 or FP16");
  }
}

// element E of QPhiX order -> index in the (spin, colour, re/im) order
__host__ __device__ constexpr int qphix_to_native(int E) { return (((E >> 1) & 3) * 3 + (E >> 3)) * 2 + (E & 1); }

template <typename S, typename R>
__device__ __forceinline__ void load_nbr_qv(const S *__restrict__ in, int qf, int sh, NbrSpinor<S, R> &n)
{
  if constexpr (std::is_same<S, double>::value) {  // FP64: element by element
    load_nbr_q(in, qf + ((int)threadIdx.x & 1), sh, n);
  } else {
    S e[24];
    load_qv_elems<S, false>(in, qf, sh, e);
#pragma unroll
    for (int E = 0; E < 24; E++) {
      if constexpr (std::is_same<S, __half>::value)
        n.v[qphix_to_native(E)] = e[E];  // raw halves, as load_nbr keeps them
      else
        n.v[qphix_to_native(E)] = (R)e[E];
    }
  }
}

template <typename S, typename R>
__device__ __forceinline__ void load_qv_reals(const S *__restrict__ base, int qf, int sh, R *out)
{
  if constexpr (std::is_same<S, double>::value) {  // FP64: element by element
    load_q_reals<S, R, true>(base, qf + ((int)threadIdx.x & 1), sh, out);
  } else {
    S e[24];
    load_qv_elems<S, true>(base, qf, sh, e);
#pragma unroll
    for (int E = 0; E < 24; E++) out[qphix_to_native(E)] = (R)to_real(e[E]);
  }
}

template <typename S, typename R>
__device__ __forceinline__ void store_qv_reals(S *__restrict__ base, int qf, int sh, const R *in)
{
  if constexpr (std::is_same<S, double>::value) {  // FP64: element by element
    store_q_reals<S, R>(base, qf + ((int)threadIdx.x & 1), sh, in);
  } else if constexpr (std::is_same<S, float>::value) {
    const int j = (int)threadIdx.x & 3;
#pragma unroll
    for (int m = 0; m < 6; m++) {
      float t[4];
#pragma unroll
      for (int k = 0; k < 4; k++) t[k] = (float)in[qphix_to_native(4 * m + k)];
      quad_transpose(t);
      u32x4 *a = reinterpret_cast<u32x4 *>(base + qf + ((4 * m + j) << sh));
      u32x4 v;
      __builtin_memcpy(&v, t, 16);
      if (QAMD_NT_STREAM)
        __builtin_nontemporal_store(v, a);
      else
        *a = v;
    }
  } else if constexpr (std::is_same<S, __half>::value) {
    const int j = (int)threadIdx.x & 3;
#pragma unroll
    for (int m = 0; m < 6; m++) {
      __half h[4];
#pragma unroll
      for (int k = 0; k < 4; k++) h[k] = to_storage<__half>(in[qphix_to_native(4 * m + k)]);
      unsigned w[2];
      __builtin_memcpy(w, h, 8);
      quad_transpose_h(w);
      u32x2 *a = reinterpret_cast<u32x2 *>(base + qf + ((4 * m + j) << sh));
      u32x2 v;
      v.x = w[0];
      v.y = w[1];
      if (QAMD_NT_STREAM)
        __builtin_nontemporal_store(v, a);
      else
        *a = v;
    }
  } else {
    static_assert(std::is_same<S, __half>::value, "vector QPhiX access: FP32 or FP16");
  }
}

// (1 + SGN*gamma_MU) psi from raw FP16 components: every projected real is
// a + s b = fma(b, s, a) with s = +-one, exact products and one rounding,
// the same value as the FP32 add of the converted components.
template <int MU, int SGN>
__device__ __forceinline__ void project_mixed(const __half *p, float one, cx<float> h[2][3])
{
  const float s = SGN > 0 ? one : -one;
  auto f = [](__half v) { return __half2float(v); };
  auto re = [&](int sp, int c) { return f(p[(sp * 3 + c) * 2]); };
  auto im = [&](int sp, int c) { return f(p[(sp * 3 + c) * 2 + 1]); };
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (MU == 0) {  // h0 = p0 + s i p3, h1 = p1 + s i p2
      h[0][c] = {fmaf(im(3, c), -s, re(0, c)), fmaf(re(3, c), s, im(0, c))};
      h[1][c] = {fmaf(im(2, c), -s, re(1, c)), fmaf(re(2, c), s, im(1, c))};
    } else if (MU == 1) {  // h0 = p0 - s p3, h1 = p1 + s p2
      h[0][c] = {fmaf(re(3, c), -s, re(0, c)), fmaf(im(3, c), -s, im(0, c))};
      h[1][c] = {fmaf(re(2, c), s, re(1, c)), fmaf(im(2, c), s, im(1, c))};
    } else if (MU == 2) {  // h0 = p0 + s i p2, h1 = p1 - s i p3
      h[0][c] = {fmaf(im(2, c), -s, re(0, c)), fmaf(re(2, c), s, im(0, c))};
      h[1][c] = {fmaf(im(3, c), s, re(1, c)), fmaf(re(3, c), -s, im(1, c))};
    } else {  // h0 = p0 + s p2, h1 = p1 + s p3
      h[0][c] = {fmaf(re(2, c), s, re(0, c)), fmaf(im(2, c), s, im(0, c))};
      h[1][c] = {fmaf(re(3, c), s, re(1, c)), fmaf(im(3, c), s, im(1, c))};
    }
  }
}

template <typename S, int MU, int FB, int ISIGN, typename R>
__device__ __forceinline__ void hop_apply(cx<R> acc[4][3], const NbrSpinor<S, R> &n, const cx<R> u[3][3],
                                          R beta, bool tbc, cx<R> phase, R one)
{
  constexpr int SGN = FB == 0 ? ISIGN : -ISIGN;
  cx<R> h[2][3];
  if constexpr (std::is_same<S, __half>::value)
    project_mixed<MU, SGN>(n.v, one, h);
  else
    project<MU, SGN>(n.v, h);
  cx<R> w[2][3];
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      if constexpr (std::is_same<S, __half>::value)  // FP16 links read by the FMAs
        w[k][i] = cdot3(u[i][0], u[i][1], u[i][2], h[k][0], h[k][1], h[k][2], -R(0) * one);
      else
        w[k][i] = cdot3(u[i][0], u[i][1], u[i][2], h[k][0], h[k][1], h[k][2]);
    }
  if (tbc) {
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
      for (int i = 0; i < 3; i++) w[k][i] = cmul_ord(phase, w[k][i]);
  }
  reconstruct<MU, SGN>(acc, w, beta);
}

// nb, nf: the neighbours' native site indices, or (SPL) their element-0
// offsets in a QPhiX FourSpinorBlock array of soalen 1 << sh (SPL = 2 and
// MU > 0: those of the first site of the neighbours' quad or pair)
template <typename S, int RECON, int MU, int ISIGN, int SPL = 0, typename R>
__device__ __forceinline__ void hop_mu(cx<R> acc[4][3], const S *__restrict__ in,
                                       const S *__restrict__ gauge, int vcb, int site, int nb,
                                       int nf, R bb, R bf, bool tbc, cx<R> phase, R one, int sh = 0)
{
  cx<R> ub[3][3], uf[3][3];
  load_link_pair<S, RECON, MU>(gauge, vcb, site, ub, uf, -R(0) * one);
  NbrSpinor<S, R> n;
  if constexpr (SPL == 2 && MU > 0)
    load_nbr_qv(in, nb, sh, n);
  else if constexpr (SPL)
    load_nbr_q(in, nb, sh, n);
  else
    load_nbr(in, vcb, nb, n);
  hop_apply<S, MU, 0, ISIGN>(acc, n, ub, bb, tbc, conj(phase), one);
  if constexpr (SPL == 2 && MU > 0)
    load_nbr_qv(in, nf, sh, n);
  else if constexpr (SPL)
    load_nbr_q(in, nf, sh, n);
  else
    load_nbr(in, vcb, nf, n);
  hop_apply<S, MU, 1, ISIGN>(acc, n, uf, bf, tbc, phase, one);
}

// the same with both neighbours read from the staged LDS rows
template <typename S, int RECON, int MU, int ISIGN, typename R>
__device__ __forceinline__ void hop_mu_lds(cx<R> acc[4][3], const Vec16 *lds, const S *__restrict__ gauge,
                                           int vcb, int site, int jb, int jf, R bb, R bf, bool tbc,
                                           cx<R> phase, R one)
{
  cx<R> ub[3][3], uf[3][3];
  load_link_pair<S, RECON, MU>(gauge, vcb, site, ub, uf, -R(0) * one);
  NbrSpinor<S, R> n;
  load_nbr_lds<S>(lds, jb, n);
  hop_apply<S, MU, 0, ISIGN>(acc, n, ub, bb, tbc, conj(phase), one);
  load_nbr_lds<S>(lds, jf, n);
  hop_apply<S, MU, 1, ISIGN>(acc, n, uf, bf, tbc, phase, one);
}

// out = A in, A = two packed Hermitian 6x6 blocks
// (codegen/dslash_common.cc:1098-1186); ACC: out += A in, each row's chain
// starting from out (the face form, codegen/dslash_common.cc:1125-1131).
template <typename S, bool ACC = false, typename R>
__device__ __forceinline__ void clover_packed(const S *clov, int vcb, int site,
                                              const cx<R> in[4][3], cx<R> out[4][3])
{
  R all[72];
  load_reals<S, clover_w<S>(), 0, 72, R, true>(clov, vcb, site, all);
#pragma unroll
  for (int b = 0; b < 2; b++) {
    const R *blk = all + 36 * b;
    cx<R> v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = in[2 * b + k / 3][k % 3];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      cx<R> res;
      if constexpr (ACC) {
        const cx<R> o = out[2 * b + r / 3][r % 3];
        res = {fmx(blk[r], v[r].re, o.re), fmx(blk[r], v[r].im, o.im)};
      } else {
        res = blk[r] * v[r];
      }
#pragma unroll
      for (int c = 0; c < 6; c++) {
        if (c < r) {
          const int od = r * (r - 1) / 2 + c;
          cx<R> a = {blk[6 + 2 * od], blk[6 + 2 * od + 1]};
          res = cmac(res, a, v[c]);
        } else if (c > r) {
          const int od = c * (c - 1) / 2 + r;
          cx<R> a = {blk[6 + 2 * od], blk[6 + 2 * od + 1]};
          res = cmac_conj(res, a, v[c]);
        }
      }
      out[2 * b + r / 3][r % 3] = res;
    }
  }
}

// out = B in, B = two full complex 6x6 blocks, row-major
// (codegen/dslash_common.cc:1188-1232); ACC: out += B in, each row's chain
// starting from out (the face form).
template <typename S, bool ACC = false, typename R>
__device__ __forceinline__ void clover_full(const S *clov, int vcb, int site,
                                            const cx<R> in[4][3], cx<R> out[4][3])
{
  constexpr int WF = fclover_w<S>();
#pragma unroll
  for (int b = 0; b < 2; b++) {
    R blk[72];
    if (b == 0)
      load_reals<S, WF, 0, 72, R, true>(clov, vcb, site, blk);
    else
      load_reals<S, WF, 72, 72, R, true>(clov, vcb, site, blk);
    cx<R> v[6];
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = in[2 * b + k / 3][k % 3];
#pragma unroll
    for (int r = 0; r < 6; r++) {
      cx<R> a0 = {blk[(r * 6) * 2], blk[(r * 6) * 2 + 1]};
      cx<R> res;
      if constexpr (ACC)
        res = cmac(out[2 * b + r / 3][r % 3], a0, v[0]);
      else
        res = cmul_ord(a0, v[0]);
#pragma unroll
      for (int c = 1; c < 6; c++) {
        cx<R> a = {blk[(r * 6 + c) * 2], blk[(r * 6 + c) * 2 + 1]};
        res = cmac(res, a, v[c]);
      }
      out[2 * b + r / 3][r % 3] = res;
    }
  }
}

// a (1 + i m g5) in, m real (g5 = +1 on spins 0,1, -1 on 2,3), in the
// reference's order: (1 + i m g5) in by one fused multiply-add per part,
// then the product by a (codegen/dslash_common.cc:1234-1489).  The
// twisted-mass dslash applies muinv (1 - isign i mu g5), its achimbdpsi form
// alpha (1 + isign i mu g5).
template <typename R>
__device__ __forceinline__ void twist_ord(const cx<R> in[4][3], cx<R> out[4][3], R a, R m)
{
#pragma unroll
  for (int s = 0; s < 4; s++) {
    const R ms = s < 2 ? m : -m;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const cx<R> v = in[s][c];
      out[s][c] = {a * fma(-ms, v.im, v.re), a * fma(ms, v.re, v.im)};
    }
  }
}

#if QAMD_DSLASH_MIN_WAVES > 0
#define QAMD_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(QAMD_DSLASH_MIN_WAVES)))
#else
#define QAMD_WAVES_ATTR
#endif

// EPI (EPI_STORE, EPI_NORM, EPI_RESIDUAL) is instantiated beyond EPI_STORE
// for the clover and twisted-mass-clover achimbdpsi kernels, the second half
// of the even-odd operator M: EPI_NORM on M p gives the CG's
// <p, M^dag M p> = |M p|^2, EPI_RESIDUAL on M^dag (M p) its residual update.
// SPL = 1 reads in and chi and writes out as QPhiX FourSpinorBlock arrays
// (the links and clover blocks stay native), with the optional per-site
// direction mask p.acc: a QPhiX host's arrays used in place, the reference's
// body with its accumulate[] lane masks (codegen/dslash.cc:248,285-290).
// SPL = 2 is the same with the spinors read and written as vectors by lane
// group (FP64, FP32, FP16; the x/y rows staged in LDS from those vectors, or
// else the x-neighbours, not aligned to a lane group, element by element).
template <typename S, int RECON, int OP, bool ACHI, int ISIGN, int EPI = EPI_STORE, int SPL = 0>
__global__ __launch_bounds__(kDslashBlock) QAMD_WAVES_ATTR void dslash_kernel(DslashParams<typename Traits<S>::R> p)
{
  using R = typename Traits<S>::R;
#if QAMD_XCD_REMAP
  const int blk = xcd_block();
#else
  const int blk = (int)blockIdx.x;
#endif
  if (blk >= p.nblocks || (p.stop && *p.stop)) return;
  const Geom &g = p.g;
  int idx;
  if (p.order == 1) {
    // XCD k owns z-planes [k zs, (k+1) zs) for every t, visited t-major, so
    // t-neighbours are zs*bpp work-groups apart and z-neighbours bpp apart
    const unsigned k = fdiv((unsigned)blk, p.f_per), r = (unsigned)blk - k * p.f_per.d;
    const unsigned t = fdiv(r, p.f_zb), r2 = r - t * p.f_zb.d;
    const unsigned zl = fdiv(r2, p.f_bpp), h = r2 - zl * p.f_bpp.d;
    const int z = (int)k * p.zs + (int)zl;
    idx = ((int)t * g.Lz + z) * g.sz + (int)h * kDslashBlock + (int)threadIdx.x;
  } else if (p.order == 2) {
    // XCD k owns t-slices [k ts, (k+1) ts), visited z-major: t-neighbours
    // are bpp work-groups apart, z-neighbours ts*bpp
    const int per = p.ts * g.Lz * p.bpp;
    const int k = blk / per, r = blk - k * per;
    const int z = r / (p.ts * p.bpp), r2 = r - z * (p.ts * p.bpp);
    const int t = k * p.ts + r2 / p.bpp, h = r2 % p.bpp;
    idx = (t * g.Lz + z) * g.sz + h * kDslashBlock + (int)threadIdx.x;
  } else if (p.order == 3) {
    // XCD k owns t-slices [k ts, (k+1) ts) and sweeps them one z-part (zs
    // planes) at a time: the spinor rows reused across three t-slices then
    // span a part of the slice, not the whole slice
    const int per = p.ts * g.Lz * p.bpp;
    const int k = blk / per, r = blk - k * per;
    const int part = p.ts * p.zs * p.bpp;
    const int zp = r / part, r2 = r - zp * part;
    const int t = k * p.ts + r2 / (p.zs * p.bpp), r3 = r2 % (p.zs * p.bpp);
    const int z = zp * p.zs + r3 / p.bpp, h = r3 % p.bpp;
    idx = (t * g.Lz + z) * g.sz + h * kDslashBlock + (int)threadIdx.x;
  } else {
    idx = blk * kDslashBlock + (int)threadIdx.x;
  }
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, p.cb);
  const S *__restrict__ in = static_cast<const S *>(p.in);
  const S *__restrict__ gauge = static_cast<const S *>(p.gauge);
  const int vcb = g.vcb;

  // neighbour cb indices (the input lives on the other checkerboard)
  const int row = idx - c.xh;
  const int xb = c.x == 0 ? g.Lx - 1 : c.x - 1;
  const int xf = c.x == g.Lx - 1 ? 0 : c.x + 1;
  const int n_xb = row + (xb >> 1), n_xf = row + (xf >> 1);
  const int n_yb = c.y == 0 ? idx + (g.Ly - 1) * g.sy : idx - g.sy;
  const int n_yf = c.y == g.Ly - 1 ? idx - (g.Ly - 1) * g.sy : idx + g.sy;
  const int n_zb = c.z == 0 ? idx + (g.Lz - 1) * g.sz : idx - g.sz;
  const int n_zf = c.z == g.Lz - 1 ? idx - (g.Lz - 1) * g.sz : idx + g.sz;
  const int n_tb = c.t == 0 ? idx + (g.Lt - 1) * g.st : idx - g.st;
  const int n_tf = c.t == g.Lt - 1 ? idx - (g.Lt - 1) * g.st : idx + g.st;
  // Direction coefficients; in a split dimension the hops that wrap around
  // the local lattice belong to the faces and are dropped here (the
  // reference's accumulate[d] = 0 on boundary blocks, codegen/dslash.cc:248).
  const R bs = p.coeff_s;
  const int cm = p.comm;
  R bxb = ((cm & 1) && c.x == 0) ? R(0) : bs;
  R bxf = ((cm & 1) && c.x == g.Lx - 1) ? R(0) : bs;
  R byb = ((cm & 2) && c.y == 0) ? R(0) : bs;
  R byf = ((cm & 2) && c.y == g.Ly - 1) ? R(0) : bs;
  R bzb = ((cm & 4) && c.z == 0) ? R(0) : bs;
  R bzf = ((cm & 4) && c.z == g.Lz - 1) ? R(0) : bs;
  R btb = c.t == 0 ? ((cm & 8) ? R(0) : p.coeff_tb * p.tbound) : p.coeff_tb;
  R btf = c.t == g.Lt - 1 ? ((cm & 8) ? R(0) : p.coeff_tf * p.tbound) : p.coeff_tf;
  // QPhiX arrays: element-0 offsets of this site and of its neighbours
  int q0 = 0, qf0 = 0;
  // the group (quad, or pair for FP64) of lanes that SPL = 2 reads together
  constexpr int QG = std::is_same<S, double>::value ? 2 : 4;
  const int xg = c.xh & ~(QG - 1);
  if constexpr (SPL) {
    q0 = qphix_site(p, c.xh, c.y, c.z, c.t);
    qf0 = qphix_site(p, xg, c.y, c.z, c.t);
    if (p.acc) {
      // the reference's accumulate[d]: a cleared bit drops direction d here
      // (a zero coefficient adds exact zeros, the neighbour being finite)
      const int blk = p.q_pxyz * c.t + p.q_pxy * c.z + p.q_nvecs * c.y + (c.xh >> p.q_sh);
      const unsigned m = p.acc[(blk << p.q_sh) + (c.xh & ((1 << p.q_sh) - 1))];
      bxb = (m & 1) ? bxb : R(0);
      bxf = (m & 2) ? bxf : R(0);
      byb = (m & 4) ? byb : R(0);
      byf = (m & 8) ? byf : R(0);
      bzb = (m & 16) ? bzb : R(0);
      bzf = (m & 32) ? bzf : R(0);
      btb = (m & 64) ? btb : R(0);
      btf = (m & 128) ? btf : R(0);
    }
  }

  cx<R> acc[4][3];
  // a chi - b D psi in the reference's order (codegen/dslash_common.cc:
  // 1830-1932): the output starts as the chi term, and every direction's hop
  // is then subtracted by its own fused multiply-adds (mbeta tables,
  // codegen/dslash_common.cc:784-890): its coefficient is negated.  The chi
  // term is formed just before the first hop (after the LDS barrier), so
  // that it is not held across the staging loop.
  auto start_acc = [&]() {
  if constexpr (ACHI) {
    R chr[24];
    if constexpr (SPL == 2)
      load_qv_reals<S, R>(static_cast<const S *>(p.chi), qf0, p.q_sh, chr);
    else if constexpr (SPL)
      load_q_reals<S, R, true>(static_cast<const S *>(p.chi), q0, p.q_sh, chr);
    else
      load_reals<S, spinor_w<S>(), 0, 24, R, true>(static_cast<const S *>(p.chi), vcb, idx, chr);
    cx<R> chi[4][3];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int k = 0; k < 3; k++) chi[s][k] = {chr[(s * 3 + k) * 2], chr[(s * 3 + k) * 2 + 1]};
    if (OP == OP_WILSON) {
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int k = 0; k < 3; k++) acc[s][k] = p.alpha * chi[s][k];
    } else if (OP == OP_TM) {  // alpha (1 + isign i mu g5) chi (codegen/dslash_common.cc:1364-1489)
      twist_ord<R>(chi, acc, p.alpha, (R)ISIGN * p.mu);
    } else {
      if (OP == OP_CLOVER)
        clover_packed<S>(static_cast<const S *>(p.clov), vcb, idx, chi, acc);
      else
        clover_full<S>(static_cast<const S *>(p.clov), vcb, idx, chi, acc);
      // + isign i rho g5 chi, one fused multiply-add per part
      // (codegen/dslash_common.cc:1873-1906)
      const R rho = (R)ISIGN * p.rho;
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const R f = s < 2 ? rho : -rho;
#pragma unroll
        for (int k = 0; k < 3; k++)
          acc[s][k] = {fmx(-f, chi[s][k].im, acc[s][k].re), fmx(f, chi[s][k].re, acc[s][k].im)};
      }
    }
  } else {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int k = 0; k < 3; k++) acc[s][k] = {R(0), R(0)};
  }
  };
  if constexpr (ACHI) {
    bxb = -bxb;
    bxf = -bxf;
    byb = -byb;
    byf = -byf;
    bzb = -bzb;
    bzf = -bzf;
    btb = -btb;
    btf = -btf;
  }

  const int tm = p.tbc_mask;
  const cx<R> ph0 = {p.tbc_re[0], p.tbc_im[0]}, ph1 = {p.tbc_re[1], p.tbc_im[1]};
  const cx<R> ph2 = {p.tbc_re[2], p.tbc_im[2]}, ph3 = {p.tbc_re[3], p.tbc_im[3]};
  if constexpr (SPL == 1) {
    const int yb = c.y == 0 ? g.Ly - 1 : c.y - 1, yf = c.y == g.Ly - 1 ? 0 : c.y + 1;
    const int zb = c.z == 0 ? g.Lz - 1 : c.z - 1, zf = c.z == g.Lz - 1 ? 0 : c.z + 1;
    const int tb = c.t == 0 ? g.Lt - 1 : c.t - 1, tf = c.t == g.Lt - 1 ? 0 : c.t + 1;
    const int sh = p.q_sh;
    start_acc();
    hop_mu<S, RECON, 0, ISIGN, 1>(acc, in, gauge, vcb, idx, qphix_site(p, xb >> 1, c.y, c.z, c.t),
                                  qphix_site(p, xf >> 1, c.y, c.z, c.t), bxb, bxf, tm & 1, ph0, p.one, sh);
    hop_mu<S, RECON, 1, ISIGN, 1>(acc, in, gauge, vcb, idx, qphix_site(p, c.xh, yb, c.z, c.t),
                                  qphix_site(p, c.xh, yf, c.z, c.t), byb, byf, tm & 2, ph1, p.one, sh);
    hop_mu<S, RECON, 2, ISIGN, 1>(acc, in, gauge, vcb, idx, qphix_site(p, c.xh, c.y, zb, c.t),
                                  qphix_site(p, c.xh, c.y, zf, c.t), bzb, bzf, tm & 4, ph2, p.one, sh);
    hop_mu<S, RECON, 3, ISIGN, 1>(acc, in, gauge, vcb, idx, qphix_site(p, c.xh, c.y, c.z, tb),
                                  qphix_site(p, c.xh, c.y, c.z, tf), btb, btf, tm & 8, ph3, p.one, sh);
  } else {
#if QAMD_LDS_XY
  // A work-group inside one (z, t) plane stages its x rows y0 .. y1 and the
  // rows y0 - 1 and y1 + 1 (y periodic) once; its x and y hops then read
  // LDS.  One that straddles two planes (when a plane does not hold whole
  // work-groups) reads its neighbours from global memory.  The host checks
  // that every work-group is full (no lane leaves before the barrier, which
  // the branch below reaches uniformly per work-group) and that the rows fit
  // the array.  SPL = 2 stages the QPhiX array's sites through the lane
  // groups' vector loads (whole groups: Lx/2 and the staged count are
  // multiples of the group), in the native chunk order.
  int y0 = 0, r1 = 0;
  if (p.lds) {
    const unsigned base = (unsigned)(idx - (int)threadIdx.x);
    const unsigned r0 = fdiv(base, g.fxh);  // rows numbered over the checkerboard
    r1 = (int)(fdiv(base + kDslashBlock - 1, g.fxh) - r0);
    y0 = (int)frem(r0, g.fy);
  }
  const bool use_lds = p.lds && y0 + r1 < g.Ly;  // uniform per work-group
  __shared__ Vec16 lds[(24 * sizeof(S) / 16) * kLdsSites];
  if (use_lds) {
    const int plane = row - c.y * g.Lxh;
    const int nst = (r1 + 3) * g.Lxh;
    const Vec16 *vin = reinterpret_cast<const Vec16 *>(in);
    constexpr int NC = 24 * (int)sizeof(S) / 16;
    for (int j = (int)threadIdx.x; j < nst; j += kDslashBlock) {
      const int k = (int)fdiv((unsigned)j, g.fxh);
      // staged row k is y0 + k - 1, in [-1, Ly] since k <= r1 + 2 and y0 + r1 < Ly
      int yk = y0 + k - 1;
      yk = yk < 0 ? yk + g.Ly : (yk >= g.Ly ? yk - g.Ly : yk);
      if constexpr (SPL == 2 && std::is_same<S, double>::value) {
        // each lane's vector holds one real part of one chunk of the pair's
        // two sites: written to LDS as it is, no transpose
        const int jb = j & 1, j0 = j - jb;
        const int qf = qphix_site(p, (j - k * g.Lxh) & ~1, yk, c.z, c.t);
        double *ld = reinterpret_cast<double *>(lds);
#pragma unroll
        for (int m = 0; m < 12; m++) {
          const u32x4 *a = reinterpret_cast<const u32x4 *>(static_cast<const double *>(in) + qf + ((2 * m + jb) << p.q_sh));
          const u32x4 v = *a;
          double t[2];
          __builtin_memcpy(t, &v, 16);
          const int ch = qphix_to_native(2 * m) >> 1;  // the chunk of elements 2m, 2m + 1
          ld[(ch * kLdsSites + j0) * 2 + jb] = t[0];
          ld[(ch * kLdsSites + j0 + 1) * 2 + jb] = t[1];
        }
      } else if constexpr (SPL == 2) {
        S e[24], nat[24];
        load_qv_elems<S, false>(in, qphix_site(p, (j - k * g.Lxh) & ~(QG - 1), yk, c.z, c.t), p.q_sh, e);
#pragma unroll
        for (int E = 0; E < 24; E++) nat[qphix_to_native(E)] = e[E];
#pragma unroll
        for (int ch = 0; ch < NC; ch++) {
          Vec16 v;
          __builtin_memcpy(&v, nat + ch * (16 / (int)sizeof(S)), 16);
          lds[ch * kLdsSites + j] = v;
        }
      } else {
        const int src = plane + yk * g.Lxh + (j - k * g.Lxh);
#pragma unroll
        for (int ch = 0; ch < NC; ch++) lds[ch * kLdsSites + j] = vin[(unsigned)(ch * vcb + src)];
      }
    }
    __syncthreads();
  }
#else
  const bool use_lds = false;
#endif
  start_acc();
  // SPL = 2: neighbours read from the QPhiX array by lane group (the first
  // site of the group, xg, in the y, z and t directions)
  const int sh = p.q_sh;
  const int yb = c.y == 0 ? g.Ly - 1 : c.y - 1, yf = c.y == g.Ly - 1 ? 0 : c.y + 1;
#if QAMD_LDS_XY
  if (use_lds) {
    const int rl = c.y - y0 + 1;  // staged row of this site
    hop_mu_lds<S, RECON, 0, ISIGN>(acc, lds, gauge, vcb, idx, rl * g.Lxh + (xb >> 1),
                                   rl * g.Lxh + (xf >> 1), bxb, bxf, tm & 1, ph0, p.one);
    hop_mu_lds<S, RECON, 1, ISIGN>(acc, lds, gauge, vcb, idx, (rl - 1) * g.Lxh + c.xh,
                                   (rl + 1) * g.Lxh + c.xh, byb, byf, tm & 2, ph1, p.one);
  } else
#endif
  if constexpr (SPL == 2) {
    hop_mu<S, RECON, 0, ISIGN, 2>(acc, in, gauge, vcb, idx, qphix_site(p, xb >> 1, c.y, c.z, c.t),
                                  qphix_site(p, xf >> 1, c.y, c.z, c.t), bxb, bxf, tm & 1, ph0, p.one, sh);
    hop_mu<S, RECON, 1, ISIGN, 2>(acc, in, gauge, vcb, idx, qphix_site(p, xg, yb, c.z, c.t),
                                  qphix_site(p, xg, yf, c.z, c.t), byb, byf, tm & 2, ph1, p.one, sh);
  } else {
    hop_mu<S, RECON, 0, ISIGN>(acc, in, gauge, vcb, idx, n_xb, n_xf, bxb, bxf, tm & 1, ph0, p.one);
    hop_mu<S, RECON, 1, ISIGN>(acc, in, gauge, vcb, idx, n_yb, n_yf, byb, byf, tm & 2, ph1, p.one);
  }
  if constexpr (SPL == 2) {
    const int zb = c.z == 0 ? g.Lz - 1 : c.z - 1, zf = c.z == g.Lz - 1 ? 0 : c.z + 1;
    const int tb = c.t == 0 ? g.Lt - 1 : c.t - 1, tf = c.t == g.Lt - 1 ? 0 : c.t + 1;
    hop_mu<S, RECON, 2, ISIGN, 2>(acc, in, gauge, vcb, idx, qphix_site(p, xg, c.y, zb, c.t),
                                  qphix_site(p, xg, c.y, zf, c.t), bzb, bzf, tm & 4, ph2, p.one, sh);
    hop_mu<S, RECON, 3, ISIGN, 2>(acc, in, gauge, vcb, idx, qphix_site(p, xg, c.y, c.z, tb),
                                  qphix_site(p, xg, c.y, c.z, tf), btb, btf, tm & 8, ph3, p.one, sh);
  } else {
    hop_mu<S, RECON, 2, ISIGN>(acc, in, gauge, vcb, idx, n_zb, n_zf, bzb, bzf, tm & 4, ph2, p.one);
    hop_mu<S, RECON, 3, ISIGN>(acc, in, gauge, vcb, idx, n_tb, n_tf, btb, btf, tm & 8, ph3, p.one);
  }
  }

  cx<R> res[4][3];
  if constexpr (!ACHI) {
    if (OP == OP_WILSON) {
#pragma unroll
      for (int s = 0; s < 4; s++)
#pragma unroll
        for (int k = 0; k < 3; k++) res[s][k] = acc[s][k];
    } else if (OP == OP_CLOVER) {
      clover_packed<S>(static_cast<const S *>(p.clov), vcb, idx, acc, res);
    } else if (OP == OP_TMCLOVER) {
      clover_full<S>(static_cast<const S *>(p.clov), vcb, idx, acc, res);
    } else {  // muinv (1 - isign i mu g5) (codegen/dslash_common.cc:1234-1359)
      twist_ord<R>(acc, res, p.muinv, -(R)ISIGN * p.mu);
    }
  } else {
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int k = 0; k < 3; k++) res[s][k] = acc[s][k];
  }

  R o[24];
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      o[(s * 3 + k) * 2] = res[s][k].re;
      o[(s * 3 + k) * 2 + 1] = res[s][k].im;
    }
  if constexpr (EPI == EPI_RESIDUAL || EPI == EPI_TRUE_RESIDUAL) {
    // r - step * result; each site reads and writes only its own r (b for
    // the true residual)
    R rv[24];
    load_reals<S, spinor_w<S>(), 0, 24, R>(
        static_cast<const S *>(EPI == EPI_RESIDUAL ? p.out : p.rin), vcb, idx, rv);
    const R a = (R)*p.step;
#pragma unroll
    for (int k = 0; k < 24; k++) o[k] = rv[k] - a * o[k];
  }
  if constexpr (SPL == 2)
    store_qv_reals<S, R>(static_cast<S *>(p.out), qf0, p.q_sh, o);
  else if constexpr (SPL)
    store_q_reals<S, R>(static_cast<S *>(p.out), q0, p.q_sh, o);
  else
    store_reals<S, spinor_w<S>(), 0, 24, R, true>(static_cast<S *>(p.out), vcb, idx, o);
  if constexpr (EPI != EPI_STORE) {
    // |out|^2 of the stored (rounded) values: lanes in FP32/FP64, the wave
    // and the work-group in FP64 in a fixed order
    R s2 = 0;
#pragma unroll
    for (int k = 0; k < 24; k++) {
      const R v = (R)to_real(to_storage<S>(o[k]));
      s2 += v * v;
    }
    double w = (double)s2;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) w += __shfl_xor(w, off, 64);
    __shared__ double wsum[kDslashBlock / 64];
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
#pragma unroll
      for (int i = 0; i < kDslashBlock / 64; i++) t += wsum[i];
      p.norm[(idx - (int)threadIdx.x) / kDslashBlock] = t;
    }
  }
}

// Host-side launch for one instantiation.
template <typename S, int RECON, int OP, bool ACHI, int ISIGN, int SPL = 0>
hipError_t launch_dslash(const DslashParams<typename Traits<S>::R> &p, hipStream_t stream)
{
  const int grid = xcd_grid(p.nblocks);
  if constexpr (SPL) {  // QPhiX arrays: plain store only
    if (p.norm || p.step) return hipErrorInvalidValue;
    hipLaunchKernelGGL((dslash_kernel<S, RECON, OP, ACHI, ISIGN, EPI_STORE, SPL>), dim3(grid),
                       dim3(kDslashBlock), 0, stream, p);
    return hipGetLastError();
  }
  if constexpr (ACHI && (OP == OP_CLOVER || OP == OP_TMCLOVER)) {
    if (p.norm && !p.step) {
      hipLaunchKernelGGL((dslash_kernel<S, RECON, OP, ACHI, ISIGN, EPI_NORM>), dim3(grid), dim3(kDslashBlock),
                         0, stream, p);
      return hipGetLastError();
    }
    // the residual update follows M^dag (isign -1), the second half of M^dag M
    if constexpr (ISIGN == -1) {
      if (p.norm && p.step && !p.rin) {
        hipLaunchKernelGGL((dslash_kernel<S, RECON, OP, ACHI, ISIGN, EPI_RESIDUAL>), dim3(grid),
                           dim3(kDslashBlock), 0, stream, p);
        return hipGetLastError();
      }
      if constexpr (std::is_same<S, double>::value) {
        if (p.norm && p.step && p.rin) {
          hipLaunchKernelGGL((dslash_kernel<S, RECON, OP, ACHI, ISIGN, EPI_TRUE_RESIDUAL>), dim3(grid),
                             dim3(kDslashBlock), 0, stream, p);
          return hipGetLastError();
        }
      }
    }
  }
  if (p.norm || p.step) return hipErrorInvalidValue;
  hipLaunchKernelGGL((dslash_kernel<S, RECON, OP, ACHI, ISIGN>), dim3(grid), dim3(kDslashBlock), 0,
                     stream, p);
  return hipGetLastError();
}

// Run-time dispatch over (recon, op, achi, isign) for one storage type;
// SPL = 1 for QPhiX spinor arrays, 2 for the same read as vectors.
template <typename S, int SPL = 0>
hipError_t dispatch_dslash(int recon, int op, bool achi, int isign,
                           const DslashParams<typename Traits<S>::R> &p, hipStream_t st)
{
#define QAMD_ISIGN(RC, OPV, AC)                                              \
  return isign > 0 ? launch_dslash<S, RC, OPV, AC, 1, SPL>(p, st)            \
                   : launch_dslash<S, RC, OPV, AC, -1, SPL>(p, st);
#define QAMD_ACHI(RC, OPV)                                                   \
  if (achi) {                                                                \
    QAMD_ISIGN(RC, OPV, true)                                                \
  } else {                                                                   \
    QAMD_ISIGN(RC, OPV, false)                                               \
  }
#define QAMD_OP(RC)                                                          \
  switch (op) {                                                              \
    case OP_WILSON: QAMD_ACHI(RC, OP_WILSON)                                 \
    case OP_CLOVER: QAMD_ACHI(RC, OP_CLOVER)                                 \
    case OP_TM: QAMD_ACHI(RC, OP_TM)                                         \
    case OP_TMCLOVER: QAMD_ACHI(RC, OP_TMCLOVER)                             \
    default: return hipErrorInvalidValue;                                    \
  }
  if (recon == 12) {
    QAMD_OP(12)
  } else {
    QAMD_OP(18)
  }
#undef QAMD_OP
#undef QAMD_ACHI
#undef QAMD_ISIGN
}

}  // namespace qamd
