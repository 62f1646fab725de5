// ref_driver.cpp -- a QPhiX-style block driver around the reference's own
// generated kernel bodies (qphix-codegen), compiled per ISA configuration by
// oracle/ref.mk into oracle/_ref/libref_<config>.so.
//
// TEST INFRASTRUCTURE ONLY: the tests use it to pin the FP64 oracle against
// the reference's kernels, and bench.py times it as the CPU baseline.
// QPhiX itself (the runtime caller of these bodies) is not part of the
// reference tree, so this file plays its part:
//   * packs lexicographic FP64 fields into the reference's block layouts
//     (codegen/include/decl_common.h:9-62): FourSpinorBlock
//     FT[3 colour][4 spin][2][SOALEN], SU3MatrixBlock FT[8][2|3][3][2][VECLEN]
//     holding the transpose of each hop matrix (codegen/dslash_common.cc:
//     1678-1727), CloverBlock / FullCloverBlock;
//   * computes the per-block neighbour offsets, including the X-direction
//     split loads whose lane shift depends on the row parity
//     (codegen/data_types.cc:140-198, loadSplitSOAFVec in inst_sp_vec16.cc);
//   * passes coeff_t_b / coeff_t_f multiplied by the time boundary factor on
//     the first / last time slice (codegen/dslash.cc:79-86);
//   * runs the block loop with OpenMP, one body call per SU3MatrixBlock:
//     VECLEN lanes = NGY = VECLEN / SOALEN rows of y, each row segment of
//     SOALEN sites addressed through the offset arrays at lane r * SOALEN
//     (the bodies load row segment r from base + offs[r * SOALEN],
//     codegen/data_types.cc:64-119, codegen/address_types.h:281-312);
//   * builds accumulate[d] from an optional per-site direction mask: bit l
//     of accumulate[d] is bit d of lane l's site (all lanes when no mask).
// Requires Lx/2 a multiple of SOALEN and Ly a multiple of NGY.
#include <immintrin.h>
#include <omp.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "decl_common.h"
#if defined(AVX2) && !defined(AVX512)
#include "avx_utils.h"
#endif

#if VECLEN == 1
// the reference's own glue removes prefetches from scalar bodies the same way
// (codegen/jinja/complete_specialization.h.j2:54-57)
#define _mm_prefetch(a, b) \
  {                        \
  }
#endif

#if defined(USE_LP_SPINOR)
// low-precision build: every array and every scalar is IEEE binary16
// (codegen/data_types.h:7-39; the bodies broadcast scalars from half,
// codegen/inst_sp_vec16.cc:190-213), converted round-to-nearest-even as the
// kernels' own stores do (codegen/inst_sp_vec16.cc:5-7,129-135)
typedef QPhiX::half FT;
static inline FT to_ft(double v) { return (FT)_cvtss_sh((float)v, _MM_FROUND_TO_NEAREST_INT); }
static inline double from_ft(FT v) { return (double)_cvtsh_ss(v); }
#elif PRECISION == 1
typedef float FT;
#else
typedef double FT;
#endif
#if !defined(USE_LP_SPINOR)
static inline FT to_ft(double v) { return (FT)v; }
static inline double from_ft(FT v) { return (double)v; }
#endif

#if VECLEN % SOALEN != 0
#error "VECLEN must be a multiple of SOALEN"
#endif

struct RefArgs {
  const void *xyBase, *zbBase, *zfBase, *tbBase, *tfBase, *chiBase;
  void *oBase;
  const void *gBase, *clBase;
  int xbOffs[VECLEN], xfOffs[VECLEN], ybOffs[VECLEN], yfOffs[VECLEN], offs[VECLEN],
      gOffs[VECLEN], pfyOffs[VECLEN];
  unsigned int accumulate[8];
  FT alpha, coeff_s, coeff_t_f, coeff_t_b, mu, muinv, prec_mass_rho;
  FT tbc_phases[4][2];
};

typedef void (*ref_body_fn)(const RefArgs &);

// Binds every name a body may use.
#define REF_BIND_COMMON                                                              \
  const Ty::FourSpinorBlock *xyBase = (const Ty::FourSpinorBlock *)A.xyBase;         \
  const Ty::FourSpinorBlock *zbBase = (const Ty::FourSpinorBlock *)A.zbBase;         \
  const Ty::FourSpinorBlock *zfBase = (const Ty::FourSpinorBlock *)A.zfBase;         \
  const Ty::FourSpinorBlock *tbBase = (const Ty::FourSpinorBlock *)A.tbBase;         \
  const Ty::FourSpinorBlock *tfBase = (const Ty::FourSpinorBlock *)A.tfBase;         \
  const Ty::FourSpinorBlock *chiBase = (const Ty::FourSpinorBlock *)A.chiBase;       \
  Ty::FourSpinorBlock *oBase = (Ty::FourSpinorBlock *)A.oBase;                       \
  const Ty::SU3MatrixBlock *gBase = (const Ty::SU3MatrixBlock *)A.gBase;             \
  const int *xbOffs = A.xbOffs, *xfOffs = A.xfOffs, *ybOffs = A.ybOffs;              \
  const int *yfOffs = A.yfOffs, *offs = A.offs, *gOffs = A.gOffs;                    \
  const int *pfyOffs = A.pfyOffs;                                                    \
  const Ty::FourSpinorBlock *pfBase1 = 0, *pfBase2 = 0, *pfBase3 = 0, *pfBase4 = 0;  \
  const Ty::FourSpinorBlock *pfBaseChi = 0;                                          \
  const int siprefdist1 = 0, siprefdist2 = 0, siprefdist3 = 0, siprefdist4 = 0;      \
  const int chiprefdist = 0, gprefdist = 0, clprefdist = 0;                          \
  const unsigned int *accumulate = A.accumulate;                                     \
  const FT alpha = A.alpha, coeff_s = A.coeff_s, coeff_t_f = A.coeff_t_f;            \
  const FT coeff_t_b = A.coeff_t_b, mu = A.mu, muinv = A.muinv;                      \
  const FT prec_mass_rho = A.prec_mass_rho;                                          \
  const FT(*tbc_phases)[2] = A.tbc_phases;

// The face bodies' parameters (codegen/jinja/*_dslash_decl.h.j2:90-115).
struct FaceArgs {
  const void *xyBase, *gBase, *clBase;
  void *oBase;
  const FT *inbuf;
  FT *outbuf;
  int offs[VECLEN], gOffs[VECLEN];
  unsigned int mask;
  int dir;
  FT beta, mu, muinv;
  FT tbc_phases[4][2];
};

typedef void (*ref_face_fn)(const FaceArgs &);

#define REF_BIND_FACE                                                                \
  const Ty::FourSpinorBlock *xyBase = (const Ty::FourSpinorBlock *)F.xyBase;         \
  Ty::FourSpinorBlock *oBase = (Ty::FourSpinorBlock *)F.oBase;                       \
  const Ty::SU3MatrixBlock *gBase = (const Ty::SU3MatrixBlock *)F.gBase;             \
  const FT *inbuf = F.inbuf;                                                         \
  FT *outbuf = F.outbuf;                                                             \
  const int *offs = F.offs, *gOffs = F.gOffs;                                        \
  const int si_prefdist = 0, hsprefdist = 0, gprefdist = 0, soprefdist = 0;          \
  const int clprefdist = 0;                                                          \
  const unsigned int mask = F.mask;                                                  \
  const int dir = F.dir;                                                             \
  const FT beta = F.beta, mu = F.mu, muinv = F.muinv;                                \
  const FT(*tbc_phases)[2] = F.tbc_phases;

using namespace QPhiX;

#include "bodies.inc"
#include "faces.inc"

namespace {

constexpr int S = SOALEN;
constexpr int NGY = VECLEN / SOALEN;  // rows of y per gauge / clover vector
constexpr int BLK = 24 * S;           // FT elements per FourSpinorBlock

struct Lat {
  int L[4], Lxh, nxb;
  long V;
  // FourSpinorBlock of row (t, z, y), x-segment xb
  long block(int t, int z, int y, int xb) const
  {
    return (((long)t * L[2] + z) * L[1] + y) * nxb + xb;
  }
  // SU3MatrixBlock / CloverBlock holding rows y / NGY * NGY .. + NGY - 1
  long gblock(int t, int z, int y, int xb) const
  {
    return (((long)t * L[2] + z) * (L[1] / NGY) + y / NGY) * nxb + xb;
  }
  long lex(int x, int y, int z, int t) const
  {
    return (((long)t * L[2] + z) * L[1] + y) * L[0] + x;
  }
  int xof(int xh, int y, int z, int t, int cb) const { return 2 * xh + ((cb + y + z + t) & 1); }
  long nblk() const { return (long)L[3] * L[2] * L[1] * nxb; }
};

template <typename T> T *amalloc(size_t n)
{
  void *p = aligned_alloc(64, ((n * sizeof(T) + 63) / 64) * 64);
  memset(p, 0, n * sizeof(T));
  return (T *)p;
}

// lex FP64 spinor [V][4][3][2] -> blocks FT[3][4][2][S] of checkerboard cb
void pack_spinor(FT *dst, const double *src, const Lat &g, int cb)
{
  for (int t = 0; t < g.L[3]; t++)
    for (int z = 0; z < g.L[2]; z++)
      for (int y = 0; y < g.L[1]; y++)
        for (int xb = 0; xb < g.nxb; xb++) {
          FT *b = dst + g.block(t, z, y, xb) * BLK;
          for (int i = 0; i < S; i++) {
            long s = g.lex(g.xof(xb * S + i, y, z, t, cb), y, z, t);
            for (int c = 0; c < 3; c++)
              for (int sp = 0; sp < 4; sp++)
                for (int r = 0; r < 2; r++)
                  b[((c * 4 + sp) * 2 + r) * S + i] = to_ft(src[((s * 4 + sp) * 3 + c) * 2 + r]);
          }
        }
}

void unpack_spinor(double *dst, const FT *src, const Lat &g, int cb)
{
  for (int t = 0; t < g.L[3]; t++)
    for (int z = 0; z < g.L[2]; z++)
      for (int y = 0; y < g.L[1]; y++)
        for (int xb = 0; xb < g.nxb; xb++) {
          const FT *b = src + g.block(t, z, y, xb) * BLK;
          for (int i = 0; i < S; i++) {
            long s = g.lex(g.xof(xb * S + i, y, z, t, cb), y, z, t);
            for (int c = 0; c < 3; c++)
              for (int sp = 0; sp < 4; sp++)
                for (int r = 0; r < 2; r++)
                  dst[((s * 4 + sp) * 3 + c) * 2 + r] = from_ft(b[((c * 4 + sp) * 2 + r) * S + i]);
          }
        }
}

// Gauge blocks FT[8][nrow][3][2][VECLEN]: S[d][i][j] = M_d[j][i], with
// M_{2mu} = U_mu(x - mu) (applied as its adjoint) and M_{2mu+1} = U_mu(x);
// the site of row y, lane i sits in lane (y % NGY) * S + i.
void pack_gauge(FT *dst, const double *U, const Lat &g, int cb, int nrow)
{
  const long gblk = 8L * nrow * 3 * 2 * VECLEN;
  for (int t = 0; t < g.L[3]; t++)
    for (int z = 0; z < g.L[2]; z++)
      for (int y = 0; y < g.L[1]; y++)
        for (int xb = 0; xb < g.nxb; xb++) {
          FT *b = dst + g.gblock(t, z, y, xb) * gblk;
          for (int i = 0; i < S; i++) {
            const int lane = (y % NGY) * S + i;
            int c4[4] = {g.xof(xb * S + i, y, z, t, cb), y, z, t};
            for (int d = 0; d < 8; d++) {
              int mu = d / 2, fb = d % 2;
              int n[4] = {c4[0], c4[1], c4[2], c4[3]};
              if (!fb) n[mu] = (n[mu] - 1 + g.L[mu]) % g.L[mu];
              long s = g.lex(n[0], n[1], n[2], n[3]);
              const double *m = U + (s * 4 + mu) * 18;
              for (int r = 0; r < nrow; r++)
                for (int col = 0; col < 3; col++)
                  for (int ri = 0; ri < 2; ri++)
                    b[(((d * nrow + r) * 3 + col) * 2 + ri) * VECLEN + lane] =
                        to_ft(m[(col * 3 + r) * 2 + ri]);
            }
          }
        }
}

// lex [V][2][36] -> CloverBlock {diag1[6][V], off1[15][2][V], diag2, off2}
void pack_clover(FT *dst, const double *cl, const Lat &g, int cb)
{
  const long cblk = 72L * VECLEN;
  for (int t = 0; t < g.L[3]; t++)
    for (int z = 0; z < g.L[2]; z++)
      for (int y = 0; y < g.L[1]; y++)
        for (int xb = 0; xb < g.nxb; xb++) {
          FT *b = dst + g.gblock(t, z, y, xb) * cblk;
          for (int i = 0; i < S; i++) {
            const int lane = (y % NGY) * S + i;
            long s = g.lex(g.xof(xb * S + i, y, z, t, cb), y, z, t);
            for (int blk = 0; blk < 2; blk++) {
              const double *src = cl + s * 72 + blk * 36;
              FT *diag = b + blk * 36 * VECLEN;
              FT *off = diag + 6 * VECLEN;
              for (int k = 0; k < 6; k++) diag[k * VECLEN + lane] = to_ft(src[k]);
              for (int k = 0; k < 15; k++)
                for (int ri = 0; ri < 2; ri++) off[(k * 2 + ri) * VECLEN + lane] = to_ft(src[6 + 2 * k + ri]);
            }
          }
        }
}

// lex [V][2][6][6][2] -> FullCloverBlock {block1[6][6][2][V], block2}
void pack_full_clover(FT *dst, const double *cl, const Lat &g, int cb)
{
  const long cblk = 144L * VECLEN;
  for (int t = 0; t < g.L[3]; t++)
    for (int z = 0; z < g.L[2]; z++)
      for (int y = 0; y < g.L[1]; y++)
        for (int xb = 0; xb < g.nxb; xb++) {
          FT *b = dst + g.gblock(t, z, y, xb) * cblk;
          for (int i = 0; i < S; i++) {
            const int lane = (y % NGY) * S + i;
            long s = g.lex(g.xof(xb * S + i, y, z, t, cb), y, z, t);
            for (int k = 0; k < 144; k++) b[k * VECLEN + lane] = to_ft(cl[s * 144 + k]);
          }
        }
}

bool make_lat(Lat &g, int Lx, int Ly, int Lz, int Lt)
{
  if ((Lx | Ly | Lz | Lt) & 1) return false;
  g.L[0] = Lx;
  g.L[1] = Ly;
  g.L[2] = Lz;
  g.L[3] = Lt;
  g.Lxh = Lx / 2;
  if (g.Lxh % S || Ly % NGY) return false;
  g.nxb = g.Lxh / S;
  g.V = (long)Lx * Ly * Lz * Lt;
  return true;
}

}  // namespace

extern "C" {

// Packs a lexicographic FP64 field into the block arrays ref_apply hands to
// the bodies (field 0 spinor, 1 gauge with recon 12|18, 2 clover, 3 full
// clover) for checkerboard cb.  Returns the number of FT elements; with
// dst == NULL only the count.  -1 for an unsupported geometry.
long ref_pack_field(int field, int recon, int cb, int Lx, int Ly, int Lz, int Lt, const double *src,
                    void *dst)
{
  Lat g;
  if (!make_lat(g, Lx, Ly, Lz, Lt) || field < 0 || field > 3) return -1;
  const long nblk = g.nblk();
  const int nrow = recon == 12 ? 2 : 3;
  const long per[4] = {nblk * BLK, nblk / NGY * 8L * nrow * 3 * 2 * VECLEN, nblk / NGY * 72L * VECLEN,
                       nblk / NGY * 144L * VECLEN};
  const long n = per[field];
  if (!dst) return n;
  FT *d = (FT *)dst;
  memset(d, 0, n * sizeof(FT));
  if (field == 0) pack_spinor(d, src, g, cb);
  if (field == 1) pack_gauge(d, src, g, cb, nrow);
  if (field == 2) pack_clover(d, src, g, cb);
  if (field == 3) pack_full_clover(d, src, g, cb);
  return n;
}

// STREAM triad c = a + 1.5 b over three FP64 arrays of n elements on
// nthreads OpenMP threads (arrays first touched by the same static schedule,
// so each thread's pages sit on its own memory node): the host's roofline
// for the kernels above.  Returns the best of reps in GB/s, counting 24
// bytes per element; -1 if the arrays cannot be allocated.
double ref_stream_triad(long n, int nthreads, int reps)
{
  double *a = (double *)aligned_alloc(64, ((n * 8 + 63) / 64) * 64);
  double *b = (double *)aligned_alloc(64, ((n * 8 + 63) / 64) * 64);
  double *c = (double *)aligned_alloc(64, ((n * 8 + 63) / 64) * 64);
  double best = -1.0;
  if (a && b && c) {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long i = 0; i < n; i++) {
      a[i] = 1.0;
      b[i] = 2.0;
      c[i] = 0.0;
    }
    double tmin = 1e300;
    for (int r = 0; r < reps; r++) {
      const double t0 = omp_get_wtime();
#pragma omp parallel for num_threads(nthreads) schedule(static)
      for (long i = 0; i < n; i++) c[i] = a[i] + 1.5 * b[i];
      const double dt = omp_get_wtime() - t0;
      tmin = dt < tmin ? dt : tmin;
    }
    best = 24.0 * (double)n / tmin / 1e9;
  }
  free(a);
  free(b);
  free(c);
  return best;
}

const char *ref_config(void) { return REF_CONFIG_NAME; }
int ref_veclen(void) { return VECLEN; }
int ref_soalen(void) { return SOALEN; }
int ref_precision(void) { return PRECISION; }
int ref_lowprec(void)
{
#if defined(USE_LP_SPINOR)
  return 1;
#else
  return 0;
#endif
}

// kind: 0 wilson, 1 clover, 2 tm, 3 tm-clover.  Inputs are lexicographic
// FP64 arrays; psi is read on checkerboard 1-cb, chi on cb, and out is
// written on cb.  Returns 0, or -1 for an unsupported geometry/argument.
// reps > 0 times the block loop (not the packing) reps more times and
// stores each call's seconds in seconds[0 .. reps-1].
// tbc_phases (8 doubles, (re, im) per dimension x y z t) selects the bodies
// generated with twisted boundary conditions in every dimension and passes
// the phases (codegen/dslash.cc:206-230); NULL selects the simple bodies.
// mask (optional, one byte per lexicographic site, read at the output
// sites): bit d enables direction d = 2 mu + fb; lane l of a body call's
// accumulate[d] is bit d of the site in lane l.  The bodies apply these
// lane masks to the x hops (and to the y hops when SOALEN < VECLEN) and
// otherwise run a direction for the whole vector when any lane's bit is set
// (codegen/dslash.cc:109,248,285-290,570-578).
int ref_apply(int kind, int achi, int isign, int recon, int cb, int Lx, int Ly, int Lz, int Lt,
              const double *psi, const double *chi, const double *U, const double *clov,
              double *out, double coeff_s, double coeff_t_f, double coeff_t_b,
              double t_boundary, double alpha, double mu, double muinv, double rho,
              int nthreads, int reps, double *seconds, const double *tbc_phases,
              const unsigned char *mask)
{
  if (kind < 0 || kind > 3 || (recon != 12 && recon != 18) || (cb != 0 && cb != 1)) return -1;
  Lat g;
  if (!make_lat(g, Lx, Ly, Lz, Lt)) return -1;
  const long nblk = g.nblk();
  const long ngblk = nblk / NGY;
  const int nrow = recon == 12 ? 2 : 3;
  FT *in = amalloc<FT>(nblk * BLK);
  FT *chib = amalloc<FT>(nblk * BLK);
  FT *o = amalloc<FT>(nblk * BLK);
  const long gblk = 8L * nrow * 3 * 2 * VECLEN;
  FT *gb = amalloc<FT>(ngblk * gblk);
  const long cblk = (kind == 3 ? 144L : 72L) * VECLEN;
  FT *cl = (kind == 1 || kind == 3) ? amalloc<FT>(ngblk * cblk) : nullptr;
  pack_spinor(in, psi, g, 1 - cb);
  if (achi) pack_spinor(chib, chi, g, cb);
  pack_gauge(gb, U, g, cb, nrow);
  if (kind == 1) pack_clover(cl, clov, g, cb);
  if (kind == 3) pack_full_clover(cl, clov, g, cb);

  ref_body_fn fn = (tbc_phases ? ref_table_tbc : ref_table)[kind][achi ? 1 : 0][isign > 0 ? 0 : 1]
                                                          [recon == 12 ? 0 : 1];
  const unsigned int all = VECLEN >= 32 ? 0xffffffffu : ((1u << VECLEN) - 1u);
  const int Lyg = Ly / NGY;

  auto run = [&]() {
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (long gi = 0; gi < ngblk; gi++) {
      long r = gi;
      const int xb = (int)(r % g.nxb);
      r /= g.nxb;
      const int yg = (int)(r % Lyg);
      r /= Lyg;
      const int z = (int)(r % Lz);
      const int t = (int)(r / Lz);
      const int y0 = yg * NGY;
      const long b0 = g.block(t, z, y0, xb);  // the vector's first row segment
      RefArgs A;
      memset(&A, 0, sizeof(A));
      for (int row = 0; row < NGY; row++) {
        const int y = y0 + row, L0 = row * S;
        const long bi = g.block(t, z, y, xb);
        const int ro = (int)((bi - b0) * BLK);  // this row's segment, from xyBase
        const int p = (cb + y + z + t) & 1;    // x = 2 xh + p on this row
        const long xprev = g.block(t, z, y, (xb + g.nxb - 1) % g.nxb);
        const long xnext = g.block(t, z, y, (xb + 1) % g.nxb);
        A.offs[L0] = ro;
        // X split loads (codegen/inst_sp_vec16.cc:609-636): the backward
        // neighbour's lanes 1..S-1 come from xyBase + xbOffs[L0 + 1] - 1 and
        // lane 0 from xyBase + xbOffs[L0]; the forward neighbour's lanes
        // 0..S-2 from xyBase + xfOffs[L0] and lane S-1 from xfOffs[L0 + S-1].
        if (p == 1) {  // x-1 = 2 xh: same lanes; x+1 = 2 xh + 2: shift by one
          A.xbOffs[L0] = ro;
          if (S > 1) A.xbOffs[L0 + 1] = ro + 1;
          A.xfOffs[L0] = ro + 1;
          A.xfOffs[L0 + S - 1] = (int)((xnext - b0) * BLK);
        } else {  // x-1 = 2 xh - 1: shift by one; x+1 = 2 xh + 1: same lanes
          A.xbOffs[L0] = (int)((xprev - b0) * BLK + (S - 1));
          if (S > 1) A.xbOffs[L0 + 1] = ro;
          A.xfOffs[L0] = ro;
          A.xfOffs[L0 + S - 1] = ro + S - 1;
        }
        A.ybOffs[L0] = (int)((g.block(t, z, (y + Ly - 1) % Ly, xb) - b0) * BLK);
        A.yfOffs[L0] = (int)((g.block(t, z, (y + 1) % Ly, xb) - b0) * BLK);
        for (int i = 0; i < S; i++)
          for (int d = 0; d < 8; d++) {
            const bool on = mask ? (mask[g.lex(g.xof(xb * S + i, y, z, t, cb), y, z, t)] >> d) & 1 : true;
            if (on) A.accumulate[d] |= 1u << (L0 + i);
          }
      }
      for (int d = 0; d < 8; d++) A.accumulate[d] &= all;
      // z / t neighbours: the neighbouring row-(y0) segment; rows further in
      // the vector follow at the same offs as in the vector's own plane
      A.xyBase = in + b0 * BLK;
      A.zbBase = in + g.block(t, (z + Lz - 1) % Lz, y0, xb) * BLK;
      A.zfBase = in + g.block(t, (z + 1) % Lz, y0, xb) * BLK;
      A.tbBase = in + g.block((t + Lt - 1) % Lt, z, y0, xb) * BLK;
      A.tfBase = in + g.block((t + 1) % Lt, z, y0, xb) * BLK;
      A.chiBase = chib + b0 * BLK;
      A.oBase = o + b0 * BLK;
      A.gBase = gb + gi * gblk;
      A.clBase = cl ? cl + gi * cblk : nullptr;
      A.alpha = to_ft(alpha);
      A.coeff_s = to_ft(coeff_s);
      A.coeff_t_b = to_ft(t == 0 ? coeff_t_b * t_boundary : coeff_t_b);
      A.coeff_t_f = to_ft(t == Lt - 1 ? coeff_t_f * t_boundary : coeff_t_f);
      A.mu = to_ft(mu);
      A.muinv = to_ft(muinv);
      A.prec_mass_rho = to_ft(rho);
      if (tbc_phases)
        for (int m = 0; m < 4; m++) {
          A.tbc_phases[m][0] = to_ft(tbc_phases[2 * m]);
          A.tbc_phases[m][1] = to_ft(tbc_phases[2 * m + 1]);
        }
      fn(A);
    }
  };
  run();
  for (int rr = 0; rr < reps; rr++) {
    const double t0 = omp_get_wtime();
    run();
    if (seconds) seconds[rr] = omp_get_wtime() - t0;
  }
  unpack_spinor(out, o, g, cb);
  free(in);
  free(chib);
  free(o);
  free(gb);
  free(cl);
  return 0;
}

// ---- faces ----------------------------------------------------------------
// The reference's face bodies driven the way the reference's face
// declarations imply (codegen/jinja/dslash_decl.h.j2:90-115): one call per
// gauge vector of the boundary slice, the vector's spinor blocks addressed
// through xyBase / oBase + offs[row * SOALEN] as in the body calls, and a
// half-spinor buffer per call of nActive lanes per real, [spin 0|1][colour]
// [re|im][nActive] (codegen/data_types.cc:508-543, 587-626):
//   X: the boundary lane of each row, lane 0 (back face, x = 0) or lane
//      SOALEN - 1 (forward face, x = Lx - 1), on the rows whose parity puts
//      a site there, (VECLEN / SOALEN + 1) / 2 of them, compress-stored /
//      expand-loaded under the runtime lane mask;
//   Y: the vector's first row (back, y = 0) or last row (forward,
//      y = Ly - 1), SOALEN lanes (all VECLEN when SOALEN = VECLEN);
//   Z, T: the whole vector.
// The X face loads use the x-direction split loads of the body (lane 0
// from offs[row * SOALEN], lanes 1.. from offs[row * SOALEN + 1] for the
// back face; lanes ..S-2 from offs[row * SOALEN], lane S-1 from
// offs[row * SOALEN + S - 1] for the forward face), pointed here at the
// row's own sites.  Host arrays are lexicographic FP64; a half-spinor array
// hs holds 12 reals ([spin 0|1][colour][re|im]) per lexicographic site and
// is read / written only at the face sites.

namespace {

struct FaceLanes {
  int lane[VECLEN];
  long site[VECLEN];
  int n;
  unsigned int mask;
};

// active lanes of gauge vector (t, z, yg, xb) for face dir on checkerboard
// cb, in lane order (the compress / expand order); n = 0 off the slice
FaceLanes face_lanes(const Lat &g, int dir, int cb, int t, int z, int yg, int xb)
{
  FaceLanes f;
  f.n = 0;
  f.mask = 0;
  const int dim = dir / 2, fw = dir % 2;
  const int Lyg = g.L[1] / NGY;
  const bool on = dim == 0 ? xb == (fw ? g.nxb - 1 : 0)
                  : dim == 1 ? yg == (fw ? Lyg - 1 : 0)
                  : dim == 2 ? z == (fw ? g.L[2] - 1 : 0)
                             : t == (fw ? g.L[3] - 1 : 0);
  if (!on) return f;
  for (int row = 0; row < NGY; row++) {
    const int y = yg * NGY + row;
    if (dim == 1 && NGY > 1 && row != (fw ? NGY - 1 : 0)) continue;
    for (int i = 0; i < S; i++) {
      const int x = g.xof(xb * S + i, y, z, t, cb);
      if (dim == 0 && x != (fw ? g.L[0] - 1 : 0)) continue;
      f.lane[f.n] = row * S + i;
      f.site[f.n] = g.lex(x, y, z, t);
      f.mask |= 1u << (row * S + i);
      f.n++;
    }
  }
  return f;
}

int face_active(int dim)
{
  return dim == 0 ? (NGY + 1) / 2 : dim == 1 ? S : VECLEN;
}

// offs of the vector whose first row segment is block b0, per row
void face_offs(const Lat &g, int dir, int t, int z, int yg, int xb, long b0, int *offs)
{
  for (int row = 0; row < NGY; row++) {
    const int L0 = row * S;
    const int ro = (int)((g.block(t, z, yg * NGY + row, xb) - b0) * BLK);
    offs[L0] = ro;
    if (dir == 0 && S > 1) offs[L0 + 1] = ro + 1;
    if (dir == 1 && S > 1) offs[L0 + S - 1] = ro + S - 1;
  }
}

}  // namespace

// Face pack through face_proj_dir_{plus,minus}: the projected half-spinors
// of the sites of checkerboard pcb on face dir = 2 mu + fb (fb 0: x_mu = 0,
// sent to the back neighbour; 1: x_mu = L_mu - 1, sent forward) into hs.
// Returns 0, or -1 for an unsupported geometry / argument.
int ref_face_pack(int isign, int dir, int pcb, int Lx, int Ly, int Lz, int Lt, const double *psi,
                  double *hs)
{
  Lat g;
  if (!make_lat(g, Lx, Ly, Lz, Lt) || dir < 0 || dir > 7 || (pcb != 0 && pcb != 1) || S < 2) return -1;
  const long nblk = g.nblk();
  FT *in = amalloc<FT>(nblk * BLK);
  FT *buf = amalloc<FT>(12 * VECLEN);
  pack_spinor(in, psi, g, pcb);
  const ref_face_fn fn = ref_face_pack_table[isign > 0 ? 0 : 1];
  const int na = face_active(dir / 2), Lyg = Ly / NGY;
  int rc = 0;
  for (int t = 0; t < Lt && !rc; t++)
    for (int z = 0; z < Lz; z++)
      for (int yg = 0; yg < Lyg; yg++)
        for (int xb = 0; xb < g.nxb; xb++) {
          const FaceLanes fl = face_lanes(g, dir, pcb, t, z, yg, xb);
          if (!fl.n) continue;
          if (fl.n != na) rc = -1;
          if (rc) break;
          const long b0 = g.block(t, z, yg * NGY, xb);
          FaceArgs F;
          memset(&F, 0, sizeof(F));
          face_offs(g, dir, t, z, yg, xb, b0, F.offs);
          F.xyBase = in + b0 * BLK;
          F.outbuf = buf;
          F.mask = fl.mask;
          F.dir = dir;
          memset(buf, 0, 12 * VECLEN * sizeof(FT));
          fn(F);
          for (int k = 0; k < na; k++)
            for (int c = 0; c < 12; c++) hs[fl.site[k] * 12 + c] = from_ft(buf[c * na + k]);
        }
  free(in);
  free(buf);
  return rc;
}

// Face finish through the family's face finish body (kind 0 face_finish_dir,
// 1 face_clov_finish_dir with clov the packed clover, 2 tm_face_finish_dir,
// 3 tm_clov_face_finish_dir with clov the full clover blocks): the sites of
// checkerboard cb on face dir receive the half-spinors hs (read at those
// sites) through the direction-dir link; out (lexicographic, cb sites) is
// read and written in place.  tbc_phases (8 doubles) selects the twisted-
// boundary bodies.  Returns 0, or -1 for an unsupported geometry / argument.
int ref_face_finish(int kind, int isign, int recon, int dir, int cb, int Lx, int Ly, int Lz, int Lt,
                    const double *hs, const double *U, const double *clov, double *out, double beta,
                    double mu, double muinv, const double *tbc_phases)
{
  Lat g;
  if (!make_lat(g, Lx, Ly, Lz, Lt) || kind < 0 || kind > 3 || dir < 0 || dir > 7 ||
      (recon != 12 && recon != 18) || (cb != 0 && cb != 1) || S < 2)
    return -1;
  const long nblk = g.nblk(), ngblk = nblk / NGY;
  const int nrow = recon == 12 ? 2 : 3;
  const long gblk = 8L * nrow * 3 * 2 * VECLEN;
  const long cblk = (kind == 3 ? 144L : 72L) * VECLEN;
  FT *o = amalloc<FT>(nblk * BLK);
  FT *gb = amalloc<FT>(ngblk * gblk);
  FT *cl = (kind == 1 || kind == 3) ? amalloc<FT>(ngblk * cblk) : nullptr;
  FT *buf = amalloc<FT>(12 * VECLEN);
  pack_spinor(o, out, g, cb);
  pack_gauge(gb, U, g, cb, nrow);
  if (kind == 1) pack_clover(cl, clov, g, cb);
  if (kind == 3) pack_full_clover(cl, clov, g, cb);
  const ref_face_fn fn = (tbc_phases ? ref_face_table_tbc : ref_face_table)[kind][isign > 0 ? 0 : 1]
                                                                           [recon == 12 ? 0 : 1];
  const int na = face_active(dir / 2), Lyg = Ly / NGY;
  int rc = 0;
  for (int t = 0; t < Lt && !rc; t++)
    for (int z = 0; z < Lz; z++)
      for (int yg = 0; yg < Lyg; yg++)
        for (int xb = 0; xb < g.nxb; xb++) {
          const FaceLanes fl = face_lanes(g, dir, cb, t, z, yg, xb);
          if (!fl.n) continue;
          if (fl.n != na) rc = -1;
          if (rc) break;
          const long b0 = g.block(t, z, yg * NGY, xb);
          const long gi = g.gblock(t, z, yg * NGY, xb);
          FaceArgs F;
          memset(&F, 0, sizeof(F));
          face_offs(g, -1, t, z, yg, xb, b0, F.offs);
          for (int k = 0; k < na; k++)
            for (int c = 0; c < 12; c++) buf[c * na + k] = to_ft(hs[fl.site[k] * 12 + c]);
          F.oBase = o + b0 * BLK;
          F.gBase = gb + gi * gblk;
          F.clBase = cl ? cl + gi * cblk : nullptr;
          F.inbuf = buf;
          F.mask = fl.mask;
          F.dir = dir;
          F.beta = to_ft(beta);
          F.mu = to_ft(mu);
          F.muinv = to_ft(muinv);
          if (tbc_phases)
            for (int m = 0; m < 4; m++) {
              F.tbc_phases[m][0] = to_ft(tbc_phases[2 * m]);
              F.tbc_phases[m][1] = to_ft(tbc_phases[2 * m + 1]);
            }
          fn(F);
        }
  if (!rc) unpack_spinor(out, o, g, cb);
  free(o);
  free(gb);
  free(cl);
  free(buf);
  return rc;
}

}  // extern "C"
