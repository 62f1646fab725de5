/*
 * qphix_amd.h -- C ABI of the MI355X (gfx950) Wilson-type Dirac operators.
 *
 * The reference (qphix-codegen) exposes its operators as C++ function
 * templates in namespace QPhiX, one call per SIMD block of VECLEN sites of
 * one checkerboard, specialised on <FT, veclen, soalen, compress12, tbc...>
 * (codegen/jinja/dslash_decl.h.j2:6-115, clov_dslash_decl.h.j2:6-113,
 * tm_dslash_decl.h.j2:6-109, tm_clov_dslash_decl.h.j2:6-113).  A GPU cannot
 * be driven per 16-site block, so every entry point below takes a whole
 * checkerboard: the caller's OpenMP block loop becomes the HIP grid.  The
 * template parameters become run-time arguments (precision, 12/18 link
 * reconstruction, isign, checkerboard), the per-block scalars keep their
 * names and meaning (coeff_s, coeff_t_f, coeff_t_b, alpha, mu, muinv,
 * prec_mass_rho, tbc_phases).
 *
 * Conventions
 *  - All field pointers are DEVICE pointers in the library's native layout
 *    (site-fastest, per checkerboard; qphix_amd_field_bytes gives sizes).
 *    The pack/unpack entry points convert from/to a lexicographic FP64
 *    layout on the device.
 *  - Every call is asynchronous on `stream` (a hipStream_t, NULL = default
 *    stream) and returns 0 on success or a negative QPHIX_AMD_E* code.
 *    Nothing ever exits the process (the reference prints and exit(1)s on
 *    a bad direction, codegen/jinja/dslash_spec.h.j2:28-31).
 *  - The caller owns every buffer; kernels write only `out`.
 *  - Calls on distinct contexts or distinct output buffers are thread-safe.
 */
#ifndef QPHIX_AMD_H
#define QPHIX_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define QPHIX_AMD_OK 0
#define QPHIX_AMD_EINVAL -1     /* bad argument (null, range, variant)     */
#define QPHIX_AMD_EGEOM -2      /* unsupported lattice geometry            */
#define QPHIX_AMD_EHIP -3       /* HIP runtime error                       */
#define QPHIX_AMD_ENOMEM -4     /* device allocation failed                */
#define QPHIX_AMD_ENOCONV -5    /* solver did not converge                 */
#define QPHIX_AMD_EALIGN -6     /* pointer not 16-byte aligned             */
#define QPHIX_AMD_ESCRATCH -7   /* lattice too large for the static scratch */

/* ---- enums ------------------------------------------------------------ */
/* storage precision; arithmetic is FP64 for F64 and FP32 otherwise
   (codegen/data_types.h:7-39: FP16 storage with FP32 arithmetic)        */
#define QPHIX_AMD_F64 0
#define QPHIX_AMD_F32 1
#define QPHIX_AMD_F16 2
#define QPHIX_AMD_BF16 3

/* field kinds for qphix_amd_field_bytes */
#define QPHIX_AMD_FIELD_SPINOR 0      /* 24 reals / site                  */
#define QPHIX_AMD_FIELD_GAUGE 1       /* 8 links x 12|18 reals / site     */
#define QPHIX_AMD_FIELD_CLOVER 2      /* 2 x (6 + 15*2) reals / site      */
#define QPHIX_AMD_FIELD_FULLCLOVER 3  /* 2 x 6 x 6 x 2 reals / site       */

typedef struct qphix_amd_ctx qphix_amd_ctx;

/* ---- context / geometry ----------------------------------------------- */
/* Lattice Lx x Ly x Lz x Lt (Lx even; every extent >= 2).  t_boundary is
   the factor (+1 periodic, -1 antiperiodic) applied to T hops that cross
   the t = 0 / t = Lt-1 boundary; the reference takes it through the
   per-block coeff_t_b / coeff_t_f (codegen/dslash.cc:79-86).            */
int qphix_amd_ctx_create(qphix_amd_ctx **ctx, int Lx, int Ly, int Lz, int Lt,
                         double t_boundary);
int qphix_amd_ctx_destroy(qphix_amd_ctx *ctx);
/* Twisted boundary conditions: tbc[mu] != 0 multiplies every forward hop in
   dimension mu by (re,im) = tbc_phases[2mu], tbc_phases[2mu+1] and every
   backward hop by its conjugate (codegen/dslash_common.cc:1491-1524; the
   reference's template flags tbc_x..tbc_t and tbc_phases[4][2]).         */
int qphix_amd_set_tbc(qphix_amd_ctx *ctx, const int tbc[4],
                      const double tbc_phases[8]);
/* Work-group traversal of the operator kernels.  No counterpart in the
   reference (its SIMD loop has no work-group schedule); this is an MI355X
   tuning knob for callers that know their workload better than the default.
     -1  automatic (the default): 1 when one t-slice holds more than 32768
         sites per checkerboard or Lt <= Lz, else 3;
      0  each of the 8 XCDs sweeps its own t-slab;
      1  the XCDs sweep t together, each owning a z-range (Lz % 8 == 0);
      2  t-slabs visited z-major (Lt % 8 == 0);
      3  t-slabs swept two z-halves at a time (Lt % 8 == 0).
   Orders 1-3 need whole 128-site work-groups per (z,t) plane
   ((Lx/2)*Ly % 128 == 0); an order the geometry does not admit runs as 0.
   Results are bit-identical for every order.  qphix_amd_traversal returns
   the order the calls on this context use (-1 for a null context).       */
int qphix_amd_set_traversal(qphix_amd_ctx *ctx, int order);
int qphix_amd_traversal(const qphix_amd_ctx *ctx);
/* Bytes of one checkerboard of a field in native layout.                 */
size_t qphix_amd_field_bytes(const qphix_amd_ctx *ctx, int field, int prec,
                             int recon);
long qphix_amd_sites_cb(const qphix_amd_ctx *ctx);
const char *qphix_amd_error_string(int code);

/* ---- layout converters (device lex FP64 <-> device native) ------------
   Lexicographic layouts, site = ((t*Lz+z)*Ly+y)*Lx+x:
     spinor      [V][4 spin][3 colour][2]
     gauge       [V][4 mu][3][3][2]   U_mu(x) row-major
     clover      [V][2 block][6 diag + 15x2 lower-triangle]
     full clover [V][2 block][6][6][2]
   pack_gauge stores, per site x of checkerboard cb, the 8 hop matrices in
   the reference's direction order d = 2*mu + (0 back | 1 forward), as the
   rows of U_mu(x - mu)^dagger and of U_mu(x)^T (codegen/include/
   decl_common.h:13, 8 links per site; the reference stores every link
   transposed, codegen/dslash_common.cc:1678-1727); recon 12 keeps rows 0, 1
   -- the reference's two stored rows -- and rebuilds the third in the
   kernel (codegen/dslash_common.cc:1657-1676).                          */
int qphix_amd_pack_spinor(qphix_amd_ctx *ctx, void *dst, int prec,
                          const double *src_lex, int cb, void *stream);
int qphix_amd_unpack_spinor(qphix_amd_ctx *ctx, double *dst_lex,
                            const void *src, int prec, int cb, void *stream);
int qphix_amd_pack_gauge(qphix_amd_ctx *ctx, void *dst, int prec, int recon,
                         const double *u_lex, int cb, void *stream);
int qphix_amd_pack_clover(qphix_amd_ctx *ctx, void *dst, int prec,
                          const double *clov_lex, int cb, void *stream);
int qphix_amd_pack_full_clover(qphix_amd_ctx *ctx, void *dst, int prec,
                               const double *fclov_lex, int cb, void *stream);

/* ---- QPhiX block layouts (device QPhiX arrays <-> device native) ------
   The per-checkerboard arrays a QPhiX host hands to the reference's kernels
   (codegen/include/decl_common.h:9-62), with FT given by `ft`:
     FourSpinorBlock  FT[3 colour][4 spin][2][soalen]
     SU3MatrixBlock   FT[8 dir][2|3 rows][3][2][veclen]; entry [d][i][j] is
                      M_d[j][i] (stored transposed, applied as U h forward
                      and U^dagger h back, codegen/dslash_common.cc:1678-1727),
                      with M_{2mu} = U_mu(x - mu) and M_{2mu+1} = U_mu(x);
                      compress12 keeps rows i = 0, 1 and the kernel rebuilds
                      row 2 = conj(row0 x row1) (codegen/dslash_common.cc:
                      1657-1676)
     CloverBlock      {diag1[6][V], off_diag1[15][2][V], diag2, off_diag2}
     FullCloverBlock  {block1[6][6][2][V], block2[6][6][2][V]}
   Block order (QPhiX's checkerboarded geometry) for the site
   x = 2 xh + ((cb + y + z + t) & 1) of checkerboard cb:
     nvecs = (Lx/2) / soalen,  ngy = veclen / soalen,
     Pxy = Ly nvecs + pad_xy,  Pxyz = Lz Pxy + pad_xyz,
     spinor blocks        t Pxyz + z Pxy + y nvecs + xh/soalen,
                          lane xh % soalen  (Lt Pxyz blocks)
     gauge/clover blocks  (t Pxyz + z Pxy)/ngy + (y/ngy) nvecs + xh/soalen,
                          lane (y % ngy) soalen + xh % soalen
                          (Lt Pxyz / ngy blocks)
   so with soalen < veclen one gauge/clover vector spans ngy rows of y, as
   the reference's kernels read it (packed gauges, codegen/data_types.cc:
   218-241).  Requirements: soalen divides Lx/2 and veclen, ngy divides Ly,
   pad_xy and pad_xyz are non-negative multiples of ngy; ft is F64, F32 or
   F16 (IEEE binary16, the reference's `half`, decl_common.h:7).  A host
   built without USE_PACKED_GAUGES / USE_PACKED_CLOVER (gauge and clover
   blocks of soalen lanes, one per spinor block, codegen/data_types.cc:4-30)
   passes veclen = soalen.  Padding
   blocks are neither read nor written.  The gauge conversion rebuilds row 2
   of compressed blocks in FP64 before storing the applied matrices.    */
typedef struct {
  int ft;         /* storage of the QPhiX arrays: QPHIX_AMD_F64/F32/F16   */
  int veclen;     /* VECLEN, lanes of SU3MatrixBlock / CloverBlock        */
  int soalen;     /* SOALEN, lanes of FourSpinorBlock                     */
  int compress12; /* SU3MatrixBlock holds 2 rows (1) or 3 rows (0)        */
  int pad_xy;     /* spinor blocks of padding after each xy plane         */
  int pad_xyz;    /* spinor blocks of padding after each xyz volume       */
} qphix_amd_qphix_layout;

/* Bytes of one checkerboard array of `field` in this QPhiX layout (0 if the
   layout is invalid for the context's lattice).                          */
size_t qphix_amd_qphix_field_bytes(const qphix_amd_ctx *ctx,
                                   const qphix_amd_qphix_layout *lay,
                                   int field);
int qphix_amd_spinor_from_qphix(qphix_amd_ctx *ctx,
                                const qphix_amd_qphix_layout *lay, void *dst,
                                int prec, const void *src, int cb,
                                void *stream);
int qphix_amd_spinor_to_qphix(qphix_amd_ctx *ctx,
                              const qphix_amd_qphix_layout *lay, void *dst,
                              const void *src, int prec, int cb,
                              void *stream);
int qphix_amd_gauge_from_qphix(qphix_amd_ctx *ctx,
                               const qphix_amd_qphix_layout *lay, void *dst,
                               int prec, int recon, const void *src, int cb,
                               void *stream);
int qphix_amd_clover_from_qphix(qphix_amd_ctx *ctx,
                                const qphix_amd_qphix_layout *lay, void *dst,
                                int prec, const void *src, int cb,
                                void *stream);
int qphix_amd_full_clover_from_qphix(qphix_amd_ctx *ctx,
                                     const qphix_amd_qphix_layout *lay,
                                     void *dst, int prec, const void *src,
                                     int cb, void *stream);

/* ---- operators on QPhiX FourSpinorBlock arrays, in place ---------------
   The operators below, with in, chi and out read and written where a QPhiX
   host keeps them: FourSpinorBlock arrays of the layout `lay` (FT = lay->ft,
   F64, F32 or F16; soalen a power of two; the block order above, padding
   included) instead of native spinors -- no conversion pass per call.  The
   links and clover blocks are native fields in the same precision
   (converted once, e.g. with qphix_amd_gauge_from_qphix), `recon` as
   there.  Same semantics, arguments and errors as the native entries of
   the same family; QPhiX arrays need FT alignment only (FP32 and FP16
   arrays with soalen >= 4 and Lx/2 a multiple of 4, aligned to 16 and 8
   bytes, are read and written as vectors by lane quad, and 16-byte aligned
   FP64 arrays with Lx/2 even have their x/y rows staged from such vectors;
   the others are read element by element; every form gives the same bits).
   accumulate (device, optional): one byte per site of the output array, in
   its site order (block * soalen + lane, padding slots included and
   ignored); bit d = 2 mu + fb set accumulates direction d at that site,
   clear drops it.  This is the reference's per-call accumulate[8]
   (codegen/jinja/dslash_decl.h.j2:6-43) transposed to one mask per site:
   QPhiX sets lane bits for the x hops (and the y hops when soalen <
   veclen) and whole-vector flags otherwise (codegen/dslash.cc:109,248,
   285-290,570-578); any such pattern is a per-site mask.  NULL accumulates
   every direction.  The context's split dimensions (qphix_amd_set_comm)
   apply as well.
   Each replaces the per-block template named in the native entry below:
     qphix_amd_qphix_dslash                 dslash_{plus,minus}_vec
     qphix_amd_qphix_dslash_achimbdpsi      dslash_achimbdpsi_{plus,minus}_vec
     qphix_amd_qphix_clov_dslash            clov_dslash_{plus,minus}_vec
     qphix_amd_qphix_clov_dslash_achimbdpsi clov_dslash_achimbdpsi_{plus,minus}_vec
     qphix_amd_qphix_tm_dslash              tm_dslash_{plus,minus}_vec
     qphix_amd_qphix_tm_dslash_achimbdpsi   tm_dslash_achimbdpsi_{plus,minus}_vec
     qphix_amd_qphix_tm_clov_dslash         tm_clov_dslash_{plus,minus}_vec
     qphix_amd_qphix_tm_clov_dslash_achimbdpsi
                                            tm_clov_dslash_achimbdpsi_{plus,minus}_vec
   (codegen/jinja/{,clov_,tm_,tm_clov_}dslash_decl.h.j2:6-94).            */
int qphix_amd_qphix_dslash(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay,
                           int recon, int isign, int cb, void *out,
                           const void *in, const void *gauge,
                           const unsigned char *accumulate, double coeff_s,
                           double coeff_t_f, double coeff_t_b, void *stream);
int qphix_amd_qphix_dslash_achimbdpsi(qphix_amd_ctx *ctx,
                                      const qphix_amd_qphix_layout *lay,
                                      int recon, int isign, int cb, void *out,
                                      const void *in, const void *chi,
                                      const void *gauge,
                                      const unsigned char *accumulate,
                                      double alpha, double coeff_s,
                                      double coeff_t_f, double coeff_t_b,
                                      void *stream);
int qphix_amd_qphix_clov_dslash(qphix_amd_ctx *ctx,
                                const qphix_amd_qphix_layout *lay, int recon,
                                int isign, int cb, void *out, const void *in,
                                const void *gauge, const void *clover,
                                const unsigned char *accumulate,
                                double coeff_s, double coeff_t_f,
                                double coeff_t_b, void *stream);
int qphix_amd_qphix_clov_dslash_achimbdpsi(
    qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon,
    int isign, int cb, void *out, const void *in, const void *chi,
    const void *gauge, const void *clover, const unsigned char *accumulate,
    double coeff_s, double coeff_t_f, double coeff_t_b, double prec_mass_rho,
    void *stream);
int qphix_amd_qphix_tm_dslash(qphix_amd_ctx *ctx,
                              const qphix_amd_qphix_layout *lay, int recon,
                              int isign, int cb, void *out, const void *in,
                              const void *gauge,
                              const unsigned char *accumulate, double coeff_s,
                              double coeff_t_f, double coeff_t_b, double mu,
                              double muinv, void *stream);
int qphix_amd_qphix_tm_dslash_achimbdpsi(
    qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon,
    int isign, int cb, void *out, const void *in, const void *chi,
    const void *gauge, const unsigned char *accumulate, double alpha,
    double coeff_s, double coeff_t_f, double coeff_t_b, double mu,
    void *stream);
int qphix_amd_qphix_tm_clov_dslash(qphix_amd_ctx *ctx,
                                   const qphix_amd_qphix_layout *lay,
                                   int recon, int isign, int cb, void *out,
                                   const void *in, const void *gauge,
                                   const void *fclover,
                                   const unsigned char *accumulate,
                                   double coeff_s, double coeff_t_f,
                                   double coeff_t_b, void *stream);
int qphix_amd_qphix_tm_clov_dslash_achimbdpsi(
    qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon,
    int isign, int cb, void *out, const void *in, const void *chi,
    const void *gauge, const void *fclover, const unsigned char *accumulate,
    double coeff_s, double coeff_t_f, double coeff_t_b, double prec_mass_rho,
    void *stream);

/* ---- operators ----------------------------------------------------------
   out (checkerboard cb) from in (checkerboard 1-cb).  hop(x) =
     sum_mu b_mu(x) (1 + s g_mu) U_mu(x-mu)^+ psi(x-mu)
          + f_mu(x) (1 - s g_mu) U_mu(x) psi(x+mu),     s = isign = +1|-1
   in the DeGrand-Rossi basis, with b = f = coeff_s for x,y,z and
   coeff_t_b / coeff_t_f for t (times t_boundary across the boundary).   */

/* replaces QPhiX::dslash_{plus,minus}_vec (codegen/jinja/dslash_decl.h.j2:6-43)
   out = hop                                                              */
int qphix_amd_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign,
                     int cb, void *out, const void *in, const void *gauge,
                     double coeff_s, double coeff_t_f, double coeff_t_b,
                     void *stream);
/* replaces QPhiX::dslash_achimbdpsi_{plus,minus}_vec
   (codegen/jinja/dslash_decl.h.j2:45-89): out = alpha chi - hop          */
int qphix_amd_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon,
                                int isign, int cb, void *out, const void *in,
                                const void *chi, const void *gauge,
                                double alpha, double coeff_s,
                                double coeff_t_f, double coeff_t_b,
                                void *stream);
/* replaces QPhiX::clov_dslash_{plus,minus}_vec
   (codegen/jinja/clov_dslash_decl.h.j2:6-46): out = A^-1 hop, where the
   caller passes the packed Hermitian A^-1 as `clover`                    */
int qphix_amd_clov_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign,
                          int cb, void *out, const void *in,
                          const void *gauge, const void *clover,
                          double coeff_s, double coeff_t_f, double coeff_t_b,
                          void *stream);
/* replaces QPhiX::clov_dslash_achimbdpsi_{plus,minus}_vec
   (codegen/jinja/clov_dslash_decl.h.j2:48-94):
   out = A chi + isign*i*prec_mass_rho*g5 chi - hop                        */
int qphix_amd_clov_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon,
                                     int isign, int cb, void *out,
                                     const void *in, const void *chi,
                                     const void *gauge, const void *clover,
                                     double coeff_s, double coeff_t_f,
                                     double coeff_t_b, double prec_mass_rho,
                                     void *stream);
/* replaces QPhiX::tm_dslash_{plus,minus}_vec
   (codegen/jinja/tm_dslash_decl.h.j2:6-45):
   out = muinv (1 - isign*i*mu*g5) hop, mu = mu_phys/alpha,
   muinv = alpha/(alpha^2 + mu_phys^2) (codegen/dslash_common.cc:1256-1266) */
int qphix_amd_tm_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign,
                        int cb, void *out, const void *in, const void *gauge,
                        double coeff_s, double coeff_t_f, double coeff_t_b,
                        double mu, double muinv, void *stream);
/* replaces QPhiX::tm_dslash_achimbdpsi_{plus,minus}_vec
   (codegen/jinja/tm_dslash_decl.h.j2:47-91):
   out = alpha (1 + isign*i*mu*g5) chi - hop                              */
int qphix_amd_tm_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon,
                                   int isign, int cb, void *out,
                                   const void *in, const void *chi,
                                   const void *gauge, double alpha,
                                   double coeff_s, double coeff_t_f,
                                   double coeff_t_b, double mu, void *stream);
/* replaces QPhiX::tm_clov_dslash_{plus,minus}_vec
   (codegen/jinja/tm_clov_dslash_decl.h.j2:6-46): out = B hop, B = two full
   complex 6x6 blocks passed as `fclover`                                 */
int qphix_amd_tm_clov_dslash(qphix_amd_ctx *ctx, int prec, int recon,
                             int isign, int cb, void *out, const void *in,
                             const void *gauge, const void *fclover,
                             double coeff_s, double coeff_t_f,
                             double coeff_t_b, void *stream);
/* replaces QPhiX::tm_clov_dslash_achimbdpsi_{plus,minus}_vec
   (codegen/jinja/tm_clov_dslash_decl.h.j2:48-94):
   out = B chi + isign*i*prec_mass_rho*g5 chi - hop                        */
int qphix_amd_tm_clov_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec,
                                        int recon, int isign, int cb,
                                        void *out, const void *in,
                                        const void *chi, const void *gauge,
                                        const void *fclover, double coeff_s,
                                        double coeff_t_f, double coeff_t_b,
                                        double prec_mass_rho, void *stream);

/* ---- non-degenerate twisted mass -----------------------------------------
   The reference declares an ndtm_ family (ndtm_dslash, ndtm_clov_dslash;
   codegen/twisted_mass_enum.h:19-41, codegen/jinja/kernels-inactive.js:1-4)
   but leaves it out of the generated variants (twisted_mass_enum.h:31-33),
   and its bodies are copies of the degenerate ones (codegen/
   dslash_common.cc:1599-1603,1643-1650).  These entry points give exactly
   that: the twisted-mass / twisted-mass-clover operators under ndtm_
   names, same arguments.                                                 */
int qphix_amd_ndtm_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign,
                          int cb, void *out, const void *in, const void *gauge,
                          double coeff_s, double coeff_t_f, double coeff_t_b,
                          double mu, double muinv, void *stream);
int qphix_amd_ndtm_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon,
                                     int isign, int cb, void *out,
                                     const void *in, const void *chi,
                                     const void *gauge, double alpha,
                                     double coeff_s, double coeff_t_f,
                                     double coeff_t_b, double mu,
                                     void *stream);
int qphix_amd_ndtm_clov_dslash(qphix_amd_ctx *ctx, int prec, int recon,
                               int isign, int cb, void *out, const void *in,
                               const void *gauge, const void *fclover,
                               double coeff_s, double coeff_t_f,
                               double coeff_t_b, void *stream);
int qphix_amd_ndtm_clov_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec,
                                          int recon, int isign, int cb,
                                          void *out, const void *in,
                                          const void *chi, const void *gauge,
                                          const void *fclover, double coeff_s,
                                          double coeff_t_f, double coeff_t_b,
                                          double prec_mass_rho, void *stream);

/* ---- faces: the lattice split over several GPUs --------------------------
   qphix_amd_set_comm marks split dimensions: comm[mu] != 0 drops, from every
   operator call above, the hops that wrap around the local lattice in mu
   (they belong to neighbouring GPUs), which gives the interior part the
   reference computes with accumulate[d] = 0 on boundary blocks
   (codegen/dslash.cc:248,368; the time-boundary sign then no longer applies
   inside the call).  The faces complete it:
     face_pack   half-spinors of the input (checkerboard 1-cb) on the slice
                 x_dim = 0 (fb 0, for the backward neighbour) or L_dim - 1
                 (fb 1, forward), projected with the receiving hop's
                 projector, (1 - isign g_dim) for fb 0 and (1 + isign g_dim)
                 for fb 1 (face_proj_dir_{plus,minus},
                 codegen/jinja/dslash_decl.h.j2:91-100, codegen/dslash.cc:
                 433-458);
     *_finish    out sites (checkerboard cb) on x_dim = 0 (fb 0, buffer from
                 the backward neighbour, U_dim(x - dim)^dagger) or L_dim - 1
                 (fb 1, from the forward neighbour, U_dim(x)): r = reconstructed
                 beta * U h (twisted-BC phase of the context included), then
                   face_finish          out -= r
                   clov_face_finish     out += invclover r
                   tm_face_finish       out -= muinv (1 - isign i mu g5) r
                   tm_clov_face_finish  out += fclover r
                 (face_finish_dir_*, face_clov_finish_dir_*,
                 tm_face_finish_dir_*, tm_clov_face_finish_dir_*:
                 codegen/jinja/{,clov_,tm_,tm_clov_}dslash_decl.h.j2, codegen/dslash.cc:463-553,
                 codegen/dslash_common.cc:1947-1993).  So completing a plain
                 or twisted-mass dslash passes beta = -coeff, the clover ones
                 beta = +coeff, and any achimbdpsi uses face_finish with
                 beta = +coeff; the caller folds the time-boundary sign into
                 beta on the GPUs at the global time boundary.
   A face buffer holds qphix_amd_face_sites(ctx, dim) half-spinors of 12
   reals ([spin][colour][re|im], codegen/data_types.cc:508-543) in the
   storage precision, site-fastest in 16-byte vectors (bytes = sites * 12 *
   sizeof(storage)).  Face site f runs over the other coordinates:
     x: (t Lz + z) Ly/2 + y/2   y: (t Lz + z) Lx/2 + x/2
     z: (t Ly + y) Lx/2 + x/2   t: (z Ly + y) Lx/2 + x/2
   identical on both sides of a cut.  fb = 2 (no reference counterpart)
   packs or finishes both faces of dim in one launch, with buf holding the
   fb 0 face and then the fb 1 face (2 * sites * 12 reals): each face is a
   small launch bound by its latency, so one launch for both halves it on
   a host's critical path.  Both faces then take one beta, so a GPU at the
   global time boundary finishes its t faces one by one.  Operators built on a context
   (qphix_amd_eo_op_create, qphix_amd_dslash_force) need comm all zero.   */
int qphix_amd_set_comm(qphix_amd_ctx *ctx, const int comm[4]);
long qphix_amd_face_sites(const qphix_amd_ctx *ctx, int dim);
int qphix_amd_face_pack(qphix_amd_ctx *ctx, int prec, int isign, int cb,
                        int dim, int fb, void *buf, const void *in,
                        void *stream);
int qphix_amd_face_finish(qphix_amd_ctx *ctx, int prec, int recon, int isign,
                          int cb, int dim, int fb, void *out, const void *buf,
                          const void *gauge, double beta, void *stream);
int qphix_amd_clov_face_finish(qphix_amd_ctx *ctx, int prec, int recon,
                               int isign, int cb, int dim, int fb, void *out,
                               const void *buf, const void *gauge,
                               const void *invclover, double beta,
                               void *stream);
int qphix_amd_tm_face_finish(qphix_amd_ctx *ctx, int prec, int recon,
                             int isign, int cb, int dim, int fb, void *out,
                             const void *buf, const void *gauge, double beta,
                             double mu, double muinv, void *stream);
int qphix_amd_tm_clov_face_finish(qphix_amd_ctx *ctx, int prec, int recon,
                                  int isign, int cb, int dim, int fb,
                                  void *out, const void *buf,
                                  const void *gauge, const void *fclover,
                                  double beta, void *stream);
/* The same on QPhiX FourSpinorBlock arrays in place (FT = lay->ft, the
   layouts of the qphix_amd_qphix_* operators): face_pack reads `in`, the
   finishes read and write `out` where a QPhiX host keeps them, so a host
   that splits its lattice completes the in-place operators without a
   conversion pass.  Buffers, links and clover blocks as above; the arrays
   need FT alignment only.  These are the reference's face templates on
   its own arrays (face_proj_dir_{plus,minus}, face_finish_dir_*,
   face_clov_finish_dir_*, tm_face_finish_dir_*, tm_clov_face_finish_dir_*:
   codegen/jinja/{,clov_,tm_,tm_clov_}dslash_decl.h.j2, codegen/dslash.cc:
   433-553), one call per face instead of one per block.                 */
int qphix_amd_qphix_face_pack(qphix_amd_ctx *ctx,
                              const qphix_amd_qphix_layout *lay, int isign,
                              int cb, int dim, int fb, void *buf,
                              const void *in, void *stream);
int qphix_amd_qphix_face_finish(qphix_amd_ctx *ctx,
                                const qphix_amd_qphix_layout *lay, int recon,
                                int isign, int cb, int dim, int fb, void *out,
                                const void *buf, const void *gauge,
                                double beta, void *stream);
int qphix_amd_qphix_clov_face_finish(qphix_amd_ctx *ctx,
                                     const qphix_amd_qphix_layout *lay,
                                     int recon, int isign, int cb, int dim,
                                     int fb, void *out, const void *buf,
                                     const void *gauge, const void *invclover,
                                     double beta, void *stream);
int qphix_amd_qphix_tm_face_finish(qphix_amd_ctx *ctx,
                                   const qphix_amd_qphix_layout *lay,
                                   int recon, int isign, int cb, int dim,
                                   int fb, void *out, const void *buf,
                                   const void *gauge, double beta, double mu,
                                   double muinv, void *stream);
int qphix_amd_qphix_tm_clov_face_finish(qphix_amd_ctx *ctx,
                                        const qphix_amd_qphix_layout *lay,
                                        int recon, int isign, int cb, int dim,
                                        int fb, void *out, const void *buf,
                                        const void *gauge,
                                        const void *fclover, double beta,
                                        void *stream);

/* ---- differentiation (no counterpart in the reference) ------------------
   For L = Re<chi, D psi>, D = the hop above with output checkerboard cb:
   dL/dpsi = D^dagger chi.  D_isign^dagger = D_{-isign} with coeff_t_f and
   coeff_t_b exchanged (gamma_mu Hermitian), landing on checkerboard 1-cb;
   `gauge` must be packed for checkerboard 1-cb.                         */
int qphix_amd_dslash_adjoint(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb,
                             void *grad_psi, const void *chi, const void *gauge,
                             double coeff_s, double coeff_t_f, double coeff_t_b,
                             void *stream);
/* Link gradient of L: accumulates scale * (per link) into the FP64
   lexicographic field force_lex [V][4][3][3][2]:
     project = 1: TA(U G), the traceless anti-Hermitian su(3) force, with
                  d/de L(exp(i e T) U) = Re tr(i T F) for Hermitian T;
     project = 0: G^dagger, the gradient with respect to the real and
                  imaginary parts of the link entries;
   where L = sum over links Re tr(U G) (outer products of psi and the
   projected chi).  chi, psi: native spinors (FP64 or FP32) on cb, 1-cb;
   u_lex: FP64 lexicographic links.                                      */
int qphix_amd_dslash_force(qphix_amd_ctx *ctx, int prec, int isign, int cb, double *force_lex,
                           const void *chi, const void *psi, const double *u_lex,
                           double coeff_s, double coeff_t_f, double coeff_t_b, double scale,
                           int project, void *stream);

/* ---- link reunitarisation (FP64, in place, lexicographic links) --------
   row 0 normalised, row 1 Gram-Schmidt against row 0 and normalised,
   row 2 = conj(row0 x row1); no counterpart in the reference.           */
int qphix_amd_reunitarize(qphix_amd_ctx *ctx, double *u_lex, void *stream);

/* ---- clover term from links (no counterpart in the reference, whose
   kernels take A or A^-1 as input) ------------------------------------
   A(x) = alpha + (csw/2) sum_{mu<nu} sigma_{mu nu} F_{mu nu}(x),
   sigma_{mu nu} = (i/2)[g_mu, g_nu], F_{mu nu} = traceless part of
   (Q - Q^dagger)/(8i), Q = sum of the four plaquette leaves at x.
   All FP64 on lexicographic device arrays.                              */
int qphix_amd_clover_from_links(qphix_amd_ctx *ctx, double *clov_lex, const double *u_lex,
                                double alpha, double csw, void *stream);
/* packed Hermitian blocks -> packed Hermitian inverse blocks            */
int qphix_amd_clover_invert(qphix_amd_ctx *ctx, double *inv_lex, const double *clov_lex,
                            void *stream);
/* full complex blocks of A + i*sign*mu*g5 (invert = 0) or of its inverse  */
int qphix_amd_tmclover_blocks(qphix_amd_ctx *ctx, double *out_lex, const double *clov_lex,
                              double mu, int sign, int invert, void *stream);

/* ---- even-odd operator and CG (no counterpart in the reference; these
   are the callers QPhiX builds on top of its generated kernels) ----------
   M = F_oo - 1/4 D_oe F_ee^-1 D_eo on the odd checkerboard (cb 1),
   F = A + i mu g5 (mu = 0: Wilson-clover with packed Hermitian blocks).
   The operator keeps an FP64 copy (12-real links) and a sloppy copy in
   prec_sloppy / recon_sloppy storage; it allocates device memory.  The
   force and CG calls on an operator keep their workspaces in it (grow-only,
   released by qphix_amd_eo_op_destroy), so repeated calls allocate
   nothing; calls on one operator must not overlap.                      */
typedef struct qphix_amd_eo_op qphix_amd_eo_op;
int qphix_amd_eo_op_create(qphix_amd_eo_op **op, qphix_amd_ctx *ctx, const double *u_lex,
                           const double *clov_lex, double mu, int prec_sloppy,
                           int recon_sloppy, void *stream);
/* Frees the operator; like hipFree it waits for all work on the device. */
int qphix_amd_eo_op_destroy(qphix_amd_eo_op *op);
/* out = M in (dagger = 0) or M^dagger in (dagger = 1); prec is
   QPHIX_AMD_F64 or the operator's sloppy precision; cb-1 native spinors. */
int qphix_amd_eo_op_apply(qphix_amd_eo_op *op, int prec, int dagger, void *out, const void *in,
                          void *stream);

typedef struct {
  double rsd_target; /* stop at |b - M^dag M x| / |b| < rsd_target (FP64)   */
  int max_iters;
  double delta;      /* reliable update when |r| < delta * max|r|; <= 0 picks
                        0.01 (FP32 sloppy), 0.05 (FP16) or 0.1 (BF16)       */
  int mixed;         /* 0: CG in FP64; 1: sloppy iterations + reliable updates */
} qphix_amd_cg_params;

typedef struct {
  int iters;            /* CG iterations                                  */
  int reliable_updates; /* FP64 true-residual recomputations             */
  double rsd_final;     /* final true relative residual (FP64)            */
  double seconds;       /* wall time of the solve (host clock)            */
  long mv_apps;         /* applications of M or M^dagger                 */
} qphix_amd_cg_result;

/* Hopping-term part of the gradient of S = phi^dag (M^dag M)^-1 phi at
   x = (M^dag M)^-1 phi (FP64 native, cb 1):
     dS = 1/2 Re[<Y, dD_oe Z> + <W, dD_eo X>],  Y = M X,
     Z = F_ee^-1 D_eo X,  W = (F_ee^-1)^dag D_oe^dag Y,
   accumulated as scale * (per link) into force_lex like
   qphix_amd_dslash_force (project = 1: su(3) force, 0: link gradient).
   This is the whole gradient when F does not depend on the links (csw = 0);
   qphix_amd_eo_op_force_clover adds the clover-term derivative.         */
int qphix_amd_eo_op_force(qphix_amd_eo_op *op, double *force_lex, const void *x,
                          const double *u_lex, double scale, int project, void *stream);

/* The whole pseudofermion force of the operator when its clover field is
   A = alpha + (csw/2) sum sigma F from qphix_amd_clover_from_links of the
   same links (any alpha; the twisted mass of the operator does not depend
   on the links): the hopping term of qphix_amd_eo_op_force plus the clover
   term dS = -2 Re <Y, dA_oo X> - 1/2 Re <W, dA_ee Z>.                   */
int qphix_amd_eo_op_force_clover(qphix_amd_eo_op *op, double *force_lex,
                                 const void *x, const double *u_lex,
                                 double csw, double scale, int project,
                                 void *stream);

/* Link derivative of the clover term on its own: accumulates into
   force_lex the gradient of L = Re sum_x w(x) chi(x)^dag A(x) psi(x), A as
   built by qphix_amd_clover_from_links (csw; alpha drops out), w(x) =
   scale_even / scale_odd by the parity of x; chi, psi lexicographic FP64
   spinors on the whole lattice.  project as for qphix_amd_dslash_force.
   Every link lies on 24 plaquette leaves (2 squares in each of its 3
   planes, each a leaf of its 4 corners), all gathered by one thread.    */
int qphix_amd_clover_force(qphix_amd_ctx *ctx, double *force_lex,
                           const double *chi_lex, const double *psi_lex,
                           const double *u_lex, double csw, double scale_even,
                           double scale_odd, int project, void *stream);

/* Scratch-free qphix_amd_clover_force: the same arguments and result
   (the same kernels), with its workspace -- the clover weights, 48 doubles
   per lattice site -- in a static device array of
   QPHIX_AMD_NOALLOC_SCRATCH_BYTES (64 MiB) per GPU instead of a per-call
   hipMalloc, so the call allocates nothing and does not synchronise.  No
   counterpart in the reference; it is for callers that cannot allocate
   device memory (code under HIP graph capture, tight HMC loops, harnesses
   that forbid hipMalloc).  Lattices of more than 174,762 sites (2 * vcb *
   48 * 8 bytes > 64 MiB; e.g. 16^3x32 fits, 16^3x48 does not) return
   QPHIX_AMD_ESCRATCH without launching anything.  Every call on one GPU
   shares that GPU's array: calls that overlap on one GPU race.           */
#define QPHIX_AMD_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int qphix_amd_clover_force_noalloc(qphix_amd_ctx *ctx, double *force_lex,
                                   const double *chi_lex,
                                   const double *psi_lex, const double *u_lex,
                                   double csw, double scale_even,
                                   double scale_odd, int project, void *stream);

/* Solves M^dag M x = b; x, b are FP64 native spinors on cb 1.  Returns
   QPHIX_AMD_ENOCONV (with res filled) when max_iters is reached.         */
int qphix_amd_cg_solve(qphix_amd_eo_op *op, void *x, const void *b,
                       const qphix_amd_cg_params *params, qphix_amd_cg_result *res,
                       void *stream);

#ifdef __cplusplus
}
#endif

#endif /* QPHIX_AMD_H */
