// capi.hip -- the extern "C" entry points declared in include/qphix_amd.h.
// Every entry validates its arguments, never exits, and returns a status.
#include <new>

#include "dslash_kernel.h"
#include "internal.h"

using qamd::DslashCall;

namespace {

int hip_status(hipError_t e) { return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP; }

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

bool valid_prec(int prec) { return prec >= QPHIX_AMD_F64 && prec <= QPHIX_AMD_BF16; }

int storage_bytes(int prec) { return prec == QPHIX_AMD_F64 ? 8 : prec == QPHIX_AMD_F32 ? 4 : 2; }

int run(qphix_amd_ctx *ctx, const DslashCall &c, void *stream)
{
  if (!ctx) return QPHIX_AMD_EINVAL;
  if (!valid_prec(c.prec) || (c.recon != 12 && c.recon != 18)) return QPHIX_AMD_EINVAL;
  if ((c.isign != 1 && c.isign != -1) || (c.cb != 0 && c.cb != 1)) return QPHIX_AMD_EINVAL;
  if (!c.out || !c.in || !c.gauge) return QPHIX_AMD_EINVAL;
  if (c.out == c.in) return QPHIX_AMD_EINVAL;
  if (c.achi && !c.chi) return QPHIX_AMD_EINVAL;
  const bool need_clov = c.op == qamd::OP_CLOVER || c.op == qamd::OP_TMCLOVER;
  if (need_clov && !c.clov) return QPHIX_AMD_EINVAL;
  if (!aligned16(c.out) || !aligned16(c.in) || !aligned16(c.gauge) ||
      (c.chi && !aligned16(c.chi)) || (c.clov && !aligned16(c.clov)))
    return QPHIX_AMD_EALIGN;
  return hip_status(qamd::run_dslash_any(ctx, c, static_cast<hipStream_t>(stream)));
}

// An operator call on QPhiX FourSpinorBlock arrays: in, chi and out in the
// layout lay (FT = lay->ft, also the storage of the native links and clover
// blocks), with the optional per-site direction mask acc.
int run_qphix(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, const unsigned char *acc, DslashCall c,
              void *stream)
{
  if (!ctx || !lay) return QPHIX_AMD_EINVAL;
  qamd::QSpinorGeom q;
  if (!qamd::qphix_spinor_geometry(ctx, lay, q)) return QPHIX_AMD_EINVAL;
  if ((c.isign != 1 && c.isign != -1) || (c.cb != 0 && c.cb != 1)) return QPHIX_AMD_EINVAL;
  if (c.recon != 12 && c.recon != 18) return QPHIX_AMD_EINVAL;
  if (!c.out || !c.in || !c.gauge || c.out == c.in || (c.achi && (!c.chi || c.chi == c.out)))
    return QPHIX_AMD_EINVAL;
  const bool need_clov = c.op == qamd::OP_CLOVER || c.op == qamd::OP_TMCLOVER;
  if (need_clov && !c.clov) return QPHIX_AMD_EINVAL;
  // QPhiX arrays are read element by element: FT alignment suffices; the
  // native links and clover blocks are read as 16-byte vectors
  const uintptr_t ftm = (uintptr_t)storage_bytes(lay->ft) - 1;
  if ((reinterpret_cast<uintptr_t>(c.out) & ftm) || (reinterpret_cast<uintptr_t>(c.in) & ftm) ||
      (c.chi && (reinterpret_cast<uintptr_t>(c.chi) & ftm)) || !aligned16(c.gauge) ||
      (c.clov && !aligned16(c.clov)))
    return QPHIX_AMD_EALIGN;
  c.prec = lay->ft;
  // vectors per lane group (16 bytes; 8 for FP16) where the layout and the
  // alignment allow them, element by element otherwise
  const int group = lay->ft == QPHIX_AMD_F64 ? 2 : 4;
  const uintptr_t vm = (uintptr_t)(group * storage_bytes(lay->ft)) - 1;
  const bool vec = (1 << q.sh) >= group && ctx->g.Lxh % group == 0 && !(reinterpret_cast<uintptr_t>(c.out) & vm) &&
                   !(reinterpret_cast<uintptr_t>(c.in) & vm) &&
                   (!c.chi || !(reinterpret_cast<uintptr_t>(c.chi) & vm));
  c.spl = vec ? 2 : 1;
  c.q_sh = q.sh;
  c.q_nvecs = q.nvecs;
  c.q_pxy = q.pxy;
  c.q_pxyz = q.pxyz;
  c.acc = acc;
  return hip_status(qamd::run_dslash_any(ctx, c, static_cast<hipStream_t>(stream)));
}

DslashCall make_call(int prec, int recon, int op, bool achi, int isign, int cb, void *out,
                     const void *in, const void *chi, const void *gauge, const void *clov,
                     double cs, double ctf, double ctb)
{
  DslashCall c{};
  c.prec = prec;
  c.recon = recon;
  c.op = op;
  c.achi = achi;
  c.isign = isign;
  c.cb = cb;
  c.out = out;
  c.in = in;
  c.chi = chi;
  c.gauge = gauge;
  c.clov = clov;
  c.coeff_s = cs;
  c.coeff_tf = ctf;
  c.coeff_tb = ctb;
  return c;
}

}  // namespace

namespace qamd {
hipError_t run_dslash_any(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st)
{
  switch (c.prec) {
    case QPHIX_AMD_F64: return run_dslash_f64(ctx, c, st);
    case QPHIX_AMD_F32: return run_dslash_f32(ctx, c, st);
    case QPHIX_AMD_F16: return run_dslash_f16(ctx, c, st);
    default: return run_dslash_bf16(ctx, c, st);
  }
}
int dslash_traversal(const qphix_amd_ctx *ctx)
{
  return dslash_traversal_for(ctx->g, ctx->traversal >= 0 ? ctx->traversal : QAMD_BLOCK_ORDER);
}
}  // namespace qamd

extern "C" {

int qphix_amd_ctx_create(qphix_amd_ctx **ctx, int Lx, int Ly, int Lz, int Lt, double t_boundary)
{
  if (!ctx) return QPHIX_AMD_EINVAL;
  *ctx = nullptr;
  // Every extent even: the wrap-around hop must change the site parity.
  if (Lx < 2 || Ly < 2 || Lz < 2 || Lt < 2) return QPHIX_AMD_EGEOM;
  if ((Lx | Ly | Lz | Lt) & 1) return QPHIX_AMD_EGEOM;
  // Every field index must fit 32 bits: 72 reals x sites per checkerboard
  // < 2^31.  The bound is applied to each partial product, so no product of
  // extents can overflow on the way (each factor is below 2^31 and each
  // partial product below 2^31 / 72).
  const long kMaxVcb = ((1L << 31) - 1) / 72;
  long vcb = Lx / 2;
  for (int L : {Ly, Lz, Lt}) {
    if (vcb > kMaxVcb / L) return QPHIX_AMD_EGEOM;
    vcb *= L;
  }
  qphix_amd_ctx *c = new (std::nothrow) qphix_amd_ctx();
  if (!c) return QPHIX_AMD_ENOMEM;
  c->g.Lx = Lx;
  c->g.Ly = Ly;
  c->g.Lz = Lz;
  c->g.Lt = Lt;
  c->g.Lxh = Lx / 2;
  c->g.vcb = (int)vcb;
  c->g.sy = Lx / 2;
  c->g.sz = Lx / 2 * Ly;
  c->g.st = Lx / 2 * Ly * Lz;
  c->g.fxh = qamd::make_fastdiv((unsigned)(Lx / 2));
  c->g.fy = qamd::make_fastdiv((unsigned)Ly);
  c->g.fz = qamd::make_fastdiv((unsigned)Lz);
  c->tbound = t_boundary;
  c->comm_mask = 0;
  c->traversal = -1;
  c->tbc_mask = 0;
  for (int m = 0; m < 4; m++) {
    c->tbc_re[m] = 1.0;
    c->tbc_im[m] = 0.0;
  }
  *ctx = c;
  return QPHIX_AMD_OK;
}

int qphix_amd_ctx_destroy(qphix_amd_ctx *ctx)
{
  delete ctx;
  return QPHIX_AMD_OK;
}

int qphix_amd_set_tbc(qphix_amd_ctx *ctx, const int tbc[4], const double tbc_phases[8])
{
  if (!ctx || !tbc || !tbc_phases) return QPHIX_AMD_EINVAL;
  ctx->tbc_mask = 0;
  for (int m = 0; m < 4; m++) {
    if (tbc[m]) ctx->tbc_mask |= 1 << m;
    ctx->tbc_re[m] = tbc_phases[2 * m];
    ctx->tbc_im[m] = tbc_phases[2 * m + 1];
  }
  return QPHIX_AMD_OK;
}

int qphix_amd_set_traversal(qphix_amd_ctx *ctx, int order)
{
  if (!ctx || order < -1 || order > 3) return QPHIX_AMD_EINVAL;
  ctx->traversal = order;
  return QPHIX_AMD_OK;
}

int qphix_amd_traversal(const qphix_amd_ctx *ctx) { return ctx ? qamd::dslash_traversal(ctx) : -1; }

long qphix_amd_sites_cb(const qphix_amd_ctx *ctx) { return ctx ? (long)ctx->g.vcb : 0; }

size_t qphix_amd_field_bytes(const qphix_amd_ctx *ctx, int field, int prec, int recon)
{
  if (!ctx || !valid_prec(prec)) return 0;
  size_t reals;
  switch (field) {
    case QPHIX_AMD_FIELD_SPINOR: reals = 24; break;
    case QPHIX_AMD_FIELD_GAUGE:
      if (recon != 12 && recon != 18) return 0;
      reals = 8 * (size_t)recon;
      break;
    case QPHIX_AMD_FIELD_CLOVER: reals = 72; break;
    case QPHIX_AMD_FIELD_FULLCLOVER: reals = 144; break;
    default: return 0;
  }
  return reals * (size_t)storage_bytes(prec) * (size_t)ctx->g.vcb;
}

const char *qphix_amd_error_string(int code)
{
  switch (code) {
    case QPHIX_AMD_OK: return "ok";
    case QPHIX_AMD_EINVAL: return "invalid argument";
    case QPHIX_AMD_EGEOM: return "unsupported lattice geometry (extents must be even and >= 2)";
    case QPHIX_AMD_EHIP: return "HIP runtime error";
    case QPHIX_AMD_ENOMEM: return "out of memory";
    case QPHIX_AMD_ENOCONV: return "solver did not converge";
    case QPHIX_AMD_EALIGN: return "field pointer not 16-byte aligned";
    case QPHIX_AMD_ESCRATCH: return "lattice too large for the static scratch of a no-allocation call";
  }
  return "unknown error";
}

int qphix_amd_pack_spinor(qphix_amd_ctx *ctx, void *dst, int prec, const double *src_lex, int cb,
                          void *stream)
{
  if (!ctx || !dst || !src_lex || !valid_prec(prec) || (cb != 0 && cb != 1)) return QPHIX_AMD_EINVAL;
  if (!aligned16(dst)) return QPHIX_AMD_EALIGN;
  return hip_status(qamd::run_pack(ctx, QPHIX_AMD_FIELD_SPINOR, prec, 0, dst, src_lex, cb,
                                   static_cast<hipStream_t>(stream)));
}

int qphix_amd_unpack_spinor(qphix_amd_ctx *ctx, double *dst_lex, const void *src, int prec, int cb,
                            void *stream)
{
  if (!ctx || !dst_lex || !src || !valid_prec(prec) || (cb != 0 && cb != 1)) return QPHIX_AMD_EINVAL;
  if (!aligned16(src)) return QPHIX_AMD_EALIGN;
  return hip_status(
      qamd::run_unpack_spinor(ctx, prec, dst_lex, src, cb, static_cast<hipStream_t>(stream)));
}

int qphix_amd_pack_gauge(qphix_amd_ctx *ctx, void *dst, int prec, int recon, const double *u_lex,
                         int cb, void *stream)
{
  if (!ctx || !dst || !u_lex || !valid_prec(prec) || (cb != 0 && cb != 1)) return QPHIX_AMD_EINVAL;
  if (recon != 12 && recon != 18) return QPHIX_AMD_EINVAL;
  if (!aligned16(dst)) return QPHIX_AMD_EALIGN;
  return hip_status(qamd::run_pack(ctx, QPHIX_AMD_FIELD_GAUGE, prec, recon, dst, u_lex, cb,
                                   static_cast<hipStream_t>(stream)));
}

int qphix_amd_pack_clover(qphix_amd_ctx *ctx, void *dst, int prec, const double *clov_lex, int cb,
                          void *stream)
{
  if (!ctx || !dst || !clov_lex || !valid_prec(prec) || (cb != 0 && cb != 1)) return QPHIX_AMD_EINVAL;
  if (!aligned16(dst)) return QPHIX_AMD_EALIGN;
  return hip_status(qamd::run_pack(ctx, QPHIX_AMD_FIELD_CLOVER, prec, 0, dst, clov_lex, cb,
                                   static_cast<hipStream_t>(stream)));
}

int qphix_amd_pack_full_clover(qphix_amd_ctx *ctx, void *dst, int prec, const double *fclov_lex,
                               int cb, void *stream)
{
  if (!ctx || !dst || !fclov_lex || !valid_prec(prec) || (cb != 0 && cb != 1)) return QPHIX_AMD_EINVAL;
  if (!aligned16(dst)) return QPHIX_AMD_EALIGN;
  return hip_status(qamd::run_pack(ctx, QPHIX_AMD_FIELD_FULLCLOVER, prec, 0, dst, fclov_lex, cb,
                                   static_cast<hipStream_t>(stream)));
}

int qphix_amd_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb, void *out,
                     const void *in, const void *gauge, double coeff_s, double coeff_t_f,
                     double coeff_t_b, void *stream)
{
  DslashCall c = make_call(prec, recon, qamd::OP_WILSON, false, isign, cb, out, in, nullptr, gauge,
                           nullptr, coeff_s, coeff_t_f, coeff_t_b);
  return run(ctx, c, stream);
}

int qphix_amd_dslash_adjoint(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb,
                             void *grad_psi, const void *chi, const void *gauge, double coeff_s,
                             double coeff_t_f, double coeff_t_b, void *stream)
{
  if (cb != 0 && cb != 1) return QPHIX_AMD_EINVAL;
  // (D_s)^dagger = D_{-s} on the other checkerboard, time coefficients swapped
  DslashCall c = make_call(prec, recon, qamd::OP_WILSON, false, -isign, 1 - cb, grad_psi, chi,
                           nullptr, gauge, nullptr, coeff_s, coeff_t_b, coeff_t_f);
  return run(ctx, c, stream);
}

int qphix_amd_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb,
                                void *out, const void *in, const void *chi, const void *gauge,
                                double alpha, double coeff_s, double coeff_t_f, double coeff_t_b,
                                void *stream)
{
  DslashCall c = make_call(prec, recon, qamd::OP_WILSON, true, isign, cb, out, in, chi, gauge,
                           nullptr, coeff_s, coeff_t_f, coeff_t_b);
  c.alpha = alpha;
  return run(ctx, c, stream);
}

int qphix_amd_clov_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb, void *out,
                          const void *in, const void *gauge, const void *clover, double coeff_s,
                          double coeff_t_f, double coeff_t_b, void *stream)
{
  DslashCall c = make_call(prec, recon, qamd::OP_CLOVER, false, isign, cb, out, in, nullptr, gauge,
                           clover, coeff_s, coeff_t_f, coeff_t_b);
  return run(ctx, c, stream);
}

int qphix_amd_clov_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb,
                                     void *out, const void *in, const void *chi,
                                     const void *gauge, const void *clover, double coeff_s,
                                     double coeff_t_f, double coeff_t_b, double prec_mass_rho,
                                     void *stream)
{
  DslashCall c = make_call(prec, recon, qamd::OP_CLOVER, true, isign, cb, out, in, chi, gauge,
                           clover, coeff_s, coeff_t_f, coeff_t_b);
  c.rho = prec_mass_rho;
  return run(ctx, c, stream);
}

int qphix_amd_tm_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb, void *out,
                        const void *in, const void *gauge, double coeff_s, double coeff_t_f,
                        double coeff_t_b, double mu, double muinv, void *stream)
{
  DslashCall c = make_call(prec, recon, qamd::OP_TM, false, isign, cb, out, in, nullptr, gauge,
                           nullptr, coeff_s, coeff_t_f, coeff_t_b);
  c.mu = mu;
  c.muinv = muinv;
  return run(ctx, c, stream);
}

int qphix_amd_tm_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb,
                                   void *out, const void *in, const void *chi, const void *gauge,
                                   double alpha, double coeff_s, double coeff_t_f,
                                   double coeff_t_b, double mu, void *stream)
{
  DslashCall c = make_call(prec, recon, qamd::OP_TM, true, isign, cb, out, in, chi, gauge, nullptr,
                           coeff_s, coeff_t_f, coeff_t_b);
  c.alpha = alpha;
  c.mu = mu;
  return run(ctx, c, stream);
}

int qphix_amd_tm_clov_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb,
                             void *out, const void *in, const void *gauge, const void *fclover,
                             double coeff_s, double coeff_t_f, double coeff_t_b, void *stream)
{
  DslashCall c = make_call(prec, recon, qamd::OP_TMCLOVER, false, isign, cb, out, in, nullptr,
                           gauge, fclover, coeff_s, coeff_t_f, coeff_t_b);
  return run(ctx, c, stream);
}

int qphix_amd_tm_clov_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon, int isign,
                                        int cb, void *out, const void *in, const void *chi,
                                        const void *gauge, const void *fclover, double coeff_s,
                                        double coeff_t_f, double coeff_t_b, double prec_mass_rho,
                                        void *stream)
{
  DslashCall c = make_call(prec, recon, qamd::OP_TMCLOVER, true, isign, cb, out, in, chi, gauge,
                           fclover, coeff_s, coeff_t_f, coeff_t_b);
  c.rho = prec_mass_rho;
  return run(ctx, c, stream);
}

// Non-degenerate twisted mass: the reference declares the ndtm_ family but
// does not generate it (codegen/twisted_mass_enum.h:19-41,
// codegen/jinja/kernels-inactive.js:1-4), and its bodies are the degenerate
// ones (codegen/dslash_common.cc:1599-1603,1643-1650), so these entry points
// are the twisted-mass kernels under the ndtm_ names.
int qphix_amd_ndtm_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb, void *out,
                          const void *in, const void *gauge, double coeff_s, double coeff_t_f,
                          double coeff_t_b, double mu, double muinv, void *stream)
{
  return qphix_amd_tm_dslash(ctx, prec, recon, isign, cb, out, in, gauge, coeff_s, coeff_t_f, coeff_t_b,
                             mu, muinv, stream);
}

int qphix_amd_ndtm_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb,
                                     void *out, const void *in, const void *chi, const void *gauge,
                                     double alpha, double coeff_s, double coeff_t_f,
                                     double coeff_t_b, double mu, void *stream)
{
  return qphix_amd_tm_dslash_achimbdpsi(ctx, prec, recon, isign, cb, out, in, chi, gauge, alpha, coeff_s,
                                        coeff_t_f, coeff_t_b, mu, stream);
}

int qphix_amd_ndtm_clov_dslash(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb,
                               void *out, const void *in, const void *gauge, const void *fclover,
                               double coeff_s, double coeff_t_f, double coeff_t_b, void *stream)
{
  return qphix_amd_tm_clov_dslash(ctx, prec, recon, isign, cb, out, in, gauge, fclover, coeff_s,
                                  coeff_t_f, coeff_t_b, stream);
}

int qphix_amd_ndtm_clov_dslash_achimbdpsi(qphix_amd_ctx *ctx, int prec, int recon, int isign,
                                          int cb, void *out, const void *in, const void *chi,
                                          const void *gauge, const void *fclover, double coeff_s,
                                          double coeff_t_f, double coeff_t_b,
                                          double prec_mass_rho, void *stream)
{
  return qphix_amd_tm_clov_dslash_achimbdpsi(ctx, prec, recon, isign, cb, out, in, chi, gauge, fclover,
                                             coeff_s, coeff_t_f, coeff_t_b, prec_mass_rho, stream);
}

// ---- operators on QPhiX FourSpinorBlock arrays, in place -------------------
int qphix_amd_qphix_dslash(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon, int isign,
                           int cb, void *out, const void *in, const void *gauge,
                           const unsigned char *accumulate, double coeff_s, double coeff_t_f,
                           double coeff_t_b, void *stream)
{
  DslashCall c = make_call(0, recon, qamd::OP_WILSON, false, isign, cb, out, in, nullptr, gauge, nullptr,
                           coeff_s, coeff_t_f, coeff_t_b);
  return run_qphix(ctx, lay, accumulate, c, stream);
}

int qphix_amd_qphix_dslash_achimbdpsi(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon,
                                      int isign, int cb, void *out, const void *in, const void *chi,
                                      const void *gauge, const unsigned char *accumulate, double alpha,
                                      double coeff_s, double coeff_t_f, double coeff_t_b, void *stream)
{
  DslashCall c = make_call(0, recon, qamd::OP_WILSON, true, isign, cb, out, in, chi, gauge, nullptr, coeff_s,
                           coeff_t_f, coeff_t_b);
  c.alpha = alpha;
  return run_qphix(ctx, lay, accumulate, c, stream);
}

int qphix_amd_qphix_clov_dslash(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon, int isign,
                                int cb, void *out, const void *in, const void *gauge, const void *clover,
                                const unsigned char *accumulate, double coeff_s, double coeff_t_f,
                                double coeff_t_b, void *stream)
{
  DslashCall c = make_call(0, recon, qamd::OP_CLOVER, false, isign, cb, out, in, nullptr, gauge, clover,
                           coeff_s, coeff_t_f, coeff_t_b);
  return run_qphix(ctx, lay, accumulate, c, stream);
}

int qphix_amd_qphix_clov_dslash_achimbdpsi(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon,
                                           int isign, int cb, void *out, const void *in, const void *chi,
                                           const void *gauge, const void *clover,
                                           const unsigned char *accumulate, double coeff_s, double coeff_t_f,
                                           double coeff_t_b, double prec_mass_rho, void *stream)
{
  DslashCall c = make_call(0, recon, qamd::OP_CLOVER, true, isign, cb, out, in, chi, gauge, clover, coeff_s,
                           coeff_t_f, coeff_t_b);
  c.rho = prec_mass_rho;
  return run_qphix(ctx, lay, accumulate, c, stream);
}

int qphix_amd_qphix_tm_dslash(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon, int isign,
                              int cb, void *out, const void *in, const void *gauge,
                              const unsigned char *accumulate, double coeff_s, double coeff_t_f,
                              double coeff_t_b, double mu, double muinv, void *stream)
{
  DslashCall c = make_call(0, recon, qamd::OP_TM, false, isign, cb, out, in, nullptr, gauge, nullptr, coeff_s,
                           coeff_t_f, coeff_t_b);
  c.mu = mu;
  c.muinv = muinv;
  return run_qphix(ctx, lay, accumulate, c, stream);
}

int qphix_amd_qphix_tm_dslash_achimbdpsi(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon,
                                         int isign, int cb, void *out, const void *in, const void *chi,
                                         const void *gauge, const unsigned char *accumulate, double alpha,
                                         double coeff_s, double coeff_t_f, double coeff_t_b, double mu,
                                         void *stream)
{
  DslashCall c = make_call(0, recon, qamd::OP_TM, true, isign, cb, out, in, chi, gauge, nullptr, coeff_s,
                           coeff_t_f, coeff_t_b);
  c.alpha = alpha;
  c.mu = mu;
  return run_qphix(ctx, lay, accumulate, c, stream);
}

int qphix_amd_qphix_tm_clov_dslash(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon,
                                   int isign, int cb, void *out, const void *in, const void *gauge,
                                   const void *fclover, const unsigned char *accumulate, double coeff_s,
                                   double coeff_t_f, double coeff_t_b, void *stream)
{
  DslashCall c = make_call(0, recon, qamd::OP_TMCLOVER, false, isign, cb, out, in, nullptr, gauge, fclover,
                           coeff_s, coeff_t_f, coeff_t_b);
  return run_qphix(ctx, lay, accumulate, c, stream);
}

int qphix_amd_qphix_tm_clov_dslash_achimbdpsi(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay,
                                              int recon, int isign, int cb, void *out, const void *in,
                                              const void *chi, const void *gauge, const void *fclover,
                                              const unsigned char *accumulate, double coeff_s,
                                              double coeff_t_f, double coeff_t_b, double prec_mass_rho,
                                              void *stream)
{
  DslashCall c = make_call(0, recon, qamd::OP_TMCLOVER, true, isign, cb, out, in, chi, gauge, fclover,
                           coeff_s, coeff_t_f, coeff_t_b);
  c.rho = prec_mass_rho;
  return run_qphix(ctx, lay, accumulate, c, stream);
}

}  // extern "C"
