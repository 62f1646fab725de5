// solver.hip -- the even-odd preconditioned (twisted-mass) clover operator
// and a CG on its normal equations, plain FP64 or mixed precision with
// reliable updates.
//
// Operator on the odd checkerboard (cb 1), F = A + i mu g5 per site:
//   M       = F_oo - 1/4 D_oe F_ee^-1 D_eo
//   M^dag   = F_oo^dag - 1/4 D_oe^- (F_ee^dag)^-1 D_eo^-
// i.e. one (tm_)clov_dslash on cb 0 followed by one
// (tm_)clov_dslash_achimbdpsi with beta = 1/4 on cb 1; the dagger uses
// isign = -1 and the blocks of -mu (D_+^dag = D_-).  With mu = 0 the
// Hermitian packed clover kernels are used.  This is the Schur complement
// that QPhiX's even-odd clover operator builds from the same two generated
// kernels (codegen/jinja/clov_dslash_decl.h.j2).
//
// Mixed-precision CG: the iterated residual r and search direction p (and
// A p, where the operator's last kernel does not update r itself) live in
// the sloppy storage type (FP16/BF16/FP32 storage, FP32 arithmetic),
// in units of a scale sigma = |r_true| / sqrt(n) fixed at each reliable
// update so that FP16's range holds them.  The iterate increment is
// accumulated in FP32 and folded into the FP64 solution at each reliable
// update, where the true residual b - M^dag M x is recomputed in FP64 and
// replaces r (beta = |r_true|^2 / |r_old|^2, p rescaled to the new units).
// All reductions are FP64 with a fixed summation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <new>
#include <vector>

#include "blas.h"
#include "dslash_kernel.h"

using namespace qamd;

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  hipError_t alloc(size_t bytes)
  {
    n = bytes;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      p = nullptr;
      n = 0;
    }
    return e;
  }
  // at least `bytes`, keeping the current allocation when it is large enough
  hipError_t reserve(size_t bytes)
  {
    if (p && n >= bytes) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    return alloc(bytes);
  }
  ~DevBuf()
  {
    if (p) (void)hipFree(p);
  }
};

int sbytes(int prec) { return prec == QPHIX_AMD_F64 ? 8 : prec == QPHIX_AMD_F32 ? 4 : 2; }

}  // namespace

struct qphix_amd_eo_op {
  qphix_amd_ctx *ctx;
  int tm;  // 1: twisted-mass clover (full blocks)
  double mu;
  int prec_s, recon_s;
  // per precision level (0 = FP64, 1 = sloppy): links on cb 0 and cb 1;
  // epilogue fields: dslash blocks on cb 0 and achimbdpsi blocks on cb 1,
  // index [dagger]
  DevBuf gauge[2][2];
  DevBuf dsl[2][2], achi[2][2];
  DevBuf tmp[2];  // cb-0 intermediate per level
  // Workspaces of the force and CG calls, kept from one call to the next
  // (grow-only) and released by qphix_amd_eo_op_destroy, so that repeated
  // solves and forces allocate nothing: three FP64 vectors (the forces' Y,
  // Z, W and the CG's FP64 residual, M^dag M x and cb-1 intermediate), the
  // clover force's lexicographic fields and K, the CG's sloppy vectors,
  // scalars and partial sums, and a pinned block for the scalars.  Calls on
  // one operator do not overlap, so they share these.
  DevBuf w64[3], wk, wrs, wp, wap, wxs, wscb, wpart, wpnorm;
  qamd::CGScalars *hsc = nullptr;
  ~qphix_amd_eo_op()
  {
    if (hsc) (void)hipHostFree(hsc);
  }
};

namespace {

int level_prec(const qphix_amd_eo_op *op, int lvl) { return lvl == 0 ? QPHIX_AMD_F64 : op->prec_s; }
int level_recon(const qphix_amd_eo_op *op, int lvl) { return lvl == 0 ? 12 : op->recon_s; }

// norm: when set, per work-group partials of |out|^2 (vcb / 128 of them);
// step: when set too (dagger only), out receives r - (*step) M^dag in, with
// r read from rin (null: out itself holds r), and norm gets the partials of
// that |r|^2
// first stage of M: out (cb 0) = F_ee^-1 D_eo in (dagger: the dagger blocks
// and isign -1)
hipError_t eo_stage1(qphix_amd_eo_op *op, int lvl, int dagger, void *out, const void *in,
                     hipStream_t st, const int *stop = nullptr)
{
  DslashCall c{};
  c.stop = stop;
  c.prec = level_prec(op, lvl);
  c.recon = level_recon(op, lvl);
  c.op = op->tm ? OP_TMCLOVER : OP_CLOVER;
  c.isign = dagger ? -1 : 1;
  c.coeff_s = c.coeff_tf = c.coeff_tb = 1.0;
  c.cb = 0;
  c.out = out;
  c.in = in;
  c.gauge = op->gauge[lvl][0].p;
  c.clov = op->dsl[lvl][dagger].p;
  return run_dslash_any(op->ctx, c, st);
}

// second stage: out (cb 1) = F_oo in - 1/4 D_oe t, t the first stage's result
hipError_t eo_stage2(qphix_amd_eo_op *op, int lvl, int dagger, void *out, const void *t, const void *in,
                     hipStream_t st, double *norm = nullptr, const double *step = nullptr,
                     const int *stop = nullptr, const void *rin = nullptr)
{
  DslashCall c{};
  c.stop = stop;
  c.prec = level_prec(op, lvl);
  c.recon = level_recon(op, lvl);
  c.op = op->tm ? OP_TMCLOVER : OP_CLOVER;
  c.isign = dagger ? -1 : 1;
  c.achi = true;
  c.cb = 1;
  c.out = out;
  c.in = t;
  c.chi = in;
  c.gauge = op->gauge[lvl][1].p;
  c.clov = op->achi[lvl][dagger].p;
  c.coeff_s = c.coeff_tf = c.coeff_tb = 0.25;
  c.rho = 0.0;
  c.norm = norm;
  c.step = step;
  c.rin = rin;
  return run_dslash_any(op->ctx, c, st);
}

// out = M in (dagger: M^dag), through the operator's cb-0 temporary
hipError_t eo_apply(qphix_amd_eo_op *op, int lvl, int dagger, void *out, const void *in,
                    hipStream_t st, double *norm = nullptr, const double *step = nullptr,
                    const int *stop = nullptr, const void *rin = nullptr)
{
  hipError_t e = eo_stage1(op, lvl, dagger, op->tmp[lvl].p, in, st, stop);
  if (e != hipSuccess) return e;
  return eo_stage2(op, lvl, dagger, out, op->tmp[lvl].p, in, st, norm, step, stop, rin);
}

// out = M^dag M in, via the cb-1 temporary mid = M in
hipError_t eo_normal(qphix_amd_eo_op *op, int lvl, void *out, void *mid, const void *in,
                     hipStream_t st, const int *stop = nullptr)
{
  hipError_t e = eo_apply(op, lvl, 0, mid, in, st, nullptr, nullptr, stop);
  if (e != hipSuccess) return e;
  return eo_apply(op, lvl, 1, out, mid, st, nullptr, nullptr, stop);
}

}  // namespace

extern "C" {

int qphix_amd_clover_from_links(qphix_amd_ctx *ctx, double *clov_lex, const double *u_lex,
                                double alpha, double csw, void *stream)
{
  if (!ctx || !clov_lex || !u_lex) return QPHIX_AMD_EINVAL;
  return run_clover_from_links(ctx, clov_lex, u_lex, alpha, csw, (hipStream_t)stream) == hipSuccess
             ? QPHIX_AMD_OK
             : QPHIX_AMD_EHIP;
}

int qphix_amd_clover_invert(qphix_amd_ctx *ctx, double *inv_lex, const double *clov_lex, void *stream)
{
  if (!ctx || !inv_lex || !clov_lex || inv_lex == clov_lex) return QPHIX_AMD_EINVAL;
  return run_clover_blocks(ctx, inv_lex, clov_lex, 0, 0.0, 1, (hipStream_t)stream) == hipSuccess
             ? QPHIX_AMD_OK
             : QPHIX_AMD_EHIP;
}

int qphix_amd_tmclover_blocks(qphix_amd_ctx *ctx, double *out_lex, const double *clov_lex, double mu,
                              int sign, int invert, void *stream)
{
  if (!ctx || !out_lex || !clov_lex || (sign != 1 && sign != -1)) return QPHIX_AMD_EINVAL;
  return run_clover_blocks(ctx, out_lex, clov_lex, invert ? 2 : 1, mu, sign, (hipStream_t)stream) ==
                 hipSuccess
             ? QPHIX_AMD_OK
             : QPHIX_AMD_EHIP;
}

int qphix_amd_eo_op_create(qphix_amd_eo_op **out_op, qphix_amd_ctx *ctx, const double *u_lex,
                           const double *clov_lex, double mu, int prec_sloppy, int recon_sloppy,
                           void *stream)
{
  if (!out_op || !ctx || !u_lex || !clov_lex) return QPHIX_AMD_EINVAL;
  if (ctx->comm_mask) return QPHIX_AMD_EINVAL;  // the operator needs the whole lattice
  if (prec_sloppy < QPHIX_AMD_F64 || prec_sloppy > QPHIX_AMD_BF16) return QPHIX_AMD_EINVAL;
  if (recon_sloppy != 12 && recon_sloppy != 18) return QPHIX_AMD_EINVAL;
  *out_op = nullptr;
  hipStream_t st = (hipStream_t)stream;
  qphix_amd_eo_op *op = new (std::nothrow) qphix_amd_eo_op();
  if (!op) return QPHIX_AMD_ENOMEM;
  op->ctx = ctx;
  op->tm = mu != 0.0;
  op->mu = mu;
  op->prec_s = prec_sloppy;
  op->recon_s = recon_sloppy;
  const long V = (long)ctx->g.vcb * 2;
  const int nreal_cl = op->tm ? 144 : 72;
  const int field = op->tm ? QPHIX_AMD_FIELD_FULLCLOVER : QPHIX_AMD_FIELD_CLOVER;
  int rc = QPHIX_AMD_OK;
  hipError_t e = hipSuccess;
  // the packed fields of both precision levels; links packed right away
  for (int lvl = 0; lvl < 2 && rc == QPHIX_AMD_OK && e == hipSuccess; lvl++) {
    const int prec = level_prec(op, lvl), recon = level_recon(op, lvl);
    const size_t gb = qphix_amd_field_bytes(ctx, QPHIX_AMD_FIELD_GAUGE, prec, recon);
    const size_t cb = qphix_amd_field_bytes(ctx, field, prec, 0);
    const size_t sb = qphix_amd_field_bytes(ctx, QPHIX_AMD_FIELD_SPINOR, prec, 0);
    for (int c = 0; c < 2; c++)
      if (op->gauge[lvl][c].alloc(gb) != hipSuccess) rc = QPHIX_AMD_ENOMEM;
    for (int d = 0; d < 2; d++)
      if (op->dsl[lvl][d].alloc(cb) != hipSuccess || op->achi[lvl][d].alloc(cb) != hipSuccess)
        rc = QPHIX_AMD_ENOMEM;
    if (op->tmp[lvl].alloc(sb) != hipSuccess) rc = QPHIX_AMD_ENOMEM;
    for (int c = 0; c < 2 && rc == QPHIX_AMD_OK && e == hipSuccess; c++)
      e = run_pack(ctx, QPHIX_AMD_FIELD_GAUGE, prec, recon, op->gauge[lvl][c].p, u_lex, c, st);
  }
  // FP64 lexicographic epilogue blocks, one dagger at a time: the cb-0
  // dslash takes the inverse blocks, the cb-1 achimbdpsi the forward ones
  // (with mu = 0 the clover field itself, for both daggers)
  DevBuf lexd, lexa;
  if (rc == QPHIX_AMD_OK &&
      (lexd.alloc(V * nreal_cl * 8) != hipSuccess || (op->tm && lexa.alloc(V * nreal_cl * 8) != hipSuccess)))
    rc = QPHIX_AMD_ENOMEM;
  for (int d = 0; d < 2 && rc == QPHIX_AMD_OK && e == hipSuccess; d++) {
    const int sign = d ? -1 : 1;
    const double *fwd = clov_lex;
    if (op->tm) {
      e = run_clover_blocks(ctx, (double *)lexd.p, clov_lex, 2, mu, sign, st);
      if (e == hipSuccess) e = run_clover_blocks(ctx, (double *)lexa.p, clov_lex, 1, mu, sign, st);
      fwd = (const double *)lexa.p;
    } else if (d == 0) {
      e = run_clover_blocks(ctx, (double *)lexd.p, clov_lex, 0, 0.0, 1, st);
    }
    for (int lvl = 0; lvl < 2 && e == hipSuccess; lvl++) {
      const int prec = level_prec(op, lvl);
      e = run_pack(ctx, field, prec, 0, op->dsl[lvl][d].p, (const double *)lexd.p, 0, st);
      if (e == hipSuccess) e = run_pack(ctx, field, prec, 0, op->achi[lvl][d].p, fwd, 1, st);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // lex temporaries are freed next
  if (rc == QPHIX_AMD_OK && e != hipSuccess) rc = QPHIX_AMD_EHIP;
  if (rc != QPHIX_AMD_OK) {
    delete op;
    return rc;
  }
  *out_op = op;
  return QPHIX_AMD_OK;
}

// Frees the operator's device memory.  hipFree waits for the whole device
// (HIP's documented behaviour) before releasing a block, so work still
// queued on any stream that uses the operator completes first; destroy is
// therefore a set-up call (once per gauge configuration), not one for a hot
// loop.
int qphix_amd_eo_op_destroy(qphix_amd_eo_op *op)
{
  delete op;
  return QPHIX_AMD_OK;
}

int qphix_amd_eo_op_apply(qphix_amd_eo_op *op, int prec, int dagger, void *out, const void *in,
                          void *stream)
{
  if (!op || !out || !in || out == in) return QPHIX_AMD_EINVAL;
  if (op->ctx->comm_mask) return QPHIX_AMD_EINVAL;  // split after the operator was built
  int lvl;
  if (prec == QPHIX_AMD_F64)
    lvl = 0;
  else if (prec == op->prec_s)
    lvl = 1;
  else
    return QPHIX_AMD_EINVAL;
  return eo_apply(op, lvl, dagger ? 1 : 0, out, in, (hipStream_t)stream) == hipSuccess ? QPHIX_AMD_OK
                                                                                       : QPHIX_AMD_EHIP;
}

namespace {
// the two terms of dS = 1/2 Re[<Y, dD_oe Z> + <W, dD_eo X>] (FP64 native
// spinors; Y, W on checkerboard 1, 0)
void hopping_terms(HopTerm t[2], const void *y, const void *z, const void *w, const void *x, double scale)
{
  t[0] = HopTerm{y, z, 1, 1, 0.5 * scale};
  t[1] = HopTerm{w, x, 0, 1, 0.5 * scale};
}
}  // namespace

int qphix_amd_eo_op_force(qphix_amd_eo_op *op, double *force_lex, const void *x,
                          const double *u_lex, double scale, int project, void *stream)
{
  if (!op || !force_lex || !x || !u_lex) return QPHIX_AMD_EINVAL;
  if (op->ctx->comm_mask) return QPHIX_AMD_EINVAL;  // needs the whole lattice
  hipStream_t st = (hipStream_t)stream;
  const long n = 24L * op->ctx->g.vcb;
  DevBuf &y = op->w64[0], &z = op->w64[1], &w = op->w64[2];
  if (y.reserve(n * 8) != hipSuccess || z.reserve(n * 8) != hipSuccess || w.reserve(n * 8) != hipSuccess)
    return QPHIX_AMD_ENOMEM;
  // Z = F_ee^-1 D_eo X, Y = M X (from Z), W = (F_ee^-1)^dag D_oe^dag Y
  hipError_t e = eo_stage1(op, 0, 0, z.p, x, st);
  if (e == hipSuccess) e = eo_stage2(op, 0, 0, y.p, z.p, x, st);
  if (e == hipSuccess) e = eo_stage1(op, 0, 1, w.p, y.p, st);
  if (e != hipSuccess) return QPHIX_AMD_EHIP;
  // dS = 1/2 Re[<Y, dD_oe Z> + <W, dD_eo X>], both terms in one kernel
  HopTerm terms[2];
  hopping_terms(terms, y.p, z.p, w.p, x, scale);
  e = run_hop_force(op->ctx, force_lex, u_lex, terms, 2, 1.0, project, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP;
}

int qphix_amd_clover_force(qphix_amd_ctx *ctx, double *force_lex, const double *chi_lex,
                           const double *psi_lex, const double *u_lex, double csw, double scale_even,
                           double scale_odd, int project, void *stream)
{
  if (!ctx || !force_lex || !chi_lex || !psi_lex || !u_lex) return QPHIX_AMD_EINVAL;
  if (ctx->comm_mask) return QPHIX_AMD_EINVAL;  // needs the whole lattice
  hipStream_t st = (hipStream_t)stream;
  const long V = 2L * ctx->g.vcb;
  DevBuf K;
  if (K.alloc(V * kKSiteReals * 8) != hipSuccess) return QPHIX_AMD_ENOMEM;
  hipError_t e = run_clover_force(ctx, force_lex, (double *)K.p, chi_lex, psi_lex, u_lex, csw, scale_even,
                                  scale_odd, 1.0, project, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);  // K is released on return
  return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP;
}

int qphix_amd_eo_op_force_clover(qphix_amd_eo_op *op, double *force_lex, const void *x,
                                 const double *u_lex, double csw, double scale, int project,
                                 void *stream)
{
  if (!op || !force_lex || !x || !u_lex) return QPHIX_AMD_EINVAL;
  if (op->ctx->comm_mask) return QPHIX_AMD_EINVAL;  // needs the whole lattice
  hipStream_t st = (hipStream_t)stream;
  const long n = 24L * op->ctx->g.vcb;
  DevBuf &y = op->w64[0], &z = op->w64[1], &w = op->w64[2];
  DevBuf &K = op->wk;
  const long V = 2L * op->ctx->g.vcb;
  if (y.reserve(n * 8) != hipSuccess || z.reserve(n * 8) != hipSuccess || w.reserve(n * 8) != hipSuccess ||
      K.reserve(V * kKSiteReals * 8) != hipSuccess)
    return QPHIX_AMD_ENOMEM;
  // Z = F_ee^-1 D_eo X, Y = M X (from Z), W = (F_ee^-1)^dag D_oe^dag Y
  hipError_t e = eo_stage1(op, 0, 0, z.p, x, st);
  if (e == hipSuccess) e = eo_stage2(op, 0, 0, y.p, z.p, x, st);
  if (e == hipSuccess) e = eo_stage1(op, 0, 1, w.p, y.p, st);
  if (e != hipSuccess) return QPHIX_AMD_EHIP;
  // hopping term: dS = 1/2 Re[<Y, dD_oe Z> + <W, dD_eo X>]; clover term:
  // dS = -2 Re <Y, dA_oo X> - 1/2 Re <W, dA_ee Z> (dF = dA: the twisted mass
  // does not depend on the links).  One gather forms both terms' G per link
  // and adds the projected force to force_lex; chi = (W, Y) and psi = (Z, X)
  // on checkerboards (0, 1).
  const double *chi_cb[2] = {(const double *)w.p, (const double *)y.p};
  const double *psi_cb[2] = {(const double *)z.p, (const double *)x};
  HopTerm terms[2];
  hopping_terms(terms, y.p, z.p, w.p, x, scale);
  e = run_clover_hop_force(op->ctx, force_lex, (double *)K.p, chi_cb, psi_cb, u_lex, csw, -0.5 * scale,
                           -2.0 * scale, 1.0, project, terms, 2, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP;
}

// CG iterations enqueued between two host reads of the scalars.  The
// reliable-update trigger is evaluated on the device after every iteration,
// and the iterations queued after it fire return at once, so the batch
// length costs only those empty launches, never a late update.
#ifndef QAMD_CG_CHECK
#define QAMD_CG_CHECK 8
#endif
// <p, M^dag M p> = |M p|^2 summed by the kernel that stores M p, instead of
// a dot product of p and M^dag M p (which reads both again); and the
// residual update r -= alpha M^dag M p with its |r|^2 done by the kernel that
// computes M^dag M p, which then is never stored (x += alpha p moves into the
// p update, which reads p anyway)
#ifndef QAMD_CG_FUSED_NORM
#define QAMD_CG_FUSED_NORM 1
#endif

int qphix_amd_cg_solve(qphix_amd_eo_op *op, void *x, const void *b, const qphix_amd_cg_params *prm,
                       qphix_amd_cg_result *res, void *stream)
{
  if (!op || !x || !b || !prm || !res || x == b) return QPHIX_AMD_EINVAL;
  if (!(prm->rsd_target > 0.0) || prm->max_iters <= 0) return QPHIX_AMD_EINVAL;
  if (op->ctx->comm_mask) return QPHIX_AMD_EINVAL;  // split after the operator was built
  hipStream_t st = (hipStream_t)stream;
  auto t0 = std::chrono::steady_clock::now();
  const long vcb = op->ctx->g.vcb;
  const long n = 24 * vcb;
  const bool mixed = prm->mixed != 0;
  const int ps = mixed ? op->prec_s : QPHIX_AMD_F64;
  const int px = ps == QPHIX_AMD_F64 ? QPHIX_AMD_F64 : QPHIX_AMD_F32;  // increment accumulator
  const int lvl = mixed ? 1 : 0;
  // Default reliable-update threshold by sloppy storage: the FP16 (10-bit
  // mantissa) and FP32 iterations stay accurate across a larger residual drop
  // than BF16 (7-bit); values from a sweep of the time to solution at
  // 32^3x64, kappa 0.135, csw 1 (FP16: 0.02 -> 123 iterations, 6 updates;
  // BF16: 0.15 -> 131 iterations, 12 updates).
  const double delta = prm->delta > 0.0 ? prm->delta
                       : ps == QPHIX_AMD_BF16 ? 0.15
                       : ps == QPHIX_AMD_F16  ? 0.02
                                              : 0.01;

  // fused |M p|^2 needs every work-group of the operator kernels full
  const bool fuse = QAMD_CG_FUSED_NORM && vcb % kDslashBlock == 0;
  const long nwg = vcb / kDslashBlock;  // (vcb < 2^31 / 72: nwg / 2048 < kMaxPartials)
  DevBuf &r64 = op->w64[0], &ax64 = op->w64[1], &mid = op->w64[2];
  DevBuf &rs = op->wrs, &p = op->wp, &ap = op->wap, &xs = op->wxs;
  DevBuf &scb = op->wscb, &part = op->wpart, &pnorm = op->wpnorm;
  if (fuse && pnorm.reserve(nwg * 8) != hipSuccess) return QPHIX_AMD_ENOMEM;
  if (r64.reserve(n * 8) != hipSuccess || (!fuse && ax64.reserve(n * 8) != hipSuccess) ||
      rs.reserve(n * sbytes(ps)) != hipSuccess || p.reserve(n * sbytes(ps)) != hipSuccess ||
      (!fuse && ap.reserve(n * sbytes(ps)) != hipSuccess) || xs.reserve(n * sbytes(px)) != hipSuccess ||
      mid.reserve(n * (size_t)sbytes(QPHIX_AMD_F64)) != hipSuccess ||
      scb.reserve(sizeof(CGScalars)) != hipSuccess || part.reserve(kMaxPartials * 8) != hipSuccess)
    return QPHIX_AMD_ENOMEM;
  CGScalars *sc = (CGScalars *)scb.p;
  double *partial = (double *)part.p;
  if (!op->hsc && hipHostMalloc((void **)&op->hsc, sizeof(CGScalars), hipHostMallocDefault) != hipSuccess) {
    op->hsc = nullptr;
    return QPHIX_AMD_ENOMEM;
  }
  CGScalars *hsc = op->hsc;

  auto fetch = [&]() -> hipError_t {
    hipError_t e = hipMemcpyAsync(hsc, sc, sizeof(CGScalars), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return e;
  };
  double target2 = 0.0;  // |r|^2 to reach (true units), set from |b|^2 below
  auto set_scalars = [&](double rr, double beta, double sig, double rmx) -> hipError_t {
    CGScalars h{};
    h.rr = rr;
    h.beta = beta;
    h.sigma = sig;
    h.rmax = rmx;
    h.delta = delta;
    h.mixed = mixed ? 1 : 0;
    h.target2 = target2;
    h.one = 1.0;
    blas_set_scalars(sc, h, st);
    return hipGetLastError();
  };

  // r64 = b - M^dag M x64, returns |r64|^2 in *r2
  auto true_residual = [&](double *r2) -> hipError_t {
    hipError_t e;
    if (fuse) {
      // the operator's last kernel stores b - 1 * M^dag (M x64) into r64 and
      // sums its |r64|^2 per work-group (M^dag M x64 is never stored)
      e = eo_apply(op, 0, 0, mid.p, x, st);
      if (e == hipSuccess)
        e = eo_apply(op, 0, 1, r64.p, mid.p, st, (double *)pnorm.p, &sc->one, nullptr, b);
      if (e != hipSuccess) return e;
      blas_finalize((const double *)pnorm.p, nwg, sc, 2, partial, st);
    } else {
      e = eo_normal(op, 0, ax64.p, mid.p, x, st);
      if (e != hipSuccess) return e;
      e = hipMemcpyAsync(r64.p, b, n * 8, hipMemcpyDeviceToDevice, st);
      if (e != hipSuccess) return e;
      blas_axpby(QPHIX_AMD_F64, ax64.p, QPHIX_AMD_F64, r64.p, -1.0, 1.0, n, partial, sc, st);
    }
    e = fetch();
    *r2 = hsc->tmp;
    return e;
  };

  long mv = 0;
  int iters = 0, updates = 0;
  hipError_t e = hipMemsetAsync(x, 0, n * 8, st);
  if (e == hipSuccess) e = hipMemsetAsync(xs.p, 0, n * sbytes(px), st);
  // |b|^2
  blas_axpby(QPHIX_AMD_F64, b, QPHIX_AMD_F64, r64.p, 1.0, 0.0, n, partial, sc, st);
  if (e == hipSuccess) e = fetch();
  if (e != hipSuccess) return QPHIX_AMD_EHIP;
  const double b2 = hsc->tmp;
  target2 = prm->rsd_target * prm->rsd_target * b2;
  double r2 = b2;
  // sloppy units: r_s = r / sigma
  double sigma = mixed ? std::sqrt(r2 / (double)n) : 1.0;
  if (!(sigma > 0.0)) sigma = 1.0;
  blas_axpby(QPHIX_AMD_F64, r64.p, ps, rs.p, 1.0 / sigma, 0.0, n, partial, nullptr, st);
  blas_axpby(ps, rs.p, ps, p.p, 1.0, 0.0, n, partial, nullptr, st);
  double rmax = std::sqrt(r2);  // max true-unit residual since the last update
  e = set_scalars(r2 / (sigma * sigma), 0.0, sigma, rmax);
  bool converged = b2 == 0.0;
  int rc = QPHIX_AMD_OK;
  int base = 0;         // iterations before the scalars were last set
  bool restart = true;  // the next iteration starts from a fresh p
  int last_len = 0;     // iterations between the last two reliable updates
  const int *stop = &sc->stop;
  while (e == hipSuccess && !converged && iters < prm->max_iters) {
    // A batch of iterations is enqueued back to back (alpha, beta, the norms
    // and the reliable-update trigger never leave the device); an iteration
    // begins with the previous one's x += alpha p, p = r + beta p.  Once the
    // trigger fires, the rest of the batch does nothing; the host reads the
    // scalars once per batch.
    // batches end near the iteration where the trigger last fired (counted
    // from the update before), so that few launches are queued past it
    int batch = QAMD_CG_CHECK;
    if (last_len > 0) batch = std::max(2, std::min(batch, last_len - (iters - base) + 1));
    batch = std::min(batch, prm->max_iters - iters);
    for (int k = 0; k < batch && e == hipSuccess; k++) {
      const bool step_p = !(restart && k == 0);
      if (fuse) {
        // mid = M p with the partials of |M p|^2 -> alpha; r -= alpha M^dag
        // mid with those of |r|^2 -> beta, and the trigger
        if (step_p) blas_cg_xpay_x(ps, px, p.p, xs.p, rs.p, n, sc, true, st, stop);
        double *pn = (double *)pnorm.p;
        e = eo_apply(op, lvl, 0, mid.p, p.p, st, pn, nullptr, stop);
        if (e != hipSuccess) break;
        blas_finalize(pn, nwg, sc, 0, partial, st, stop);
        e = eo_apply(op, lvl, 1, rs.p, mid.p, st, pn, &sc->alpha, stop);
        if (e != hipSuccess) break;
        blas_finalize(pn, nwg, sc, 1, partial, st, stop);
      } else {
        if (step_p) blas_cg_xpay(ps, p.p, rs.p, n, sc, st, stop);
        e = eo_normal(op, lvl, ap.p, mid.p, p.p, st, stop);
        if (e != hipSuccess) break;
        blas_dot(ps, p.p, ap.p, n, partial, sc, 0, st, stop);
        blas_cg_update(ps, px, rs.p, xs.p, p.p, ap.p, n, partial, sc, st, stop);
      }
    }
    if (e != hipSuccess) break;
    restart = false;
    e = fetch();
    if (e != hipSuccess) break;
    iters = base + hsc->iters;
    if (!hsc->stop) continue;  // the batch ran through
    const double rr = hsc->rr;  // |r_s|^2
    if (!std::isfinite(rr)) {
      rc = QPHIX_AMD_ENOCONV;
      break;
    }
    // reliable update: fold the increment into x (with the last
    // iteration's x += alpha p, which the fused step leaves to the next
    // iteration) and clear it, then recompute the true residual in FP64
    blas_cg_fold(ps, px, p.p, xs.p, static_cast<double *>(x), sigma, n, sc, fuse, st);
    e = true_residual(&r2);
    if (e != hipSuccess) break;
    updates++;
    if (r2 < target2) {
      converged = true;
      break;
    }
    const double sig_new = mixed ? std::sqrt(r2 / (double)n) : 1.0;
    // beta from the true residual against the previous iterated one
    const double beta = r2 / (hsc->rr_prev * sigma * sigma);
    rmax = std::sqrt(r2);
    e = set_scalars(r2 / (sig_new * sig_new), beta * sigma / sig_new, sig_new, rmax);
    if (e != hipSuccess) break;
    // r in the new units, then p = r + beta p
    blas_cg_restart(ps, rs.p, p.p, static_cast<const double *>(r64.p), 1.0 / sig_new, n, sc, st);
    sigma = sig_new;
    last_len = iters - base;
    base = iters;
    restart = true;
  }
  mv = 2L * iters + 2L * updates;
  if (e == hipSuccess && !converged && rc == QPHIX_AMD_OK) {
    // report the true residual of the final iterate (its last x += alpha p
    // first, when an iteration is pending)
    blas_cg_fold(ps, px, p.p, xs.p, static_cast<double *>(x), sigma, n, sc, fuse && !restart, st);
    e = true_residual(&r2);
    converged = r2 < target2;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  auto t1 = std::chrono::steady_clock::now();
  res->iters = iters;
  res->reliable_updates = updates;
  res->rsd_final = b2 > 0 ? std::sqrt(r2 / b2) : 0.0;
  res->seconds = std::chrono::duration<double>(t1 - t0).count();
  res->mv_apps = mv;
  if (e != hipSuccess) return QPHIX_AMD_EHIP;
  if (rc != QPHIX_AMD_OK) return rc;
  return converged ? QPHIX_AMD_OK : QPHIX_AMD_ENOCONV;
}

}  // extern "C"
