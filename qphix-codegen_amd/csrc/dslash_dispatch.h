// dslash_dispatch.h -- converts a DslashCall into kernel parameters for one
// storage type; each dslash_<prec>.hip instantiates it once.
#pragma once

#include "dslash_kernel.h"
#include "internal.h"

namespace qamd {

template <typename S>
hipError_t run_dslash_typed(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st)
{
  using R = typename Traits<S>::R;
  DslashParams<R> p;
  p.out = c.out;
  p.in = c.in;
  p.chi = c.chi;
  p.gauge = c.gauge;
  p.clov = c.clov;
  p.g = ctx->g;
  p.cb = c.cb;
  p.nblocks = blocks_for(ctx->g.vcb, kDslashBlock);
  p.zs = p.ts = p.bpp = 1;
  // LDS rows: every work-group full (no lane leaves before the barrier),
  // and the rows a work-group touches plus one on either side fit the
  // staging array.  A work-group starts at a multiple of 128, so its offset
  // in a row is a multiple of gcd(128, Lx/2); every traversal order keeps
  // the 128 sites of a work-group consecutive.
  p.lds = 0;
  if (QAMD_LDS_XY && ctx->g.Lxh > 0 && ctx->g.vcb % kDslashBlock == 0) {
    int step = kDslashBlock, b = ctx->g.Lxh;
    while (b) {
      const int t = step % b;
      step = b;
      b = t;
    }
    int worst = 0;
    for (int o = 0; o < ctx->g.Lxh; o += step) {
      const int rows = (o + kDslashBlock - 1) / ctx->g.Lxh + 3;
      worst = rows * ctx->g.Lxh > worst ? rows * ctx->g.Lxh : worst;
    }
    p.lds = worst <= kLdsSites;
  }
  p.order = dslash_traversal(ctx);
  if (p.order != 0) p.bpp = ctx->g.sz / kDslashBlock;
  if (p.order == 1) {
    p.zs = ctx->g.Lz / kNumXcd;
  } else if (p.order == 2) {
    p.ts = ctx->g.Lt / kNumXcd;
  } else if (p.order == 3) {
    p.ts = ctx->g.Lt / kNumXcd;
    p.zs = ctx->g.Lz / QAMD_ZPARTS;
  }
  p.f_per = make_fastdiv((unsigned)(p.zs * ctx->g.Lt * p.bpp));
  p.f_zb = make_fastdiv((unsigned)(p.zs * p.bpp));
  p.f_bpp = make_fastdiv((unsigned)p.bpp);
  p.coeff_s = (R)c.coeff_s;
  p.coeff_tf = (R)c.coeff_tf;
  p.coeff_tb = (R)c.coeff_tb;
  p.tbound = (R)ctx->tbound;
  p.comm = ctx->comm_mask;
  p.tbc_mask = ctx->tbc_mask;
  for (int m = 0; m < 4; m++) {
    p.tbc_re[m] = (R)ctx->tbc_re[m];
    p.tbc_im[m] = (R)ctx->tbc_im[m];
  }
  p.alpha = (R)c.alpha;
  p.mu = (R)c.mu;
  p.muinv = (R)c.muinv;
  p.rho = (R)c.rho;
  p.one = (R)1;
  p.norm = c.norm;
  p.step = c.step;
  p.rin = c.rin;
  p.stop = c.stop;
  p.q_sh = c.q_sh;
  p.q_nvecs = c.q_nvecs;
  p.q_pxy = c.q_pxy;
  p.q_pxyz = c.q_pxyz;
  p.acc = c.acc;
  if (c.spl) {
    // QPhiX FourSpinorBlock arrays (FP64, FP32 or FP16, the reference's
    // types)
    if constexpr (std::is_same<S, __hip_bfloat16>::value) {
      return hipErrorInvalidValue;
    } else {
      // whole lane groups read as vectors (SPL = 2) stage the x/y rows like
      // the native layout; element by element (SPL = 1) they are not staged
      if (c.spl == 2) return dispatch_dslash<S, 2>(c.recon, c.op, c.achi, c.isign, p, st);
      if (c.spl != 1) return hipErrorInvalidValue;
      p.lds = 0;
      return dispatch_dslash<S, 1>(c.recon, c.op, c.achi, c.isign, p, st);
    }
  }
  return dispatch_dslash<S>(c.recon, c.op, c.achi, c.isign, p, st);
}

}  // namespace qamd
