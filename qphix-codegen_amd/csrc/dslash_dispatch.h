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
  // z-slab traversal needs whole work-groups per (z,t) plane and Lz % 8 == 0
  p.order = 0;
  p.zs = p.ts = p.bpp = 1;
  // LDS rows: whole rows per work-group, whole work-groups per (z,t) plane
  const int rows = ctx->g.Lxh > 0 ? kDslashBlock / ctx->g.Lxh : 0;
  p.lds = QAMD_LDS_XY && ctx->g.Lxh <= kLdsRowsMax && kDslashBlock % ctx->g.Lxh == 0 && rows > 0 &&
          ctx->g.Ly % rows == 0 && QAMD_BLOCK_ORDER == 0;
  if (ctx->g.sz % kDslashBlock == 0) {
    p.bpp = ctx->g.sz / kDslashBlock;
    if (QAMD_BLOCK_ORDER == 1 && ctx->g.Lz % kNumXcd == 0) {
      p.order = 1;
      p.zs = ctx->g.Lz / kNumXcd;
    } else if (QAMD_BLOCK_ORDER == 2 && ctx->g.Lt % kNumXcd == 0) {
      p.order = 2;
      p.ts = ctx->g.Lt / kNumXcd;
    } else if (QAMD_BLOCK_ORDER == 3 && ctx->g.Lt % kNumXcd == 0 && ctx->g.Lz % QAMD_ZPARTS == 0) {
      p.order = 3;
      p.ts = ctx->g.Lt / kNumXcd;
      p.zs = ctx->g.Lz / QAMD_ZPARTS;
    }
  }
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
  return dispatch_dslash<S>(c.recon, c.op, c.achi, c.isign, p, st);
}

}  // namespace qamd
