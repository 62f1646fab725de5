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
  p.coeff_s = (R)c.coeff_s;
  p.coeff_tf = (R)c.coeff_tf;
  p.coeff_tb = (R)c.coeff_tb;
  p.tbound = (R)ctx->tbound;
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
