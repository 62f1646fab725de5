// gauge.hip -- FP64 link reunitarisation on the lexicographic layout.
// No counterpart in the reference.  Each link is projected back onto SU(3):
// row 0 normalised, row 1 orthogonalised against row 0 and normalised,
// row 2 = conj(row0 x row1) -- the same third-row rule the 12-real link
// compression relies on (codegen/dslash_common.cc:1664-1674).
#include "internal.h"

namespace qamd {
namespace {

__global__ __launch_bounds__(256) void reunitarize_k(double *U, long nlinks)
{
  const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= nlinks) return;
  double *m = U + l * 18;
  cx<double> r0[3], r1[3], r2[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    r0[j] = {m[j * 2], m[j * 2 + 1]};
    r1[j] = {m[6 + j * 2], m[6 + j * 2 + 1]};
  }
  double n0 = 0.0;
#pragma unroll
  for (int j = 0; j < 3; j++) n0 += r0[j].re * r0[j].re + r0[j].im * r0[j].im;
  n0 = 1.0 / sqrt(n0);
#pragma unroll
  for (int j = 0; j < 3; j++) r0[j] = n0 * r0[j];
  cx<double> d = {0.0, 0.0};  // r0^dagger r1
#pragma unroll
  for (int j = 0; j < 3; j++) d = d + cmul_conj(r0[j], r1[j]);
#pragma unroll
  for (int j = 0; j < 3; j++) r1[j] = r1[j] - d * r0[j];
  double n1 = 0.0;
#pragma unroll
  for (int j = 0; j < 3; j++) n1 += r1[j].re * r1[j].re + r1[j].im * r1[j].im;
  n1 = 1.0 / sqrt(n1);
#pragma unroll
  for (int j = 0; j < 3; j++) r1[j] = n1 * r1[j];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
    r2[j] = conj(r0[j1] * r1[j2] - r0[j2] * r1[j1]);
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    m[j * 2] = r0[j].re;
    m[j * 2 + 1] = r0[j].im;
    m[6 + j * 2] = r1[j].re;
    m[6 + j * 2 + 1] = r1[j].im;
    m[12 + j * 2] = r2[j].re;
    m[12 + j * 2 + 1] = r2[j].im;
  }
}

}  // namespace
}  // namespace qamd

extern "C" int qphix_amd_reunitarize(qphix_amd_ctx *ctx, double *u_lex, void *stream)
{
  if (!ctx || !u_lex) return QPHIX_AMD_EINVAL;
  const long nlinks = (long)ctx->g.vcb * 2 * 4;
  hipLaunchKernelGGL(qamd::reunitarize_k, dim3(qamd::blocks_for(nlinks, 256)), dim3(256), 0,
                     (hipStream_t)stream, u_lex, nlinks);
  return hipGetLastError() == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP;
}
