// pack.hip -- converters between the lexicographic FP64 device layout and
// the native per-checkerboard layout (site-fastest vector rows).
//
// The gauge packer stores, at each site x of the checkerboard, the 8 hop
// matrices in the reference's direction order d = 2*mu + fb
// (codegen/dslash.cc:60-61,232-234): fb = 0 holds the rows of
// U_mu(x - mu)^dagger (the matrix the back hop applies), fb = 1 the rows of
// U_mu(x)^T (the kernel applies its transpose).  The reference keeps 8 links
// per site as well (codegen/include/decl_common.h:13) and stores the
// transpose of every link (codegen/dslash_common.cc:1678-1727), so with 12
// reals both keep the same two rows -- columns 0, 1 of each link -- and
// rebuild the same third one.
#include "internal.h"

namespace qamd {

constexpr int kPackBlock = 256;

template <typename S>
__global__ __launch_bounds__(kPackBlock) void pack_spinor_k(S *dst, const double *src, Geom g, int cb)
{
  const int idx = blockIdx.x * kPackBlock + threadIdx.x;
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, cb);
  const long lex = lex_index(g, c);
  double v[24];
#pragma unroll
  for (int k = 0; k < 24; k++) v[k] = src[lex * 24 + k];
  store_reals<S, spinor_w<S>(), 0, 24>(dst, g.vcb, idx, v);
}

template <typename S>
__global__ __launch_bounds__(kPackBlock) void unpack_spinor_k(double *dst, const S *src, Geom g, int cb)
{
  const int idx = blockIdx.x * kPackBlock + threadIdx.x;
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, cb);
  const long lex = lex_index(g, c);
  double v[24];
  load_reals<S, spinor_w<S>(), 0, 24>(src, g.vcb, idx, v);
#pragma unroll
  for (int k = 0; k < 24; k++) dst[lex * 24 + k] = v[k];
}

__device__ __forceinline__ long lex_shift(const Geom &g, SiteCoord c, int mu, int d)
{
  if (mu == 0) c.x = (c.x + d + g.Lx) % g.Lx;
  if (mu == 1) c.y = (c.y + d + g.Ly) % g.Ly;
  if (mu == 2) c.z = (c.z + d + g.Lz) % g.Lz;
  if (mu == 3) c.t = (c.t + d + g.Lt) % g.Lt;
  return lex_index(g, c);
}

// Both links of direction MU form one storage unit (d = 2 MU, 2 MU + 1).
template <typename S, int RECON, int MU>
__device__ __forceinline__ void pack_link_pair(S *dst, const double *U, const Geom &g,
                                               const SiteCoord &c, int idx)
{
  double v[2 * RECON];
#pragma unroll
  for (int FB = 0; FB < 2; FB++) {
    const long s = FB ? lex_index(g, c) : lex_shift(g, c, MU, -1);
    const double *m = U + (s * 4 + MU) * 18;
    double *o = v + FB * RECON;
#pragma unroll
    for (int r = 0; r < RECON / 6; r++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (FB) {  // transpose: (U^T)[r][k] = U[k][r]
          o[(r * 3 + k) * 2] = m[(k * 3 + r) * 2];
          o[(r * 3 + k) * 2 + 1] = m[(k * 3 + r) * 2 + 1];
        } else {  // dagger: (U^+)[r][k] = conj(U[k][r])
          o[(r * 3 + k) * 2] = m[(k * 3 + r) * 2];
          o[(r * 3 + k) * 2 + 1] = -m[(k * 3 + r) * 2 + 1];
        }
      }
  }
  store_reals<S, gauge_w<S, RECON>(), 2 * MU * RECON, 2 * RECON>(dst, g.vcb, idx, v);
}

template <typename S, int RECON>
__global__ __launch_bounds__(kPackBlock) void pack_gauge_k(S *dst, const double *U, Geom g, int cb)
{
  const int idx = blockIdx.x * kPackBlock + threadIdx.x;
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, cb);
  pack_link_pair<S, RECON, 0>(dst, U, g, c, idx);
  pack_link_pair<S, RECON, 1>(dst, U, g, c, idx);
  pack_link_pair<S, RECON, 2>(dst, U, g, c, idx);
  pack_link_pair<S, RECON, 3>(dst, U, g, c, idx);
}

// Clover (72 reals) or full clover (144 reals) of a site as one unit.
template <typename S, int NB>
__global__ __launch_bounds__(kPackBlock) void pack_clover_k(S *dst, const double *src, Geom g, int cb)
{
  const int idx = blockIdx.x * kPackBlock + threadIdx.x;
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, cb);
  const long lex = lex_index(g, c);
  double v[2 * NB];
#pragma unroll
  for (int k = 0; k < 2 * NB; k++) v[k] = src[lex * 2 * NB + k];
  constexpr int W = NB == 36 ? clover_w<S>() : fclover_w<S>();
  store_reals<S, W, 0, 2 * NB>(dst, g.vcb, idx, v);
}

template <typename S>
hipError_t pack_typed(const Geom &g, int field, int recon, void *dst, const double *src, int cb,
                      hipStream_t st)
{
  const int grid = blocks_for(g.vcb, kPackBlock);
  S *d = static_cast<S *>(dst);
  switch (field) {
    case QPHIX_AMD_FIELD_SPINOR:
      hipLaunchKernelGGL(pack_spinor_k<S>, dim3(grid), dim3(kPackBlock), 0, st, d, src, g, cb);
      break;
    case QPHIX_AMD_FIELD_GAUGE:
      if (recon == 12)
        hipLaunchKernelGGL((pack_gauge_k<S, 12>), dim3(grid), dim3(kPackBlock), 0, st, d, src, g, cb);
      else
        hipLaunchKernelGGL((pack_gauge_k<S, 18>), dim3(grid), dim3(kPackBlock), 0, st, d, src, g, cb);
      break;
    case QPHIX_AMD_FIELD_CLOVER:
      hipLaunchKernelGGL((pack_clover_k<S, 36>), dim3(grid), dim3(kPackBlock), 0, st, d, src, g, cb);
      break;
    case QPHIX_AMD_FIELD_FULLCLOVER:
      hipLaunchKernelGGL((pack_clover_k<S, 72>), dim3(grid), dim3(kPackBlock), 0, st, d, src, g, cb);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t run_pack(const qphix_amd_ctx *ctx, int field, int prec, int recon, void *dst,
                    const double *src, int cb, hipStream_t st)
{
  switch (prec) {
    case QPHIX_AMD_F64: return pack_typed<double>(ctx->g, field, recon, dst, src, cb, st);
    case QPHIX_AMD_F32: return pack_typed<float>(ctx->g, field, recon, dst, src, cb, st);
    case QPHIX_AMD_F16: return pack_typed<__half>(ctx->g, field, recon, dst, src, cb, st);
    case QPHIX_AMD_BF16: return pack_typed<__hip_bfloat16>(ctx->g, field, recon, dst, src, cb, st);
  }
  return hipErrorInvalidValue;
}

hipError_t run_unpack_spinor(const qphix_amd_ctx *ctx, int prec, double *dst, const void *src,
                             int cb, hipStream_t st)
{
  const Geom &g = ctx->g;
  const int grid = blocks_for(g.vcb, kPackBlock);
  switch (prec) {
    case QPHIX_AMD_F64:
      hipLaunchKernelGGL(unpack_spinor_k<double>, dim3(grid), dim3(kPackBlock), 0, st, dst,
                         static_cast<const double *>(src), g, cb);
      break;
    case QPHIX_AMD_F32:
      hipLaunchKernelGGL(unpack_spinor_k<float>, dim3(grid), dim3(kPackBlock), 0, st, dst,
                         static_cast<const float *>(src), g, cb);
      break;
    case QPHIX_AMD_F16:
      hipLaunchKernelGGL(unpack_spinor_k<__half>, dim3(grid), dim3(kPackBlock), 0, st, dst,
                         static_cast<const __half *>(src), g, cb);
      break;
    case QPHIX_AMD_BF16:
      hipLaunchKernelGGL(unpack_spinor_k<__hip_bfloat16>, dim3(grid), dim3(kPackBlock), 0, st,
                         dst, static_cast<const __hip_bfloat16 *>(src), g, cb);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace qamd
