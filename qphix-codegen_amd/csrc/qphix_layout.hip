// qphix_layout.hip -- converters between QPhiX block layouts (the arrays a
// QPhiX host passes to the reference's kernels) and the native layout.
//
// Block types (codegen/include/decl_common.h:9-62): FourSpinorBlock
// FT[3][4][2][SOALEN], SU3MatrixBlock FT[8][2|3][3][2][VECLEN] holding the
// transpose of each hop matrix, CloverBlock {diag1, off_diag1, diag2,
// off_diag2}, FullCloverBlock {block1, block2}.  The block order and lane
// assignment are written out in include/qphix_amd.h next to
// qphix_amd_qphix_layout.  One thread per site of the checkerboard; each
// thread reads (or writes) its lane of every block row, so the lanes of a
// block are touched by consecutive threads.
#include <type_traits>

#include "internal.h"

namespace qamd {
namespace {

constexpr int kLayBlock = 256;

struct QLay {
  int veclen, soalen, ngy, nvecs, compress12;
  long pxy, pxyz;
};

// Block index and lane of site c in spinor / gauge-clover arrays.
__device__ __forceinline__ void spinor_slot(const QLay &q, const SiteCoord &c, long &blk, int &lane)
{
  blk = c.t * q.pxyz + c.z * q.pxy + (long)c.y * q.nvecs + c.xh / q.soalen;
  lane = c.xh % q.soalen;
}

__device__ __forceinline__ void matrix_slot(const QLay &q, const SiteCoord &c, long &blk, int &lane)
{
  blk = (c.t * q.pxyz + c.z * q.pxy) / q.ngy + (long)(c.y / q.ngy) * q.nvecs + c.xh / q.soalen;
  lane = (c.y % q.ngy) * q.soalen + c.xh % q.soalen;
}

template <typename T> __device__ __forceinline__ double qget(const T *p) { return (double)to_real(*p); }

template <typename S, typename T>
__global__ __launch_bounds__(kLayBlock) void spinor_from_qphix_k(S *dst, const T *src, Geom g, QLay q,
                                                                 int cb)
{
  const int idx = blockIdx.x * kLayBlock + threadIdx.x;
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, cb);
  long blk;
  int lane;
  spinor_slot(q, c, blk, lane);
  const T *b = src + blk * 24 * q.soalen + lane;
  double v[24];
#pragma unroll
  for (int col = 0; col < 3; col++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int ri = 0; ri < 2; ri++)
        v[(sp * 3 + col) * 2 + ri] = qget(b + ((col * 4 + sp) * 2 + ri) * q.soalen);
  store_reals<S, spinor_w<S>(), 0, 24>(dst, g.vcb, idx, v);
}

template <typename S, typename T>
__global__ __launch_bounds__(kLayBlock) void spinor_to_qphix_k(T *dst, const S *src, Geom g, QLay q, int cb)
{
  const int idx = blockIdx.x * kLayBlock + threadIdx.x;
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, cb);
  long blk;
  int lane;
  spinor_slot(q, c, blk, lane);
  T *b = dst + blk * 24 * q.soalen + lane;
  double v[24];
  load_reals<S, spinor_w<S>(), 0, 24>(src, g.vcb, idx, v);
#pragma unroll
  for (int col = 0; col < 3; col++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int ri = 0; ri < 2; ri++)
        b[((col * 4 + sp) * 2 + ri) * q.soalen] = to_storage<T>(v[(sp * 3 + col) * 2 + ri]);
}

// Direction MU of the native gauge unit, the rows of U_mu(x - mu)^dagger =
// conj(S_{2mu}) and of U_mu(x)^T = S_{2mu+1}, where S is the stored block
// entry (S[i][j] = M[j][i]).
template <typename S, int RECON, int NROW, int MU, typename T>
__device__ __forceinline__ void gauge_dir(S *dst, const T *b, int V, const Geom &g, int idx)
{
  double v[2 * RECON];
#pragma unroll
  for (int FB = 0; FB < 2; FB++) {
    const int d = 2 * MU + FB;
    double sr[3][3], si[3][3];
#pragma unroll
    for (int i = 0; i < NROW; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        sr[i][j] = qget(b + (((d * NROW + i) * 3 + j) * 2 + 0) * V);
        si[i][j] = qget(b + (((d * NROW + i) * 3 + j) * 2 + 1) * V);
      }
    if (NROW == 2) {  // row 2 = conj(row0 x row1)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        const double re = sr[0][j1] * sr[1][j2] - si[0][j1] * si[1][j2] -
                          (sr[0][j2] * sr[1][j1] - si[0][j2] * si[1][j1]);
        const double im = sr[0][j1] * si[1][j2] + si[0][j1] * sr[1][j2] -
                          (sr[0][j2] * si[1][j1] + si[0][j2] * sr[1][j1]);
        sr[2][j] = re;
        si[2][j] = -im;
      }
    }
    double *o = v + FB * RECON;
#pragma unroll
    for (int r = 0; r < RECON / 6; r++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        if (FB) {  // U^T = S
          o[(r * 3 + k) * 2] = sr[r][k];
          o[(r * 3 + k) * 2 + 1] = si[r][k];
        } else {  // U^dagger = conj(S)
          o[(r * 3 + k) * 2] = sr[r][k];
          o[(r * 3 + k) * 2 + 1] = -si[r][k];
        }
      }
  }
  store_reals<S, gauge_w<S, RECON>(), 2 * MU * RECON, 2 * RECON>(dst, g.vcb, idx, v);
}

template <typename S, int RECON, int NROW, typename T>
__global__ __launch_bounds__(kLayBlock) void gauge_from_qphix_k(S *dst, const T *src, Geom g, QLay q, int cb)
{
  const int idx = blockIdx.x * kLayBlock + threadIdx.x;
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, cb);
  long blk;
  int lane;
  matrix_slot(q, c, blk, lane);
  const T *b = src + blk * 8L * NROW * 6 * q.veclen + lane;
  gauge_dir<S, RECON, NROW, 0>(dst, b, q.veclen, g, idx);
  gauge_dir<S, RECON, NROW, 1>(dst, b, q.veclen, g, idx);
  gauge_dir<S, RECON, NROW, 2>(dst, b, q.veclen, g, idx);
  gauge_dir<S, RECON, NROW, 3>(dst, b, q.veclen, g, idx);
}

// CloverBlock (NR = 72) and FullCloverBlock (NR = 144) flatten, lane
// innermost, to the native per-site order [block][...] element by element;
// the native unit is written one W-vector at a time.
template <typename S, int NR, typename T>
__global__ __launch_bounds__(kLayBlock) void clover_from_qphix_k(S *dst, const T *src, Geom g, QLay q, int cb)
{
  const int idx = blockIdx.x * kLayBlock + threadIdx.x;
  if (idx >= g.vcb) return;
  const SiteCoord c = decode(g, idx, cb);
  long blk;
  int lane;
  matrix_slot(q, c, blk, lane);
  const T *b = src + blk * (long)NR * q.veclen + lane;
  constexpr int W = NR == 72 ? clover_w<S>() : fclover_w<S>();
  using Vec = typename RawVec<W * sizeof(S)>::T;
  Vec *out = reinterpret_cast<Vec *>(dst);
  for (int k0 = 0; k0 < NR; k0 += W) {
    S tmp[W];
#pragma unroll
    for (int e = 0; e < W; e++) tmp[e] = to_storage<S>(qget(b + (long)(k0 + e) * q.veclen));
    Vec v;
    __builtin_memcpy(&v, tmp, sizeof(Vec));
    out[(unsigned)((k0 / W) * g.vcb + idx)] = v;
  }
}

template <typename F> hipError_t with_ft(int ft, F &&f)
{
  switch (ft) {
    case QPHIX_AMD_F64: return f((double *)nullptr);
    case QPHIX_AMD_F32: return f((float *)nullptr);
    case QPHIX_AMD_F16: return f((__half *)nullptr);
  }
  return hipErrorInvalidValue;
}

template <typename F> hipError_t with_native(int prec, F &&f)
{
  switch (prec) {
    case QPHIX_AMD_F64: return f((double *)nullptr);
    case QPHIX_AMD_F32: return f((float *)nullptr);
    case QPHIX_AMD_F16: return f((__half *)nullptr);
    case QPHIX_AMD_BF16: return f((__hip_bfloat16 *)nullptr);
  }
  return hipErrorInvalidValue;
}

enum Dir { FROM_QPHIX, TO_QPHIX };

hipError_t convert(const qphix_amd_ctx *ctx, const QLay &q, int ft, int field, int prec, int recon,
                   void *dst, const void *src, int cb, Dir dir, hipStream_t st)
{
  const Geom &g = ctx->g;
  const dim3 grid(blocks_for(g.vcb, kLayBlock)), block(kLayBlock);
  return with_ft(ft, [&](auto *tq) -> hipError_t {
    using T = std::remove_pointer_t<decltype(tq)>;
    return with_native(prec, [&](auto *sn) -> hipError_t {
      using S = std::remove_pointer_t<decltype(sn)>;
      switch (field) {
        case QPHIX_AMD_FIELD_SPINOR:
          if (dir == FROM_QPHIX)
            hipLaunchKernelGGL((spinor_from_qphix_k<S, T>), grid, block, 0, st, (S *)dst,
                               (const T *)src, g, q, cb);
          else
            hipLaunchKernelGGL((spinor_to_qphix_k<S, T>), grid, block, 0, st, (T *)dst,
                               (const S *)src, g, q, cb);
          break;
        case QPHIX_AMD_FIELD_GAUGE:
          if (recon == 12 && q.compress12)
            hipLaunchKernelGGL((gauge_from_qphix_k<S, 12, 2, T>), grid, block, 0, st, (S *)dst,
                               (const T *)src, g, q, cb);
          else if (recon == 12)
            hipLaunchKernelGGL((gauge_from_qphix_k<S, 12, 3, T>), grid, block, 0, st, (S *)dst,
                               (const T *)src, g, q, cb);
          else if (q.compress12)
            hipLaunchKernelGGL((gauge_from_qphix_k<S, 18, 2, T>), grid, block, 0, st, (S *)dst,
                               (const T *)src, g, q, cb);
          else
            hipLaunchKernelGGL((gauge_from_qphix_k<S, 18, 3, T>), grid, block, 0, st, (S *)dst,
                               (const T *)src, g, q, cb);
          break;
        case QPHIX_AMD_FIELD_CLOVER:
          hipLaunchKernelGGL((clover_from_qphix_k<S, 72, T>), grid, block, 0, st, (S *)dst,
                             (const T *)src, g, q, cb);
          break;
        case QPHIX_AMD_FIELD_FULLCLOVER:
          hipLaunchKernelGGL((clover_from_qphix_k<S, 144, T>), grid, block, 0, st, (S *)dst,
                             (const T *)src, g, q, cb);
          break;
        default:
          return hipErrorInvalidValue;
      }
      return hipGetLastError();
    });
  });
}

// Validates the layout against the lattice; fills q.
bool make_layout(const qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, QLay &q)
{
  if (!ctx || !lay) return false;
  if (lay->ft != QPHIX_AMD_F64 && lay->ft != QPHIX_AMD_F32 && lay->ft != QPHIX_AMD_F16) return false;
  if (lay->soalen < 1 || lay->veclen < lay->soalen || lay->veclen % lay->soalen) return false;
  if (lay->compress12 != 0 && lay->compress12 != 1) return false;
  const Geom &g = ctx->g;
  if (g.Lxh % lay->soalen) return false;
  const int ngy = lay->veclen / lay->soalen;
  if (g.Ly % ngy || lay->pad_xy < 0 || lay->pad_xyz < 0 || lay->pad_xy % ngy || lay->pad_xyz % ngy)
    return false;
  q.veclen = lay->veclen;
  q.soalen = lay->soalen;
  q.ngy = ngy;
  q.nvecs = g.Lxh / lay->soalen;
  q.compress12 = lay->compress12;
  q.pxy = (long)g.Ly * q.nvecs + lay->pad_xy;
  q.pxyz = (long)g.Lz * q.pxy + lay->pad_xyz;
  return true;
}

size_t qphix_bytes(const qphix_amd_ctx *ctx, const QLay &q, int ft, int field)
{
  const size_t s = ft == QPHIX_AMD_F64 ? 8 : ft == QPHIX_AMD_F32 ? 4 : 2;
  const size_t nsp = (size_t)ctx->g.Lt * q.pxyz;  // spinor blocks
  const size_t nmat = nsp / q.ngy;                // gauge / clover blocks
  switch (field) {
    case QPHIX_AMD_FIELD_SPINOR: return nsp * 24 * q.soalen * s;
    case QPHIX_AMD_FIELD_GAUGE: return nmat * 8 * (q.compress12 ? 2 : 3) * 6 * q.veclen * s;
    case QPHIX_AMD_FIELD_CLOVER: return nmat * 72 * q.veclen * s;
    case QPHIX_AMD_FIELD_FULLCLOVER: return nmat * 144 * q.veclen * s;
  }
  return 0;
}

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

int entry(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int field, int prec, int recon,
          void *dst, const void *src, int cb, Dir dir, void *stream)
{
  QLay q;
  if (!make_layout(ctx, lay, q)) return QPHIX_AMD_EINVAL;
  if (!dst || !src || prec < QPHIX_AMD_F64 || prec > QPHIX_AMD_BF16 || (cb != 0 && cb != 1))
    return QPHIX_AMD_EINVAL;
  if (field == QPHIX_AMD_FIELD_GAUGE && recon != 12 && recon != 18) return QPHIX_AMD_EINVAL;
  // the native side is read / written with 16-byte vectors
  if (!aligned16(dir == FROM_QPHIX ? dst : src)) return QPHIX_AMD_EALIGN;
  const hipError_t e = convert(ctx, q, lay->ft, field, prec, recon, dst, src, cb, dir,
                               static_cast<hipStream_t>(stream));
  return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP;
}

}  // namespace

bool qphix_spinor_geometry(const qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, QSpinorGeom &out)
{
  QLay q;
  if (!make_layout(ctx, lay, q)) return false;
  int sh = 0;
  while ((1 << sh) < q.soalen) sh++;
  if ((1 << sh) != q.soalen) return false;  // the operator kernels shift by log2(soalen)
  const long elems = (long)ctx->g.Lt * q.pxyz * 24 * q.soalen;
  if (elems >= (1L << 31)) return false;  // 32-bit element offsets
  out.sh = sh;
  out.nvecs = q.nvecs;
  out.pxy = (int)q.pxy;
  out.pxyz = (int)q.pxyz;
  out.sites = elems / 24;
  return true;
}

}  // namespace qamd

extern "C" {

size_t qphix_amd_qphix_field_bytes(const qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay,
                                   int field)
{
  qamd::QLay q;
  if (!qamd::make_layout(ctx, lay, q)) return 0;
  return qamd::qphix_bytes(ctx, q, lay->ft, field);
}

int qphix_amd_spinor_from_qphix(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, void *dst,
                                int prec, const void *src, int cb, void *stream)
{
  return qamd::entry(ctx, lay, QPHIX_AMD_FIELD_SPINOR, prec, 0, dst, src, cb, qamd::FROM_QPHIX,
                     stream);
}

int qphix_amd_spinor_to_qphix(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, void *dst,
                              const void *src, int prec, int cb, void *stream)
{
  return qamd::entry(ctx, lay, QPHIX_AMD_FIELD_SPINOR, prec, 0, dst, src, cb, qamd::TO_QPHIX,
                     stream);
}

int qphix_amd_gauge_from_qphix(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, void *dst,
                               int prec, int recon, const void *src, int cb, void *stream)
{
  return qamd::entry(ctx, lay, QPHIX_AMD_FIELD_GAUGE, prec, recon, dst, src, cb,
                     qamd::FROM_QPHIX, stream);
}

int qphix_amd_clover_from_qphix(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, void *dst,
                                int prec, const void *src, int cb, void *stream)
{
  return qamd::entry(ctx, lay, QPHIX_AMD_FIELD_CLOVER, prec, 0, dst, src, cb, qamd::FROM_QPHIX,
                     stream);
}

int qphix_amd_full_clover_from_qphix(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay,
                                     void *dst, int prec, const void *src, int cb, void *stream)
{
  return qamd::entry(ctx, lay, QPHIX_AMD_FIELD_FULLCLOVER, prec, 0, dst, src, cb,
                     qamd::FROM_QPHIX, stream);
}

}  // extern "C"
