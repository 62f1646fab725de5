// face.hip -- the multi-GPU face operations of the dslash family: packing the
// projected half-spinors of a boundary slice (the reference's
// face_proj_dir_{plus,minus}) and completing the boundary sites from a
// received face (face_finish_dir_*, face_clov_finish_dir_*,
// tm_face_finish_dir_*, tm_clov_face_finish_dir_*; codegen/jinja/
// *_dslash_decl.h.j2, bodies from codegen/dslash.cc:433-553 and
// codegen/dslash_common.cc:1934-1993), plus the split-dimension mask that
// turns every operator call into the interior (body) part.
//
// Face buffers: one half-spinor (12 reals, [spin 0|1][colour][re|im], the
// reference's PackHalfSpinor order, codegen/data_types.cc:508-543) per face
// site, stored site-fastest in W-vectors like the native spinor.  Face site
// f of dimension mu enumerates the sites of one checkerboard on the slice
// x_mu = 0 or L_mu - 1 by their other coordinates:
//   mu = x: f = (t Lz + z) Ly/2 + y/2       mu = y: f = (t Lz + z) Lx/2 + x/2
//   mu = z: f = (t Ly + y) Lx/2 + x/2       mu = t: f = (z Ly + y) Lx/2 + x/2
// so a site and its neighbour across the cut share f.
#include <type_traits>

#include "dslash_kernel.h"
#include "internal.h"

namespace qamd {
namespace {

// One wave per work-group: a face is small (16,384 sites for a 32^3 t-face),
// and its sites' chains of dependent loads are latency-bound, so it is spread
// over as many CUs as possible (256 work-groups for that face, one per CU).
constexpr int kFaceBlock = 64;

template <typename S> constexpr int half_w() { return pick_w<S>(12); }

__host__ __device__ inline int face_count(const Geom &g, int mu)
{
  switch (mu) {
    case 0: return g.Lt * g.Lz * (g.Ly / 2);
    case 1: return g.Lt * g.Lz * g.Lxh;
    case 2: return g.Lt * g.Ly * g.Lxh;
    default: return g.Lz * g.Ly * g.Lxh;
  }
}

// Site of checkerboard cb with face index f on the slice x_mu = 0 (side 0)
// or L_mu - 1 (side 1).
__device__ __forceinline__ SiteCoord face_site(const Geom &g, int mu, int side, int f, int cb)
{
  SiteCoord c;
  if (mu == 0) {
    const int lyh = g.Ly / 2;
    const int yh = f % lyh, r = f / lyh;
    c.z = r % g.Lz;
    c.t = r / g.Lz;
    c.x = side ? g.Lx - 1 : 0;
    c.y = 2 * yh + ((cb + c.x + c.z + c.t) & 1);
    c.xh = c.x >> 1;
    return c;
  }
  c.xh = f % g.Lxh;
  const int r = f / g.Lxh;
  if (mu == 1) {
    c.z = r % g.Lz;
    c.t = r / g.Lz;
    c.y = side ? g.Ly - 1 : 0;
  } else if (mu == 2) {
    c.y = r % g.Ly;
    c.t = r / g.Ly;
    c.z = side ? g.Lz - 1 : 0;
  } else {
    c.y = r % g.Ly;
    c.z = r / g.Ly;
    c.t = side ? g.Lt - 1 : 0;
  }
  c.x = 2 * c.xh + ((cb + c.y + c.z + c.t) & 1);
  return c;
}

__device__ __forceinline__ int native_index(const Geom &g, const SiteCoord &c)
{
  return ((c.t * g.Lz + c.z) * g.Ly + c.y) * g.Lxh + c.xh;
}

// Spinor addressing of the face kernels: the native layout, or (sh >= 0)
// QPhiX FourSpinorBlock arrays of soalen 1 << sh with block strides nvecs
// (y), pxy (z), pxyz (t), read and written element by element in place
// (codegen/include/decl_common.h:11).
struct FaceSpinorGeom {
  int sh, nvecs, pxy, pxyz;
};

__device__ __forceinline__ int qphix_offset(const FaceSpinorGeom &q, const SiteCoord &c)
{
  const int blk = c.t * q.pxyz + c.z * q.pxy + c.y * q.nvecs + (c.xh >> q.sh);
  return ((blk * 24) << q.sh) + (c.xh & ((1 << q.sh) - 1));
}

template <typename S, typename R>
__device__ __forceinline__ void load_site(const S *a, const Geom &g, const FaceSpinorGeom &q, const SiteCoord &c,
                                          R *v)
{
  if (q.sh >= 0)
    load_q_reals<S, R>(a, qphix_offset(q, c), q.sh, v);
  else
    load_reals<S, spinor_w<S>(), 0, 24>(a, g.vcb, native_index(g, c), v);
}

template <typename R>
__device__ __forceinline__ void project_rt(int mu, int sgn, const R *psi, cx<R> h[2][3])
{
  switch (mu * 2 + (sgn > 0)) {
    case 0: project<0, -1>(psi, h); break;
    case 1: project<0, 1>(psi, h); break;
    case 2: project<1, -1>(psi, h); break;
    case 3: project<1, 1>(psi, h); break;
    case 4: project<2, -1>(psi, h); break;
    case 5: project<2, 1>(psi, h); break;
    case 6: project<3, -1>(psi, h); break;
    default: project<3, 1>(psi, h); break;
  }
}

template <typename R>
__device__ __forceinline__ void reconstruct_rt(int mu, int sgn, cx<R> acc[4][3], const cx<R> w[2][3], R b)
{
  switch (mu * 2 + (sgn > 0)) {
    case 0: reconstruct<0, -1>(acc, w, b); break;
    case 1: reconstruct<0, 1>(acc, w, b); break;
    case 2: reconstruct<1, -1>(acc, w, b); break;
    case 3: reconstruct<1, 1>(acc, w, b); break;
    case 4: reconstruct<2, -1>(acc, w, b); break;
    case 5: reconstruct<2, 1>(acc, w, b); break;
    case 6: reconstruct<3, -1>(acc, w, b); break;
    default: reconstruct<3, 1>(acc, w, b); break;
  }
}

// buf[f] = (1 + sgn gamma_mu) in(site f), in on checkerboard cb_in; side 2:
// both faces in one launch, face 0 (sgn = -isign) then face 1 (sgn = isign),
// each in its own nface-site buffer, one after the other.
template <typename S>
__global__ __launch_bounds__(kFaceBlock) void face_pack_k(S *buf, const S *in, Geom g, FaceSpinorGeom q, int cb_in,
                                                          int mu, int side, int isign, int nface)
{
  using R = typename Traits<S>::R;
  int f = blockIdx.x * kFaceBlock + threadIdx.x;
  if (side == 2) {
    if (f >= 2 * nface) return;
    side = f >= nface;
    if (side) {
      f -= nface;
      buf += (size_t)12 * nface;
    }
  } else if (f >= nface) {
    return;
  }
  // sent backwards: the receiver's forward hop, projector (1 - isign gamma);
  // sent forwards: its backward hop, (1 + isign gamma) (codegen/dslash.cc:451-457)
  const int sgn = side == 0 ? -isign : isign;
  const SiteCoord c = face_site(g, mu, side, f, cb_in);
  R psi[24];
  load_site(in, g, q, c, psi);
  cx<R> h[2][3];
  project_rt(mu, sgn, psi, h);
  R o[12];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      o[(s * 3 + k) * 2] = h[s][k].re;
      o[(s * 3 + k) * 2 + 1] = h[s][k].im;
    }
  store_reals<S, half_w<S>(), 0, 12>(buf, nface, f, o);
}

template <typename R> struct FaceArgs {
  int cb, mu, side, isign, nface, tbc;
  R beta, ph_re, ph_im, tm_mu, tm_muinv, one;
};

// out(site f) += the face's hop, in the reference's face finish order
// (codegen/dslash.cc:463-553, codegen/dslash_common.cc:1098-1186,1188-1232,
// 1234-1300):
//   w = U h (the hop's cdot3 chains; U_mu(x - mu)^dagger from the back
//   neighbour, U_mu(x) from the front), times the twisted-boundary phase
//   (conjugated from the back), then
//   Wilson:  out -= beta R(w), one fused multiply-add per real into out;
//   clover:  d = beta R(w), out += A d, each row's chain starting from out;
//   TM:      d = -beta R(w), out += muinv (1 - isign i mu g5) d, each real
//            one fused multiply-add of muinv into out;
//   TM clover: d = beta R(w), out += B d, each row's chain starting from out;
// R the reconstruction of the half-spinor to the full spinor.
template <typename S, int RECON, int OP>
__global__ __launch_bounds__(kFaceBlock) void face_finish_k(S *out, const S *buf, const S *gauge,
                                                            const S *clov, Geom g, FaceSpinorGeom q,
                                                            FaceArgs<typename Traits<S>::R> a)
{
  using R = typename Traits<S>::R;
  int f = blockIdx.x * kFaceBlock + threadIdx.x;
  int side = a.side;
  if (side == 2) {  // both faces: face 0's buffer, then face 1's
    if (f >= 2 * a.nface) return;
    side = f >= a.nface;
    if (side) {
      f -= a.nface;
      buf += (size_t)12 * a.nface;
    }
  } else if (f >= a.nface) {
    return;
  }
  const SiteCoord c = face_site(g, a.mu, side, f, a.cb);
  const int idx = native_index(g, c);
  const int vcb = g.vcb;
  R hr[12];
  load_reals<S, half_w<S>(), 0, 12>(buf, a.nface, f, hr);
  cx<R> h[2][3];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int k = 0; k < 3; k++) h[s][k] = {hr[(s * 3 + k) * 2], hr[(s * 3 + k) * 2 + 1]};
  const R nz = -R(0) * a.one;  // opaque -0.0, as in the operator kernels
  cx<R> ub[3][3], uf[3][3];
  switch (a.mu) {
    case 0: load_link_pair<S, RECON, 0>(gauge, vcb, idx, ub, uf, nz); break;
    case 1: load_link_pair<S, RECON, 1>(gauge, vcb, idx, ub, uf, nz); break;
    case 2: load_link_pair<S, RECON, 2>(gauge, vcb, idx, ub, uf, nz); break;
    default: load_link_pair<S, RECON, 3>(gauge, vcb, idx, ub, uf, nz); break;
  }
  const cx<R> ph = {a.ph_re, side ? a.ph_im : -a.ph_im};
  cx<R> w[2][3];
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const cx<R> u0 = side ? uf[i][0] : ub[i][0];
      const cx<R> u1 = side ? uf[i][1] : ub[i][1];
      const cx<R> u2 = side ? uf[i][2] : ub[i][2];
      if constexpr (std::is_same<S, __half>::value)
        w[k][i] = cdot3(u0, u1, u2, h[k][0], h[k][1], h[k][2], nz);
      else
        w[k][i] = cdot3(u0, u1, u2, h[k][0], h[k][1], h[k][2]);
      if (a.tbc) w[k][i] = cmul_ord(ph, w[k][i]);
    }
  const int sgn = side ? -a.isign : a.isign;
  R o[24];
  load_site(out, g, q, c, o);
  cx<R> acc[4][3];
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int k = 0; k < 3; k++) acc[s][k] = {o[(s * 3 + k) * 2], o[(s * 3 + k) * 2 + 1]};
  if (OP == OP_WILSON) {
    reconstruct_rt(a.mu, sgn, acc, w, -a.beta);
  } else {
    cx<R> d[4][3];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
      for (int k = 0; k < 3; k++) d[s][k] = {R(0), R(0)};
    reconstruct_rt(a.mu, sgn, d, w, OP == OP_TM ? -a.beta : a.beta);
    if (OP == OP_CLOVER) {
      clover_packed<S, true>(clov, vcb, idx, d, acc);
    } else if (OP == OP_TMCLOVER) {
      clover_full<S, true>(clov, vcb, idx, d, acc);
    } else {  // (1 - isign i mu g5) d by one fused multiply-add per real, then muinv
      const R m = -(R)a.isign * a.tm_mu;
#pragma unroll
      for (int s = 0; s < 4; s++) {
        const R ms = s < 2 ? m : -m;
#pragma unroll
        for (int k = 0; k < 3; k++) {
          const cx<R> v = d[s][k];
          acc[s][k] = {fmx(a.tm_muinv, fmx(-ms, v.im, v.re), acc[s][k].re),
                       fmx(a.tm_muinv, fmx(ms, v.re, v.im), acc[s][k].im)};
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int k = 0; k < 3; k++) {
      o[(s * 3 + k) * 2] = acc[s][k].re;
      o[(s * 3 + k) * 2 + 1] = acc[s][k].im;
    }
  if (q.sh >= 0) {
    const int qo = qphix_offset(q, c);
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int cc = 0; cc < 3; cc++)
#pragma unroll
        for (int r = 0; r < 2; r++)
          out[qo + (((cc * 4 + sp) * 2 + r) << q.sh)] = to_storage<S>(o[(sp * 3 + cc) * 2 + r]);
  } else {
    store_reals<S, spinor_w<S>(), 0, 24>(out, vcb, idx, o);
  }
}

template <typename F> hipError_t with_prec(int prec, F &&fn)
{
  switch (prec) {
    case QPHIX_AMD_F64: return fn((double *)nullptr);
    case QPHIX_AMD_F32: return fn((float *)nullptr);
    case QPHIX_AMD_F16: return fn((__half *)nullptr);
    case QPHIX_AMD_BF16: return fn((__hip_bfloat16 *)nullptr);
  }
  return hipErrorInvalidValue;
}

template <typename S, int OP>
hipError_t launch_finish(int recon, const Geom &g, const FaceSpinorGeom &q, S *out, const S *buf, const S *gauge,
                         const S *clov, const FaceArgs<typename Traits<S>::R> &a, hipStream_t st)
{
  const dim3 grid(blocks_for(a.side == 2 ? 2L * a.nface : a.nface, kFaceBlock)), block(kFaceBlock);
  if (recon == 12)
    hipLaunchKernelGGL((face_finish_k<S, 12, OP>), grid, block, 0, st, out, buf, gauge, clov, g, q, a);
  else
    hipLaunchKernelGGL((face_finish_k<S, 18, OP>), grid, block, 0, st, out, buf, gauge, clov, g, q, a);
  return hipGetLastError();
}

constexpr FaceSpinorGeom kNativeSpinor = {-1, 0, 0, 0};

// The FourSpinorBlock geometry of a QPhiX layout, or false when the lattice
// does not admit it (or its FT is not a QPhiX one).
bool face_qgeom(const qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, FaceSpinorGeom &out)
{
  QSpinorGeom q;
  if (!lay || lay->ft == QPHIX_AMD_BF16 || !qphix_spinor_geometry(ctx, lay, q)) return false;
  out = {q.sh, q.nvecs, q.pxy, q.pxyz};
  return true;
}

int storage_size(int prec) { return prec == QPHIX_AMD_F64 ? 8 : prec == QPHIX_AMD_F32 ? 4 : 2; }

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

int status(hipError_t e) { return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP; }

// lay: null for native spinors, else out is a QPhiX array of that layout
int finish(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int op, int prec, int recon, int isign, int cb,
           int dim, int fb, void *out, const void *buf, const void *gauge, const void *clov, double beta,
           double mu, double muinv, void *stream)
{
  if (!ctx || prec < QPHIX_AMD_F64 || prec > QPHIX_AMD_BF16 || (recon != 12 && recon != 18))
    return QPHIX_AMD_EINVAL;
  if ((isign != 1 && isign != -1) || (cb != 0 && cb != 1) || dim < 0 || dim > 3 || fb < 0 || fb > 2)
    return QPHIX_AMD_EINVAL;
  if (!out || !buf || !gauge || out == buf) return QPHIX_AMD_EINVAL;
  if ((op == OP_CLOVER || op == OP_TMCLOVER) && !clov) return QPHIX_AMD_EINVAL;
  FaceSpinorGeom q = kNativeSpinor;
  if (lay && !face_qgeom(ctx, lay, q)) return QPHIX_AMD_EINVAL;
  // QPhiX arrays need their FT's alignment only (element accesses)
  const bool out_ok = lay ? (reinterpret_cast<uintptr_t>(out) % storage_size(prec)) == 0 : aligned16(out);
  if (!out_ok || !aligned16(buf) || !aligned16(gauge) || (clov && !aligned16(clov)))
    return QPHIX_AMD_EALIGN;
  const Geom &g = ctx->g;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return status(with_prec(prec, [&](auto *tag) -> hipError_t {
    using S = std::remove_pointer_t<decltype(tag)>;
    using R = typename Traits<S>::R;
    FaceArgs<R> a;
    a.cb = cb;
    a.mu = dim;
    a.side = fb;
    a.isign = isign;
    a.nface = face_count(g, dim);
    a.tbc = (ctx->tbc_mask >> dim) & 1;
    a.beta = (R)beta;
    a.ph_re = (R)ctx->tbc_re[dim];
    a.ph_im = (R)ctx->tbc_im[dim];
    a.tm_mu = (R)mu;
    a.tm_muinv = (R)muinv;
    a.one = R(1);
    S *o = static_cast<S *>(out);
    const S *b = static_cast<const S *>(buf), *u = static_cast<const S *>(gauge);
    const S *cl = static_cast<const S *>(clov);
    switch (op) {
      case OP_WILSON: return launch_finish<S, OP_WILSON>(recon, g, q, o, b, u, cl, a, st);
      case OP_CLOVER: return launch_finish<S, OP_CLOVER>(recon, g, q, o, b, u, cl, a, st);
      case OP_TM: return launch_finish<S, OP_TM>(recon, g, q, o, b, u, cl, a, st);
      default: return launch_finish<S, OP_TMCLOVER>(recon, g, q, o, b, u, cl, a, st);
    }
  }));
}

}  // namespace
}  // namespace qamd

using qamd::OP_CLOVER;
using qamd::OP_TM;
using qamd::OP_TMCLOVER;
using qamd::OP_WILSON;

extern "C" {

int qphix_amd_set_comm(qphix_amd_ctx *ctx, const int comm[4])
{
  if (!ctx || !comm) return QPHIX_AMD_EINVAL;
  int m = 0;
  for (int d = 0; d < 4; d++)
    if (comm[d]) m |= 1 << d;
  ctx->comm_mask = m;
  return QPHIX_AMD_OK;
}

long qphix_amd_face_sites(const qphix_amd_ctx *ctx, int dim)
{
  if (!ctx || dim < 0 || dim > 3) return 0;
  return qamd::face_count(ctx->g, dim);
}

}  // extern "C"

namespace qamd {
namespace {
// lay: null for a native spinor in, else a QPhiX array of that layout
int pack(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int prec, int isign, int cb, int dim, int fb,
         void *buf, const void *in, void *stream)
{
  if (!ctx || prec < QPHIX_AMD_F64 || prec > QPHIX_AMD_BF16) return QPHIX_AMD_EINVAL;
  if ((isign != 1 && isign != -1) || (cb != 0 && cb != 1) || dim < 0 || dim > 3 || fb < 0 || fb > 2)
    return QPHIX_AMD_EINVAL;
  if (!buf || !in || buf == in) return QPHIX_AMD_EINVAL;
  FaceSpinorGeom q = kNativeSpinor;
  if (lay && !face_qgeom(ctx, lay, q)) return QPHIX_AMD_EINVAL;
  const bool in_ok = lay ? (reinterpret_cast<uintptr_t>(in) % storage_size(prec)) == 0 : aligned16(in);
  if (!aligned16(buf) || !in_ok) return QPHIX_AMD_EALIGN;
  const Geom &g = ctx->g;
  const int nface = qamd::face_count(g, dim);
  const long nthreads = fb == 2 ? 2L * nface : nface;
  hipStream_t st = static_cast<hipStream_t>(stream);
  return status(with_prec(prec, [&](auto *tag) -> hipError_t {
    using S = std::remove_pointer_t<decltype(tag)>;
    hipLaunchKernelGGL(face_pack_k<S>, dim3(blocks_for(nthreads, kFaceBlock)), dim3(kFaceBlock), 0, st,
                       static_cast<S *>(buf), static_cast<const S *>(in), g, q, 1 - cb, dim, fb, isign, nface);
    return hipGetLastError();
  }));
}
}  // namespace
}  // namespace qamd

extern "C" {

int qphix_amd_face_pack(qphix_amd_ctx *ctx, int prec, int isign, int cb, int dim, int fb, void *buf,
                        const void *in, void *stream)
{
  return qamd::pack(ctx, nullptr, prec, isign, cb, dim, fb, buf, in, stream);
}

int qphix_amd_qphix_face_pack(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int isign, int cb, int dim,
                              int fb, void *buf, const void *in, void *stream)
{
  if (!lay) return QPHIX_AMD_EINVAL;
  return qamd::pack(ctx, lay, lay->ft, isign, cb, dim, fb, buf, in, stream);
}

int qphix_amd_face_finish(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb, int dim, int fb,
                          void *out, const void *buf, const void *gauge, double beta, void *stream)
{
  return qamd::finish(ctx, nullptr, OP_WILSON, prec, recon, isign, cb, dim, fb, out, buf, gauge, nullptr, beta,
                      0.0, 0.0, stream);
}

int qphix_amd_clov_face_finish(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb, int dim,
                               int fb, void *out, const void *buf, const void *gauge,
                               const void *invclov, double beta, void *stream)
{
  return qamd::finish(ctx, nullptr, OP_CLOVER, prec, recon, isign, cb, dim, fb, out, buf, gauge, invclov, beta,
                      0.0, 0.0, stream);
}

int qphix_amd_tm_face_finish(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb, int dim, int fb,
                             void *out, const void *buf, const void *gauge, double beta, double mu,
                             double muinv, void *stream)
{
  return qamd::finish(ctx, nullptr, OP_TM, prec, recon, isign, cb, dim, fb, out, buf, gauge, nullptr, beta, mu,
                      muinv, stream);
}

int qphix_amd_tm_clov_face_finish(qphix_amd_ctx *ctx, int prec, int recon, int isign, int cb, int dim,
                                  int fb, void *out, const void *buf, const void *gauge,
                                  const void *fclov, double beta, void *stream)
{
  return qamd::finish(ctx, nullptr, OP_TMCLOVER, prec, recon, isign, cb, dim, fb, out, buf, gauge, fclov, beta,
                      0.0, 0.0, stream);
}

// QPhiX arrays in place: out a FourSpinorBlock array of the layout lay
// (FT = lay->ft), buffers, links and clover blocks as in the native calls
int qphix_amd_qphix_face_finish(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon, int isign,
                                int cb, int dim, int fb, void *out, const void *buf, const void *gauge,
                                double beta, void *stream)
{
  if (!lay) return QPHIX_AMD_EINVAL;
  return qamd::finish(ctx, lay, OP_WILSON, lay->ft, recon, isign, cb, dim, fb, out, buf, gauge, nullptr, beta,
                      0.0, 0.0, stream);
}

int qphix_amd_qphix_clov_face_finish(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon, int isign,
                                     int cb, int dim, int fb, void *out, const void *buf, const void *gauge,
                                     const void *invclov, double beta, void *stream)
{
  if (!lay) return QPHIX_AMD_EINVAL;
  return qamd::finish(ctx, lay, OP_CLOVER, lay->ft, recon, isign, cb, dim, fb, out, buf, gauge, invclov, beta,
                      0.0, 0.0, stream);
}

int qphix_amd_qphix_tm_face_finish(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon, int isign,
                                   int cb, int dim, int fb, void *out, const void *buf, const void *gauge,
                                   double beta, double mu, double muinv, void *stream)
{
  if (!lay) return QPHIX_AMD_EINVAL;
  return qamd::finish(ctx, lay, OP_TM, lay->ft, recon, isign, cb, dim, fb, out, buf, gauge, nullptr, beta, mu,
                      muinv, stream);
}

int qphix_amd_qphix_tm_clov_face_finish(qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, int recon,
                                        int isign, int cb, int dim, int fb, void *out, const void *buf,
                                        const void *gauge, const void *fclov, double beta, void *stream)
{
  if (!lay) return QPHIX_AMD_EINVAL;
  return qamd::finish(ctx, lay, OP_TMCLOVER, lay->ft, recon, isign, cb, dim, fb, out, buf, gauge, fclov, beta,
                      0.0, 0.0, stream);
}

}  // extern "C"
