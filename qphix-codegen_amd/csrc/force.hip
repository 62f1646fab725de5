// force.hip -- gradient of L = Re<chi, D psi> with respect to the gauge
// links (the outer-product fermion-force stencil; the per-link algebra is in
// force_hop.h).  No counterpart in the reference, whose kernels only apply D;
// D is the hop of the reference's body (codegen/dslash_common.cc:1729-1826)
// with chi on checkerboard cb and psi on 1-cb.  Output, accumulated with a
// scale into the FP64 lexicographic [V][4][3][3][2] link field:
//   project = 1: F = TA(U G), the traceless anti-Hermitian su(3) force:
//                d/de L(exp(i e T) U) = Re tr(i T F) for Hermitian T;
//   project = 0: G^dagger, the gradient with respect to the real and
//                imaginary parts of the link entries.
#include "blas.h"
#include "force_hop.h"

namespace qamd {

namespace {

constexpr int kForceBlock = 128;

struct ForceParams {
  double *force;
  const double *U;
  Geom g;
  long V;  // sites of the lattice
  HopCoef k;
  HopTerm term[2];
  int nterm;
  double scale;
  int project;
};

// One thread per link U_mu(y), the four links of a site in adjacent lanes:
// a work-group's kForceSites sites are consecutive, so its link and force
// records ([4][18] reals per site, link j of the group at reals
// [18 j, 18 j + 18)) form one contiguous run.  That run moves between
// memory and LDS in coalesced copies (each load or store instruction of a
// wave covers 512 or 1024 consecutive bytes); a lane alone would read its
// 144-byte record in pieces 144 bytes apart from its neighbours'.
constexpr int kForceSites = kForceBlock / 4;
constexpr int kForceRec = 18;  // reals of one link record

using f64x2 = double __attribute__((ext_vector_type(2)));

// Coalesced copy of the group's first n reals (n = 18 x valid links) from
// global memory into LDS; 16-byte pieces when the run is 16-byte aligned.
__device__ __forceinline__ void force_run_load(double *lds, const double *src, int n)
{
  if (((uintptr_t)src & 15) == 0 && (n & 1) == 0) {
    const f64x2 *s2 = reinterpret_cast<const f64x2 *>(src);
    f64x2 *l2 = reinterpret_cast<f64x2 *>(lds);
    for (int i = threadIdx.x; i < n / 2; i += kForceBlock) l2[i] = s2[i];
  } else {
    for (int i = threadIdx.x; i < n; i += kForceBlock) lds[i] = src[i];
  }
}

// dst[i] += scale * lds[i] for the group's first n reals, coalesced
__device__ __forceinline__ void force_run_add(double *dst, const double *lds, int n, double scale)
{
  if (((uintptr_t)dst & 15) == 0 && (n & 1) == 0) {
    f64x2 *d2 = reinterpret_cast<f64x2 *>(dst);
    const f64x2 *l2 = reinterpret_cast<const f64x2 *>(lds);
    for (int i = threadIdx.x; i < n / 2; i += kForceBlock) {
      f64x2 o = d2[i];
      const f64x2 v = l2[i];
      o.x += scale * v.x;
      o.y += scale * v.y;
      d2[i] = o;
    }
  } else {
    for (int i = threadIdx.x; i < n; i += kForceBlock) dst[i] += scale * lds[i];
  }
}

template <typename S>
__global__ __launch_bounds__(kForceBlock) void force_k(ForceParams p)
{
  __shared__ __align__(16) double run[kForceBlock * kForceRec];
  // the lattice swept t-slice by t-slice by all XCDs (neighbour reuse in cache)
  int unused;
  const int mu = threadIdx.x & 3;
  const long y = sweep_site(p.g, 1, unused, kForceSites, threadIdx.x >> 2);
  // the valid sites of a group are a prefix of it
  const int nlinks = __syncthreads_count(y >= 0);
  if (nlinks == 0) return;  // uniform over the group
  const long y0 = sweep_site(p.g, 1, unused, kForceSites, 0);
  int c[4];
  long r = y < 0 ? 0 : y;  // lanes past the end compute nothing
  c[0] = (int)(r % p.g.Lx);
  r /= p.g.Lx;
  c[1] = (int)(r % p.g.Ly);
  r /= p.g.Ly;
  c[2] = (int)(r % p.g.Lz);
  c[3] = (int)(r / p.g.Lz);
  const long rec0 = y0 * 4 * kForceRec;
  const int nreal = nlinks * kForceRec;
  if (p.project) force_run_load(run, p.U + rec0, nreal);
  const int cbi_y = (((c[3] * p.g.Lz + c[2]) * p.g.Ly + c[1]) * p.g.Lxh) + (c[0] >> 1);
  cx<double> G[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++) G[i][j] = {0.0, 0.0};
  if (y >= 0) {
    hop_g_add<S>(G, p.term[0], p.k, p.g, c, cbi_y, mu);
    if (p.nterm > 1) hop_g_add<S>(G, p.term[1], p.k, p.g, c, cbi_y, mu);
  }
  // this lane's record in the LDS run (the link on entry, the force
  // contribution on exit)
  double *rec = run + threadIdx.x * kForceRec;
  double f[kForceRec];
  if (!p.project) {
    // G^dagger
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        const cx<double> v = conj(G[j][i]);
        f[(i * 3 + j) * 2] = v.re;
        f[(i * 3 + j) * 2 + 1] = v.im;
      }
  } else {
    // W = U G, F = (W - W^+)/2 - tr(W - W^+)/6
    __syncthreads();  // the link run is in LDS
    cx<double> W[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
      cx<double> u[3];
#pragma unroll
      for (int k = 0; k < 3; k++) u[k] = {rec[(i * 3 + k) * 2], rec[(i * 3 + k) * 2 + 1]};
#pragma unroll
      for (int j = 0; j < 3; j++) W[i][j] = u[0] * G[0][j] + u[1] * G[1][j] + u[2] * G[2][j];
    }
    const double trim = (W[0][0].im + W[1][1].im + W[2][2].im) / 3.0;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) {
        f[(i * 3 + j) * 2] = 0.5 * (W[i][j].re - W[j][i].re);
        f[(i * 3 + j) * 2 + 1] = 0.5 * (W[i][j].im + W[j][i].im) - (i == j ? trim : 0.0);
      }
    __syncthreads();  // every lane has read its link
  }
#pragma unroll
  for (int e = 0; e < kForceRec; e++) rec[e] = f[e];
  __syncthreads();
  force_run_add(p.force + rec0, run, nreal, p.scale);
}

template <typename S>
hipError_t launch_force(const ForceParams &p, hipStream_t st)
{
  hipLaunchKernelGGL((force_k<S>), dim3(sweep_grid(p.g, 1, kForceSites)), dim3(kForceBlock), 0, st, p);
  return hipGetLastError();
}

HopCoef hop_coef(const qphix_amd_ctx *ctx, double coeff_s, double coeff_t_f, double coeff_t_b)
{
  HopCoef k{};
  k.coeff_s = coeff_s;
  k.coeff_tf = coeff_t_f;
  k.coeff_tb = coeff_t_b;
  k.tbound = ctx->tbound;
  k.tbc_mask = ctx->tbc_mask;
  for (int m = 0; m < 4; m++) {
    k.tbc_re[m] = ctx->tbc_re[m];
    k.tbc_im[m] = ctx->tbc_im[m];
  }
  return k;
}

ForceParams force_params(const qphix_amd_ctx *ctx, double *force, const double *u, double scale, int project)
{
  ForceParams p{};
  p.force = force;
  p.U = u;
  p.g = ctx->g;
  p.V = 2L * ctx->g.vcb;
  p.scale = scale;
  p.project = project ? 1 : 0;
  return p;
}

}  // namespace

hipError_t run_hop_force(const qphix_amd_ctx *ctx, double *force_lex, const double *u_lex, const HopTerm *terms,
                         int nterm, double scale, int project, hipStream_t st)
{
  if (nterm < 1 || nterm > 2) return hipErrorInvalidValue;
  ForceParams p = force_params(ctx, force_lex, u_lex, scale, project);
  p.k = hop_coef(ctx, 1.0, 1.0, 1.0);
  for (int i = 0; i < nterm; i++) p.term[i] = terms[i];
  p.nterm = nterm;
  return launch_force<double>(p, st);
}

HopCoef hop_unit_coef(const qphix_amd_ctx *ctx) { return hop_coef(ctx, 1.0, 1.0, 1.0); }

}  // namespace qamd

using namespace qamd;

extern "C" int qphix_amd_dslash_force(qphix_amd_ctx *ctx, int prec, int isign, int cb,
                                      double *force_lex, const void *chi, const void *psi,
                                      const double *u_lex, double coeff_s, double coeff_t_f,
                                      double coeff_t_b, double scale, int project, void *stream)
{
  if (!ctx || !force_lex || !chi || !psi || !u_lex) return QPHIX_AMD_EINVAL;
  if (ctx->comm_mask) return QPHIX_AMD_EINVAL;  // needs the whole lattice
  if ((isign != 1 && isign != -1) || (cb != 0 && cb != 1)) return QPHIX_AMD_EINVAL;
  if (prec != QPHIX_AMD_F64 && prec != QPHIX_AMD_F32) return QPHIX_AMD_EINVAL;
  ForceParams p = force_params(ctx, force_lex, u_lex, scale, project);
  p.k = hop_coef(ctx, coeff_s, coeff_t_f, coeff_t_b);
  p.term[0] = HopTerm{chi, psi, cb, isign, 1.0};
  p.nterm = 1;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = prec == QPHIX_AMD_F64 ? launch_force<double>(p, st) : launch_force<float>(p, st);
  return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP;
}
