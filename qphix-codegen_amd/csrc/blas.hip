// blas.hip -- streaming vector kernels for the solvers: every vector is a
// native-layout spinor on one checkerboard, viewed as a flat array of
// n = 24 * vcb reals in its storage type.  Each thread handles groups of 8
// reals (one 16-byte load for FP16/BF16, two for FP32, four for FP64).
//
// Reductions are FP64 and deterministic: every work-group writes one
// partial sum, and a single work-group adds the partials in a fixed order.
// Solver scalars (alpha, beta, residual norms) stay in device memory, so the
// iteration is enqueued without host round trips.
#include "blas.h"

namespace qamd {

namespace {

constexpr int kBlasBlock = 256;

// 8 consecutive reals starting at element 8*g, as 16-byte loads/stores
template <typename S, typename R>
__device__ __forceinline__ void ld8(const S *p, long g, R v[8])
{
  constexpr int W = 16 / (int)sizeof(S) < 8 ? 16 / (int)sizeof(S) : 8;
  const u32x4 *vp = reinterpret_cast<const u32x4 *>(p + g * 8);
#pragma unroll
  for (int c = 0; c < 8 / W; c++) {
    u32x4 x = vp[c];
    S tmp[W];
    __builtin_memcpy(tmp, &x, 16);
#pragma unroll
    for (int e = 0; e < W; e++) v[c * W + e] = (R)to_real(tmp[e]);
  }
}

template <typename S, typename R>
__device__ __forceinline__ void st8(S *p, long g, const R v[8])
{
  constexpr int W = 16 / (int)sizeof(S) < 8 ? 16 / (int)sizeof(S) : 8;
  u32x4 *vp = reinterpret_cast<u32x4 *>(p + g * 8);
#pragma unroll
  for (int c = 0; c < 8 / W; c++) {
    S tmp[W];
#pragma unroll
    for (int e = 0; e < W; e++) tmp[e] = to_storage<S>(v[c * W + e]);
    u32x4 x;
    __builtin_memcpy(&x, tmp, 16);
    vp[c] = x;
  }
}

// work-group sum of a double, result valid in thread 0
__device__ __forceinline__ double block_sum(double v)
{
  __shared__ double sm[kBlasBlock / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sm[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlasBlock / 64; w++) r += sm[w];
  return r;
}

template <typename S>
__global__ __launch_bounds__(kBlasBlock) void dot_k(const S *a, const S *b, long ngroups,
                                                    double *partial)
{
  using R = typename Traits<S>::R;
  double acc = 0.0;
  for (long g = (long)blockIdx.x * kBlasBlock + threadIdx.x; g < ngroups;
       g += (long)gridDim.x * kBlasBlock) {
    R x[8], y[8];
    ld8(a, g, x);
    ld8(b, g, y);
    R s = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) s += x[k] * y[k];
    acc += (double)s;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// y = a*x + b*y (a, b from the host), optional |y|^2
template <typename SX, typename SY>
__global__ __launch_bounds__(kBlasBlock) void axpby_k(const SX *x, SY *y, double a, double b,
                                                      long ngroups, double *partial)
{
  using R = typename Traits<SY>::R;
  double acc = 0.0;
  for (long g = (long)blockIdx.x * kBlasBlock + threadIdx.x; g < ngroups;
       g += (long)gridDim.x * kBlasBlock) {
    double xv[8], yv[8];
    ld8(x, g, xv);
    if (b != 0.0)
      ld8(y, g, yv);
    else
#pragma unroll
      for (int k = 0; k < 8; k++) yv[k] = 0.0;
    R o[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      o[k] = (R)(a * xv[k] + b * yv[k]);
      acc += (double)o[k] * (double)o[k];
    }
    st8(y, g, o);
  }
  if (partial) {
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// CG step: r -= alpha Ap; x += alpha p; partial |r|^2.  alpha = sc->alpha.
template <typename S, typename SXS>
__global__ __launch_bounds__(kBlasBlock) void cg_update_k(S *r, SXS *x, const S *p, const S *Ap,
                                                          const CGScalars *sc, long ngroups,
                                                          double *partial)
{
  using R = typename Traits<S>::R;
  const R alpha = (R)sc->alpha;
  double acc = 0.0;
  for (long g = (long)blockIdx.x * kBlasBlock + threadIdx.x; g < ngroups;
       g += (long)gridDim.x * kBlasBlock) {
    R rv[8], pv[8], av[8];
    ld8(r, g, rv);
    ld8(p, g, pv);
    ld8(Ap, g, av);
    typename Traits<SXS>::R xv[8];
    ld8(x, g, xv);
    R s = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      rv[k] = rv[k] - alpha * av[k];
      xv[k] = xv[k] + (typename Traits<SXS>::R)(alpha * pv[k]);
    }
    st8(r, g, rv);
    st8(x, g, xv);
    // norm of the stored (rounded) residual
    ld8(r, g, rv);
#pragma unroll
    for (int k = 0; k < 8; k++) s += rv[k] * rv[k];
    acc += (double)s;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// p = r + beta p, beta = sc->beta
template <typename S>
__global__ __launch_bounds__(kBlasBlock) void cg_xpay_k(S *p, const S *r, const CGScalars *sc,
                                                        long ngroups)
{
  using R = typename Traits<S>::R;
  const R beta = (R)sc->beta;
  for (long g = (long)blockIdx.x * kBlasBlock + threadIdx.x; g < ngroups;
       g += (long)gridDim.x * kBlasBlock) {
    R rv[8], pv[8];
    ld8(r, g, rv);
    ld8(p, g, pv);
#pragma unroll
    for (int k = 0; k < 8; k++) pv[k] = rv[k] + beta * pv[k];
    st8(p, g, pv);
  }
}

// Sum the partials in a fixed order and update the scalars:
//   kind 0: pAp  -> alpha = rr / pAp
//   kind 1: rr_new -> beta = rr_new / rr, rr = rr_new
//   kind 2: plain sum into sc->tmp
__global__ __launch_bounds__(kBlasBlock) void finalize_k(const double *partial, int n, CGScalars *sc,
                                                         int kind)
{
  double v = 0.0;
  for (int i = threadIdx.x; i < n; i += kBlasBlock) v += partial[i];
  const double t = block_sum(v);
  if (threadIdx.x == 0) {
    if (kind == 0) {
      sc->pAp = t;
      sc->alpha = sc->rr / t;
    } else if (kind == 1) {
      sc->beta = t / sc->rr;
      sc->rr_prev = sc->rr;
      sc->rr = t;
    } else {
      sc->tmp = t;
    }
  }
}

template <typename S> const S *cp(const void *p) { return static_cast<const S *>(p); }
template <typename S> S *mp(void *p) { return static_cast<S *>(p); }

}  // namespace

int blas_grid(long ngroups)
{
  long b = (ngroups + kBlasBlock - 1) / kBlasBlock;
  return (int)(b < kMaxPartials ? b : kMaxPartials);
}

#define QAMD_PREC_SWITCH(prec, MACRO) \
  switch (prec) {                     \
    case QPHIX_AMD_F64: MACRO(double); break; \
    case QPHIX_AMD_F32: MACRO(float); break;  \
    case QPHIX_AMD_F16: MACRO(__half); break; \
    default: MACRO(__hip_bfloat16); break;    \
  }

void blas_dot(int prec, const void *a, const void *b, long n, double *partial, CGScalars *sc,
              int kind, hipStream_t st)
{
  const long ng = n / 8;
  const int grid = blas_grid(ng);
#define QAMD_DOT(T) \
  hipLaunchKernelGGL(dot_k<T>, dim3(grid), dim3(kBlasBlock), 0, st, cp<T>(a), cp<T>(b), ng, partial)
  QAMD_PREC_SWITCH(prec, QAMD_DOT)
#undef QAMD_DOT
  hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kBlasBlock), 0, st, partial, grid, sc, kind);
}

void blas_axpby(int prec_x, const void *x, int prec_y, void *y, double a, double b, long n,
                double *partial, CGScalars *sc, hipStream_t st)
{
  const long ng = n / 8;
  const int grid = blas_grid(ng);
  double *pp = sc ? partial : nullptr;
#define QAMD_AXPBY_Y(TY)                                                                    \
  switch (prec_x) {                                                                         \
    case QPHIX_AMD_F64:                                                                     \
      hipLaunchKernelGGL((axpby_k<double, TY>), dim3(grid), dim3(kBlasBlock), 0, st,        \
                         cp<double>(x), mp<TY>(y), a, b, ng, pp);                            \
      break;                                                                                \
    case QPHIX_AMD_F32:                                                                     \
      hipLaunchKernelGGL((axpby_k<float, TY>), dim3(grid), dim3(kBlasBlock), 0, st,         \
                         cp<float>(x), mp<TY>(y), a, b, ng, pp);                             \
      break;                                                                                \
    case QPHIX_AMD_F16:                                                                     \
      hipLaunchKernelGGL((axpby_k<__half, TY>), dim3(grid), dim3(kBlasBlock), 0, st,        \
                         cp<__half>(x), mp<TY>(y), a, b, ng, pp);                            \
      break;                                                                                \
    default:                                                                                \
      hipLaunchKernelGGL((axpby_k<__hip_bfloat16, TY>), dim3(grid), dim3(kBlasBlock), 0, st, \
                         cp<__hip_bfloat16>(x), mp<TY>(y), a, b, ng, pp);                    \
      break;                                                                                \
  }
  QAMD_PREC_SWITCH(prec_y, QAMD_AXPBY_Y)
#undef QAMD_AXPBY_Y
  if (sc) hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kBlasBlock), 0, st, partial, grid, sc, 2);
}

void blas_cg_update(int prec, int prec_x, void *r, void *x, const void *p, const void *Ap, long n,
                    double *partial, CGScalars *sc, hipStream_t st)
{
  const long ng = n / 8;
  const int grid = blas_grid(ng);
#define QAMD_CGU(T)                                                                          \
  if (prec_x == QPHIX_AMD_F64)                                                               \
    hipLaunchKernelGGL((cg_update_k<T, double>), dim3(grid), dim3(kBlasBlock), 0, st, mp<T>(r), \
                       mp<double>(x), cp<T>(p), cp<T>(Ap), sc, ng, partial);                 \
  else                                                                                       \
    hipLaunchKernelGGL((cg_update_k<T, float>), dim3(grid), dim3(kBlasBlock), 0, st, mp<T>(r), \
                       mp<float>(x), cp<T>(p), cp<T>(Ap), sc, ng, partial);
  QAMD_PREC_SWITCH(prec, QAMD_CGU)
#undef QAMD_CGU
  hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kBlasBlock), 0, st, partial, grid, sc, 1);
}

void blas_cg_xpay(int prec, void *p, const void *r, long n, CGScalars *sc, hipStream_t st)
{
  const long ng = n / 8;
  const int grid = blas_grid(ng);
#define QAMD_XPAY(T) \
  hipLaunchKernelGGL(cg_xpay_k<T>, dim3(grid), dim3(kBlasBlock), 0, st, mp<T>(p), cp<T>(r), sc, ng)
  QAMD_PREC_SWITCH(prec, QAMD_XPAY)
#undef QAMD_XPAY
}

}  // namespace qamd
