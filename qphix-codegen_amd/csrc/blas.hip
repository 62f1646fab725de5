// blas.hip -- streaming vector kernels for the solvers.  Every vector is a
// native-layout spinor on one checkerboard (n = 24 * vcb reals); a thread
// handles whole sites, loading each operand through its own layout, so the
// storage types of the operands may differ (the native vector width depends
// on the storage type).
//
// Reductions are FP64 and deterministic: every work-group writes one
// partial sum, and a single work-group adds the partials in a fixed order.
// Solver scalars (alpha, beta, residual norms) stay in device memory, so the
// iteration is enqueued without host round trips.
#include <type_traits>

#include "blas.h"

namespace qamd {

namespace {

constexpr int kBlasBlock = 256;

// A site's 24 reals are handled in three chunks of 8 (a whole number of
// native vectors for every storage type), so that even an FP64 update of
// four operands keeps ~100 VGPRs and 4+ waves per SIMD.
constexpr int kChunk = 8;

template <int C, typename S, typename R>
__device__ __forceinline__ void ldc(const S *__restrict__ p, int vcb, int site, R v[kChunk])
{
  load_reals<S, spinor_w<S>(), C * kChunk, kChunk>(p, vcb, site, v);
}

template <int C, typename S, typename R>
__device__ __forceinline__ void stc(S *__restrict__ p, int vcb, int site, const R v[kChunk])
{
  store_reals<S, spinor_w<S>(), C * kChunk, kChunk>(p, vcb, site, v);
}

// f(integral_constant<int, c>) for the chunks c = 0, 1, 2
template <typename F>
__device__ __forceinline__ void for_chunks(F &&f)
{
  f(std::integral_constant<int, 0>{});
  f(std::integral_constant<int, 1>{});
  f(std::integral_constant<int, 2>{});
}

// v rounded to storage S and back, in registers: exactly the values a load
// of what stc() stores would return
template <typename S, typename R>
__device__ __forceinline__ void round_to(R v[kChunk])
{
#pragma unroll
  for (int k = 0; k < kChunk; k++) v[k] = (R)to_real(to_storage<S>(v[k]));
}

#define QAMD_SITE_LOOP(site, vcb) \
  for (int site = blockIdx.x * kBlasBlock + threadIdx.x; site < (vcb); site += gridDim.x * kBlasBlock)

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
__global__ __launch_bounds__(kBlasBlock) void dot_k(const S *__restrict__ a, const S *__restrict__ b, int vcb,
                                                    double *partial)
{
  using R = typename Traits<S>::R;
  double acc = 0.0;
  QAMD_SITE_LOOP(site, vcb)
  {
    R s = 0;
    for_chunks([&](auto c) {
      constexpr int C = decltype(c)::value;
      R x[kChunk], y[kChunk];
      ldc<C>(a, vcb, site, x);
      ldc<C>(b, vcb, site, y);
#pragma unroll
      for (int k = 0; k < kChunk; k++) s += x[k] * y[k];
    });
    acc += (double)s;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// y = a*x + b*y (a, b from the host), optional |y|^2
template <typename SX, typename SY>
__global__ __launch_bounds__(kBlasBlock) void axpby_k(const SX *__restrict__ x, SY *__restrict__ y, double a,
                                                      double b, int vcb, double *partial)
{
  using R = typename Traits<SY>::R;
  double acc = 0.0;
  QAMD_SITE_LOOP(site, vcb)
  {
    for_chunks([&](auto c) {
      constexpr int C = decltype(c)::value;
      double xv[kChunk], yv[kChunk];
      ldc<C>(x, vcb, site, xv);
      if (b != 0.0)
        ldc<C>(y, vcb, site, yv);
      else
#pragma unroll
        for (int k = 0; k < kChunk; k++) yv[k] = 0.0;
      R o[kChunk];
#pragma unroll
      for (int k = 0; k < kChunk; k++) o[k] = (R)(a * xv[k] + b * yv[k]);
      stc<C>(y, vcb, site, o);
      // norm of the stored (rounded) values
      round_to<SY>(o);
#pragma unroll
      for (int k = 0; k < kChunk; k++) acc += (double)o[k] * (double)o[k];
    });
  }
  if (partial) {
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
  }
}

// CG step where the operator stored Ap (lattices whose checkerboard is not
// a whole number of operator work-groups): r -= alpha Ap; x += alpha p;
// partial |r|^2.  alpha = sc->alpha.
template <typename S, typename SXS>
__global__ __launch_bounds__(kBlasBlock) void cg_update_k(S *__restrict__ r, SXS *__restrict__ x,
                                                          const S *__restrict__ p, const S *__restrict__ Ap,
                                                          const CGScalars *sc, int vcb, double *partial,
                                                          const int *stop)
{
  if (stop && *stop) return;
  using R = typename Traits<S>::R;
  using RX = typename Traits<SXS>::R;
  const R alpha = (R)sc->alpha;
  double acc = 0.0;
  QAMD_SITE_LOOP(site, vcb)
  {
    R s = 0;
    for_chunks([&](auto c) {
      constexpr int C = decltype(c)::value;
      R rv[kChunk], pv[kChunk], av[kChunk];
      RX xv[kChunk];
      ldc<C>(r, vcb, site, rv);
      ldc<C>(p, vcb, site, pv);
      ldc<C>(Ap, vcb, site, av);
      ldc<C>(x, vcb, site, xv);
#pragma unroll
      for (int k = 0; k < kChunk; k++) {
        rv[k] = rv[k] - alpha * av[k];
        xv[k] = xv[k] + (RX)(alpha * pv[k]);
      }
      stc<C>(r, vcb, site, rv);
      stc<C>(x, vcb, site, xv);
      // norm of the stored (rounded) residual
      round_to<S>(rv);
#pragma unroll
      for (int k = 0; k < kChunk; k++) s += rv[k] * rv[k];
    });
    acc += (double)s;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// p = r + beta p, beta = sc->beta
template <typename S>
__global__ __launch_bounds__(kBlasBlock) void cg_xpay_k(S *__restrict__ p, const S *__restrict__ r,
                                                        const CGScalars *sc, int vcb, const int *stop)
{
  if (stop && *stop) return;
  using R = typename Traits<S>::R;
  const R beta = (R)sc->beta;
  QAMD_SITE_LOOP(site, vcb)
  {
    for_chunks([&](auto c) {
      constexpr int C = decltype(c)::value;
      R rv[kChunk], pv[kChunk];
      ldc<C>(r, vcb, site, rv);
      ldc<C>(p, vcb, site, pv);
#pragma unroll
      for (int k = 0; k < kChunk; k++) pv[k] = rv[k] + beta * pv[k];
      stc<C>(p, vcb, site, pv);
    });
  }
}

// CG step after an operator launch that already updated r (and summed
// |r|^2): x += alpha p, then (P) p = r + beta p; alpha, beta = sc->alpha,
// sc->beta.  The arithmetic is that of cg_update_k (x) and cg_xpay_k (p).
template <typename S, typename SXS, bool P>
__global__ __launch_bounds__(kBlasBlock) void cg_xpay_x_k(S *__restrict__ p, SXS *__restrict__ x,
                                                          const S *__restrict__ r, const CGScalars *sc,
                                                          int vcb, const int *stop)
{
  if (stop && *stop) return;
  using R = typename Traits<S>::R;
  using RX = typename Traits<SXS>::R;
  const R alpha = (R)sc->alpha, beta = (R)sc->beta;
  QAMD_SITE_LOOP(site, vcb)
  {
    for_chunks([&](auto c) {
      constexpr int C = decltype(c)::value;
      R pv[kChunk], rv[kChunk];
      RX xv[kChunk];
      ldc<C>(p, vcb, site, pv);
      ldc<C>(x, vcb, site, xv);
      if constexpr (P) ldc<C>(r, vcb, site, rv);
#pragma unroll
      for (int k = 0; k < kChunk; k++) xv[k] = xv[k] + (RX)(alpha * pv[k]);
      stc<C>(x, vcb, site, xv);
      if constexpr (P) {
#pragma unroll
        for (int k = 0; k < kChunk; k++) pv[k] = rv[k] + beta * pv[k];
        stc<C>(p, vcb, site, pv);
      }
    });
  }
}

// A reliable update's fold in one pass: (PENDING) the last iteration's
// xs += alpha p, then x = a xs + b x into the FP64 solution (a = sigma,
// b = 1), then xs = 0.  The arithmetic is that of cg_xpay_x_k (xs) and
// axpby_k (x); xs has its arithmetic type as storage (FP32 or FP64), so the
// updated xs needs no rounding.
template <typename S, typename SXS, bool PENDING>
__global__ __launch_bounds__(kBlasBlock) void cg_fold_k(const S *__restrict__ p, SXS *__restrict__ xs,
                                                        double *__restrict__ x, const CGScalars *sc,
                                                        double a, double b, int vcb)
{
  using R = typename Traits<S>::R;
  using RX = typename Traits<SXS>::R;
  static_assert(std::is_same<RX, SXS>::value, "the increment is stored in its arithmetic type");
  const R alpha = (R)sc->alpha;
  QAMD_SITE_LOOP(site, vcb)
  {
    for_chunks([&](auto c) {
      constexpr int C = decltype(c)::value;
      RX xv[kChunk];
      ldc<C>(xs, vcb, site, xv);
      if constexpr (PENDING) {
        R pv[kChunk];
        ldc<C>(p, vcb, site, pv);
#pragma unroll
        for (int k = 0; k < kChunk; k++) xv[k] = xv[k] + (RX)(alpha * pv[k]);
      }
      double yv[kChunk], o[kChunk];
      ldc<C>(x, vcb, site, yv);
#pragma unroll
      for (int k = 0; k < kChunk; k++) o[k] = a * (double)xv[k] + b * yv[k];
      stc<C>(x, vcb, site, o);
      RX z[kChunk];
#pragma unroll
      for (int k = 0; k < kChunk; k++) z[k] = (RX)0;
      stc<C>(xs, vcb, site, z);
    });
  }
}

// A restart after a reliable update in one pass: rs = a r64 + b 0 (a =
// 1 / sigma, b = 0) in storage S, then p = rs + beta p with rs as stored.
// The arithmetic is that of axpby_k with b = 0 (rs) and cg_xpay_k (p).
template <typename S>
__global__ __launch_bounds__(kBlasBlock) void cg_restart_k(S *__restrict__ rs, S *__restrict__ p,
                                                           const double *__restrict__ r64,
                                                           const CGScalars *sc, double a, double b, int vcb)
{
  using R = typename Traits<S>::R;
  const R beta = (R)sc->beta;
  QAMD_SITE_LOOP(site, vcb)
  {
    for_chunks([&](auto c) {
      constexpr int C = decltype(c)::value;
      double xv[kChunk];
      ldc<C>(r64, vcb, site, xv);
      R o[kChunk];
#pragma unroll
      for (int k = 0; k < kChunk; k++) o[k] = (R)(a * xv[k] + b * 0.0);
      stc<C>(rs, vcb, site, o);
      round_to<S>(o);
      R pv[kChunk];
      ldc<C>(p, vcb, site, pv);
#pragma unroll
      for (int k = 0; k < kChunk; k++) pv[k] = o[k] + beta * pv[k];
      stc<C>(p, vcb, site, pv);
    });
  }
}

// First stage of a long sum: work-group w adds partial[w*2048 .. +2048)
// (strided over its threads, then the fixed work-group tree) into out[w].
constexpr int kReduceChunk = 8 * kBlasBlock;

__global__ __launch_bounds__(kBlasBlock) void reduce_partials_k(const double *__restrict__ partial, int n,
                                                                double *__restrict__ out)
{
  const int base = blockIdx.x * kReduceChunk;
  const int end = base + kReduceChunk < n ? base + kReduceChunk : n;
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < kReduceChunk / kBlasBlock; k++) {
    const int i = base + k * kBlasBlock + (int)threadIdx.x;
    if (i < end) v += partial[i];
  }
  const double t = block_sum(v);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Sum the partials in a fixed order and update the scalars:
//   kind 0: pAp  -> alpha = rr / pAp
//   kind 1: rr_new -> beta = rr_new / rr, rr = rr_new, one more iteration
//           done, and the reliable-update trigger (see CGScalars)
//   kind 2: plain sum into sc->tmp
__global__ __launch_bounds__(kBlasBlock) void finalize_k(const double *partial, int n, CGScalars *sc,
                                                         int kind, const int *stop)
{
  if (stop && *stop) return;
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
      sc->iters += 1;
      const double rest = sqrt(t) * sc->sigma;
      const bool reliable = sc->mixed ? (rest < sc->delta * sc->rmax || rest * rest < sc->target2)
                                      : rest * rest < sc->target2;
      if (!isfinite(t) || reliable)
        sc->stop = 1;
      else if (rest > sc->rmax)
        sc->rmax = rest;
    } else {
      sc->tmp = t;
    }
  }
}

// *sc = v (one thread; the values travel as the launch's argument)
__global__ void set_scalars_k(CGScalars *sc, CGScalars v) { *sc = v; }

template <typename S> const S *cp(const void *p) { return static_cast<const S *>(p); }
template <typename S> S *mp(void *p) { return static_cast<S *>(p); }

}  // namespace

int blas_grid(long nsites)
{
  long b = (nsites + kBlasBlock - 1) / kBlasBlock;
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
              int kind, hipStream_t st, const int *stop)
{
  const int ng = (int)(n / 24);  // sites
  const int grid = blas_grid(ng);
#define QAMD_DOT(T) \
  hipLaunchKernelGGL(dot_k<T>, dim3(grid), dim3(kBlasBlock), 0, st, cp<T>(a), cp<T>(b), ng, partial)
  QAMD_PREC_SWITCH(prec, QAMD_DOT)
#undef QAMD_DOT
  hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kBlasBlock), 0, st, partial, grid, sc, kind, stop);
}

void blas_finalize(const double *partial, long n, CGScalars *sc, int kind, double *scratch, hipStream_t st,
                   const int *stop)
{
  if (n > kReduceChunk) {  // one work-group per 2048 partials first
    const int g = (int)((n + kReduceChunk - 1) / kReduceChunk);
    hipLaunchKernelGGL(reduce_partials_k, dim3(g), dim3(kBlasBlock), 0, st, partial, (int)n, scratch);
    hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kBlasBlock), 0, st, scratch, g, sc, kind, stop);
    return;
  }
  hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kBlasBlock), 0, st, partial, (int)n, sc, kind, stop);
}

void blas_axpby(int prec_x, const void *x, int prec_y, void *y, double a, double b, long n,
                double *partial, CGScalars *sc, hipStream_t st)
{
  const int ng = (int)(n / 24);  // sites
  // one site per thread unless there are partial sums (at most kMaxPartials)
  const int grid = sc ? blas_grid(ng) : blocks_for(ng, kBlasBlock);
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
  if (sc) hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kBlasBlock), 0, st, partial, grid, sc, 2, nullptr);
}

void blas_cg_update(int prec, int prec_x, void *r, void *x, const void *p, const void *Ap, long n,
                    double *partial, CGScalars *sc, hipStream_t st, const int *stop)
{
  const int ng = (int)(n / 24);  // sites
  const int grid = blas_grid(ng);
#define QAMD_CGU(T)                                                                          \
  if (prec_x == QPHIX_AMD_F64)                                                               \
    hipLaunchKernelGGL((cg_update_k<T, double>), dim3(grid), dim3(kBlasBlock), 0, st, mp<T>(r), \
                       mp<double>(x), cp<T>(p), cp<T>(Ap), sc, ng, partial, stop);           \
  else                                                                                       \
    hipLaunchKernelGGL((cg_update_k<T, float>), dim3(grid), dim3(kBlasBlock), 0, st, mp<T>(r), \
                       mp<float>(x), cp<T>(p), cp<T>(Ap), sc, ng, partial, stop);
  QAMD_PREC_SWITCH(prec, QAMD_CGU)
#undef QAMD_CGU
  hipLaunchKernelGGL(finalize_k, dim3(1), dim3(kBlasBlock), 0, st, partial, grid, sc, 1, stop);
}

void blas_cg_xpay_x(int prec, int prec_x, void *p, void *x, const void *r, long n, CGScalars *sc,
                    bool update_p, hipStream_t st, const int *stop)
{
  const int ng = (int)(n / 24);  // sites
  const int grid = blocks_for(ng, kBlasBlock);  // no partials: one site per thread
#define QAMD_XPX_P(T, TX, P)                                                                    \
  hipLaunchKernelGGL((cg_xpay_x_k<T, TX, P>), dim3(grid), dim3(kBlasBlock), 0, st, mp<T>(p), \
                     mp<TX>(x), cp<T>(r), sc, ng, stop)
#define QAMD_XPX_X(T, TX)        \
  if (update_p)                  \
    QAMD_XPX_P(T, TX, true);     \
  else                           \
    QAMD_XPX_P(T, TX, false);
#define QAMD_XPX(T)                   \
  if (prec_x == QPHIX_AMD_F64) {      \
    QAMD_XPX_X(T, double)             \
  } else {                            \
    QAMD_XPX_X(T, float)              \
  }
  QAMD_PREC_SWITCH(prec, QAMD_XPX)
#undef QAMD_XPX
#undef QAMD_XPX_X
#undef QAMD_XPX_P
}

void blas_cg_fold(int prec, int prec_x, const void *p, void *xs, double *x, double sigma, long n,
                  CGScalars *sc, bool pending, hipStream_t st)
{
  const int ng = (int)(n / 24);  // sites
  const int grid = blocks_for(ng, kBlasBlock);
#define QAMD_FOLD_P(T, TX, P)                                                                          \
  hipLaunchKernelGGL((cg_fold_k<T, TX, P>), dim3(grid), dim3(kBlasBlock), 0, st, cp<T>(p), mp<TX>(xs), x, \
                     sc, sigma, 1.0, ng)
#define QAMD_FOLD_X(T, TX)       \
  if (pending)                   \
    QAMD_FOLD_P(T, TX, true);    \
  else                           \
    QAMD_FOLD_P(T, TX, false);
#define QAMD_FOLD(T)                  \
  if (prec_x == QPHIX_AMD_F64) {      \
    QAMD_FOLD_X(T, double)            \
  } else {                            \
    QAMD_FOLD_X(T, float)             \
  }
  QAMD_PREC_SWITCH(prec, QAMD_FOLD)
#undef QAMD_FOLD
#undef QAMD_FOLD_X
#undef QAMD_FOLD_P
}

void blas_cg_restart(int prec, void *rs, void *p, const double *r64, double scale, long n, CGScalars *sc,
                     hipStream_t st)
{
  const int ng = (int)(n / 24);  // sites
  const int grid = blocks_for(ng, kBlasBlock);
#define QAMD_RESTART(T)                                                                               \
  hipLaunchKernelGGL(cg_restart_k<T>, dim3(grid), dim3(kBlasBlock), 0, st, mp<T>(rs), mp<T>(p), r64, sc, \
                     scale, 0.0, ng)
  QAMD_PREC_SWITCH(prec, QAMD_RESTART)
#undef QAMD_RESTART
}

void blas_cg_xpay(int prec, void *p, const void *r, long n, CGScalars *sc, hipStream_t st, const int *stop)
{
  const int ng = (int)(n / 24);  // sites
  const int grid = blocks_for(ng, kBlasBlock);  // no partials: one site per thread
#define QAMD_XPAY(T) \
  hipLaunchKernelGGL(cg_xpay_k<T>, dim3(grid), dim3(kBlasBlock), 0, st, mp<T>(p), cp<T>(r), sc, ng, stop)
  QAMD_PREC_SWITCH(prec, QAMD_XPAY)
#undef QAMD_XPAY
}

void blas_set_scalars(CGScalars *sc, const CGScalars &v, hipStream_t st)
{
  hipLaunchKernelGGL(set_scalars_k, dim3(1), dim3(1), 0, st, sc, v);
}

}  // namespace qamd
