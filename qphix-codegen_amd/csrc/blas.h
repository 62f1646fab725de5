// blas.h -- vector kernels and device-resident scalars of the solvers.
#pragma once

#include "force_hop.h"
#include "internal.h"

namespace qamd {

// Scalars of a CG iteration, resident in device memory (FP64).
struct CGScalars {
  double rr;       // current |r|^2 (solver units)
  double rr_prev;  // previous |r|^2
  double pAp;      // <p, A p>
  double alpha, beta;
  double tmp;      // result of a plain reduction
  // The reliable-update trigger, evaluated on the device after each
  // iteration's |r|^2 (finalize, kind 1): with the iterated residual in true
  // units rest = sqrt(rr) sigma, the iterations stop (stop = 1) when
  // rest < delta rmax or rest^2 < target2 (mixed; plain CG: rest^2 <
  // target2) or rr is not finite, else rmax = max(rmax, rest).
  double sigma, rmax, delta, target2;
  double one;  // 1: the step of the true residual b - 1 * M^dag M x
  int mixed;
  int stop;   // set: the iteration kernels still queued return at once
  int iters;  // iterations completed since the host last set the scalars
};

// work-groups of a vector kernel (and partial sums of a reduction)
#ifndef QAMD_MAX_PARTIALS
#define QAMD_MAX_PARTIALS 1024
#endif
constexpr int kMaxPartials = QAMD_MAX_PARTIALS;

int blas_grid(long nsites);
// The kernels of a CG iteration take `stop` (&sc->stop there, else null):
// when it points to a set flag they return without doing anything.

// Re<a, b> over n = 24 * vcb reals (whole sites); kind: 0 -> alpha = rr / result,
// 2 -> sc->tmp = result.
void blas_dot(int prec, const void *a, const void *b, long n, double *partial, CGScalars *sc,
              int kind, hipStream_t st, const int *stop = nullptr);
// y = a x + b y (types may differ); with sc != null, sc->tmp = |y|^2.
void blas_axpby(int prec_x, const void *x, int prec_y, void *y, double a, double b, long n,
                double *partial, CGScalars *sc, hipStream_t st);
// Sum n partials in a fixed order into sc (kind as for blas_dot); above
// 2048 partials in two stages through scratch (n / 2048 doubles, at most
// kMaxPartials).
void blas_finalize(const double *partial, long n, CGScalars *sc, int kind, double *scratch, hipStream_t st,
                   const int *stop = nullptr);
// r -= alpha Ap, x += alpha p, then beta = |r|^2 / rr, rr = |r|^2.
void blas_cg_update(int prec, int prec_x, void *r, void *x, const void *p, const void *Ap, long n,
                    double *partial, CGScalars *sc, hipStream_t st, const int *stop = nullptr);
// p = r + beta p
void blas_cg_xpay(int prec, void *p, const void *r, long n, CGScalars *sc, hipStream_t st,
                  const int *stop = nullptr);
// x += alpha p, then (update_p) p = r + beta p: the vector half of a CG step
// whose r -= alpha Ap and |r|^2 the operator's last kernel did
// A reliable update's fold: (pending) xs += alpha p, then x += sigma xs
// (x FP64, xs the FP32/FP64 increment), then xs = 0, in one pass.
void blas_cg_fold(int prec, int prec_x, const void *p, void *xs, double *x, double sigma, long n,
                  CGScalars *sc, bool pending, hipStream_t st);
// The restart after a reliable update: rs = scale * r64 in storage prec,
// then p = rs + beta p, in one pass.
void blas_cg_restart(int prec, void *rs, void *p, const double *r64, double scale, long n, CGScalars *sc,
                     hipStream_t st);
// *sc = v in stream order, without a host synchronisation
void blas_set_scalars(CGScalars *sc, const CGScalars &v, hipStream_t st);
void blas_cg_xpay_x(int prec, int prec_x, void *p, void *x, const void *r, long n, CGScalars *sc,
                    bool update_p, hipStream_t st, const int *stop = nullptr);

hipError_t run_clover_from_links(const qphix_amd_ctx *ctx, double *clov, const double *U,
                                 double alpha, double csw, hipStream_t st);
hipError_t run_clover_blocks(const qphix_amd_ctx *ctx, double *out, const double *clov, int mode,
                             double mu, int sign, hipStream_t st);
// Reals per plane and site of the clover force's weights K (anti-Hermitian
// traceless 3x3: 8 independent reals), and per site (6 planes).
constexpr int kKReals = 8;
constexpr int kKSiteReals = 6 * kKReals;

// Accumulates scale * link gradient of Re sum_x w(x) chi^dag A psi (lex FP64
// fields; w = w_even / w_odd by parity); K: workspace of kKSiteReals * V
// doubles (component-major, K[(plane * kKReals + e) * V + site]).
hipError_t run_clover_force(const qphix_amd_ctx *ctx, double *force, double *K, const double *chi,
                            const double *psi, const double *U, double csw, double w_even, double w_odd,
                            double scale, int project, hipStream_t st);

// The same with chi and psi given as native FP64 spinors per checkerboard,
// and the hopping-term force of nhop terms (FP64 native spinors, unit
// coefficients, each weighted by its w) added to each link's G before the
// projection: the whole clover-operator force in one gather.
hipError_t run_clover_hop_force(const qphix_amd_ctx *ctx, double *force, double *K, const double *const chi_cb[2],
                                const double *const psi_cb[2], const double *U, double csw, double w_even,
                                double w_odd, double scale, int project, const HopTerm *hop, int nhop,
                                hipStream_t st);

// The hopping-term force of qphix_amd_dslash_force (FP64 native spinors,
// unit coefficients) of one or two terms, summed before the projection.
hipError_t run_hop_force(const qphix_amd_ctx *ctx, double *force_lex, const double *u_lex, const HopTerm *terms,
                         int nterm, double scale, int project, hipStream_t st);
// unit hop coefficients with the context's time boundary and twisted phases
HopCoef hop_unit_coef(const qphix_amd_ctx *ctx);

}  // namespace qamd
