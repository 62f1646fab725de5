// internal.h -- host-side glue between the C ABI and the kernel templates.
#pragma once

#include "common.h"

struct qphix_amd_ctx {
  qamd::Geom g;
  double tbound;
  int comm_mask;  // split dimensions (qphix_amd_set_comm)
  int tbc_mask;
  double tbc_re[4], tbc_im[4];
  int traversal;  // work-group order requested (qphix_amd_set_traversal), -1 auto
};

namespace qamd {

// One operator call, with every scalar still in double.
struct DslashCall {
  int prec, recon, op;
  bool achi;
  int isign, cb;
  void *out;
  const void *in, *chi, *gauge, *clov;
  double coeff_s, coeff_tf, coeff_tb, alpha, mu, muinv, rho;
  // optional: per work-group |out|^2 partials, vcb / 128 of them (only when
  // vcb is a multiple of 128)
  double *norm;
  // optional, with norm, for the clover achimbdpsi kernels with isign -1:
  // out holds a residual r and receives r - (*step) * result; with rin set
  // too (FP64 only), out receives rin - (*step) * result instead (the CG's
  // true residual b - M^dag M x)
  const double *step;
  const void *rin;
  // optional (CG iterations): when it points to a set flag the launch does
  // nothing
  const int *stop;
  // spl = 1: in, chi and out are QPhiX FourSpinorBlock arrays of the
  // geometry q (converted from a qphix_amd_qphix_layout); acc, when set, is
  // the per-site direction mask in their site order (block * soalen + lane);
  // spl = 2: the same read and written as vectors by lane group (FP32 and
  // FP16 with soalen >= 4 and Lx/2 a multiple of 4, FP64 with soalen >= 2 and
  // Lx/2 even; arrays aligned to the group's vector)
  int spl;
  int q_sh, q_nvecs, q_pxy, q_pxyz;
  const unsigned char *acc;
};

// FourSpinorBlock geometry of a QPhiX layout on a lattice, for the
// operator kernels that read and write those arrays in place: soalen =
// 1 << sh, block strides nvecs (y), pxy (z), pxyz (t) in blocks, and the
// number of site slots (padding included).
struct QSpinorGeom {
  int sh, nvecs, pxy, pxyz;
  long sites;
};
// false for a layout the lattice does not admit, a soalen that is not a
// power of two, or arrays of 2^31 elements or more
bool qphix_spinor_geometry(const qphix_amd_ctx *ctx, const qphix_amd_qphix_layout *lay, QSpinorGeom &out);

// The work-group traversal order the operator calls use on this geometry.
int dslash_traversal(const qphix_amd_ctx *ctx);

hipError_t run_dslash_f64(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st);
hipError_t run_dslash_f32(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st);
hipError_t run_dslash_f16(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st);
hipError_t run_dslash_bf16(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st);

// dispatch on c.prec
hipError_t run_dslash_any(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st);

hipError_t run_pack(const qphix_amd_ctx *ctx, int field, int prec, int recon, void *dst,
                    const double *src, int cb, hipStream_t st);
hipError_t run_unpack_spinor(const qphix_amd_ctx *ctx, int prec, double *dst, const void *src,
                             int cb, hipStream_t st);

inline int blocks_for(long n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace qamd
