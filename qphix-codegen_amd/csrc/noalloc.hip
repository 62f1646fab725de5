// noalloc.hip -- qphix_amd_clover_force_noalloc: the clover-term link
// derivative with its workspace in a static device array instead of a
// per-call hipMalloc, for callers that cannot allocate device memory (code
// under HIP graph capture, tight HMC loops, harnesses that forbid it).
//
// The workspace is the field of clover weights K, 6 planes x 8 reals per
// site (kKSiteReals = 48; component-major, K[(plane * 8 + e) * V + site]),
// written in full by the first kernel of the call before the second reads
// it, so nothing in it needs clearing between calls.  The array is QPHIX_AMD_NOALLOC_
// SCRATCH_BYTES long on every GPU the library runs on and is held until the
// process ends.  Every call on one GPU uses that GPU's array: calls that
// overlap on one GPU (two streams, two threads) race on it.
#include <mutex>

#include "blas.h"

static __device__ __align__(256) double scratch[QPHIX_AMD_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];  // device address of `scratch`, per device

// The device address of the current device's `scratch`, looked up once per
// device (with that device current) on the first call there.  The array's
// name in host code is a host-side stand-in, never a device address.
hipError_t scratch_address(void **out)
{
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace

extern "C" int qphix_amd_clover_force_noalloc(qphix_amd_ctx *ctx, double *force_lex, const double *chi_lex,
                                              const double *psi_lex, const double *u_lex, double csw,
                                              double scale_even, double scale_odd, int project, void *stream)
{
  if (!ctx || !force_lex || !chi_lex || !psi_lex || !u_lex) return QPHIX_AMD_EINVAL;
  if (ctx->comm_mask) return QPHIX_AMD_EINVAL;  // needs the whole lattice
  const size_t need = (size_t)2 * ctx->g.vcb * qamd::kKSiteReals * sizeof(double);
  if (need > QPHIX_AMD_NOALLOC_SCRATCH_BYTES) return QPHIX_AMD_ESCRATCH;
  void *K = nullptr;
  if (scratch_address(&K) != hipSuccess) return QPHIX_AMD_EHIP;
  hipError_t e = qamd::run_clover_force(ctx, force_lex, static_cast<double *>(K), chi_lex, psi_lex, u_lex,
                                        csw, scale_even, scale_odd, 1.0, project, (hipStream_t)stream);
  return e == hipSuccess ? QPHIX_AMD_OK : QPHIX_AMD_EHIP;
}
