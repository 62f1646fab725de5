// dslash_f64.hip -- dslash family instantiated for double storage.
#include "dslash_dispatch.h"

namespace qamd {
hipError_t run_dslash_f64(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st)
{
  return run_dslash_typed<double>(ctx, c, st);
}
}  // namespace qamd
