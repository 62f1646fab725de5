// dslash_bf16.hip -- dslash family instantiated for __hip_bfloat16 storage.
#include "dslash_dispatch.h"

namespace qamd {
hipError_t run_dslash_bf16(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st)
{
  return run_dslash_typed<__hip_bfloat16>(ctx, c, st);
}
}  // namespace qamd
