// dslash_f16.hip -- dslash family instantiated for __half storage.
#include "dslash_dispatch.h"

namespace qamd {
hipError_t run_dslash_f16(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st)
{
  return run_dslash_typed<__half>(ctx, c, st);
}
}  // namespace qamd
