// dslash_f32.hip -- dslash family instantiated for float storage.
#include "dslash_dispatch.h"

namespace qamd {
hipError_t run_dslash_f32(const qphix_amd_ctx *ctx, const DslashCall &c, hipStream_t st)
{
  return run_dslash_typed<float>(ctx, c, st);
}
}  // namespace qamd
