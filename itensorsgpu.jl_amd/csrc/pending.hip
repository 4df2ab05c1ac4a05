// Entry points declared in itg.h whose kernels live in later translation
// units; each returns ITG_EUNSUPPORTED until implemented.
#include "itg_internal.h"

extern "C" int64_t itg_mlp_param_count(int i, int h, int o) {
  return (int64_t)h * i + h + (int64_t)h * h + h + (int64_t)o * h + o;
}
extern "C" int itg_mlp_forward(itg_ctx* ctx, int, int, int, int, const float*, const float*,
                               float*, float*) {
  return itg_fail(ctx, ITG_EUNSUPPORTED, "itg_mlp_forward not built");
}
extern "C" int itg_mlp_backward(itg_ctx* ctx, int, int, int, int, const float*, const float*,
                                const float*, const float*, const float*, float*, double*) {
  return itg_fail(ctx, ITG_EUNSUPPORTED, "itg_mlp_backward not built");
}
extern "C" int itg_mlp_adam(itg_ctx* ctx, int64_t, float*, const float*, float*, float*, int,
                            float, float, float, float) {
  return itg_fail(ctx, ITG_EUNSUPPORTED, "itg_mlp_adam not built");
}
