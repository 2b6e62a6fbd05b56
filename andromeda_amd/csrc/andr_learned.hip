// Learned quadrature: placeholder until the network kernels land.
#include "andr_learned.h"

int andr_launch_stokes_system_learned(andr_ctx *ctx, const SysJob *, const int2 *, int, LearnedStats *) {
  return andr_fail(ctx, ANDR_ERR_STATE, "learned quadrature not built");
}
extern "C" int andr_quad_training_pairs(andr_ctx *ctx, const andr_mesh *, double, double, int64_t, float *, int32_t *,
                                        int64_t *) {
  return andr_fail(ctx, ANDR_ERR_STATE, "not built");
}
extern "C" int andr_mlp_train(andr_ctx *ctx, int64_t, const float *, const int32_t *, int32_t, int32_t, float, uint64_t,
                              float *, float *) {
  return andr_fail(ctx, ANDR_ERR_STATE, "not built");
}
extern "C" int andr_mlp_infer(andr_ctx *ctx, int64_t, const float *, const float *, float *) {
  return andr_fail(ctx, ANDR_ERR_STATE, "not built");
}
extern "C" int andr_mlp_loss_grad(andr_ctx *ctx, int64_t, const float *, const int32_t *, const float *, float *,
                                  float *) {
  return andr_fail(ctx, ANDR_ERR_STATE, "not built");
}
extern "C" int andr_set_quadrature_model(andr_ctx *ctx, const float *, float, double) {
  return andr_fail(ctx, ANDR_ERR_STATE, "not built");
}
