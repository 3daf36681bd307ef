// Learned opening criterion (placeholder until the MFMA kernels land).
#include "flk_internal.h"

extern "C" {
int32_t flk_mlp_set_weights(flk_ctx *, const float *, const float *) { return FLK_ERR_STATE; }
int32_t flk_mlp_get_weights(const flk_ctx *, float *, float *) { return FLK_ERR_STATE; }
int32_t flk_mlp_forward(flk_ctx *, const float *, int64_t, float *) { return FLK_ERR_STATE; }
int32_t flk_mlp_gradient(flk_ctx *, const float *, const float *, int64_t, float *, float *) { return FLK_ERR_STATE; }
int32_t flk_mlp_train(flk_ctx *, const float *, const float *, int64_t, int32_t, int32_t, float, uint64_t, float *) { return FLK_ERR_STATE; }
int32_t flk_mlp_samples(flk_ctx *, int32_t, int32_t, int64_t, uint64_t, float *, float *, int64_t *) { return FLK_ERR_STATE; }
}
