// pmh_mlp.h -- the learned sizer: MLP training/inference on the device and
// the sketch features it reads.
#pragma once

#include "pmh_internal.h"

namespace pmh {

constexpr int kMlpIn = PMH_MLP_IN;          // features (padded)
constexpr int kMlpHidden = PMH_MLP_HIDDEN;  // one wavefront lane per hidden unit
constexpr int kMlpParams = PMH_MLP_PARAMS;

struct Mlp;
Mlp *mlp_new(const float *init);
void mlp_free(Mlp *M);
void mlp_get_params(Mlp *M, float *out);
void mlp_forward_host(Mlp *M, const float *x, uint32_t n, float *y);
float mlp_grad_host(Mlp *M, const float *x, const float *y, uint32_t n, float lambda, float *grad);
void mlp_set_data(Mlp *M, const float *x, const float *y, uint32_t n);
float mlp_train(Mlp *M, uint32_t epochs, uint32_t batch, float lr, float lambda, uint32_t seed);

// The model "auto" consults after its sample sketch (nullptr uninstalls).
void sizer_install(const float *params);
bool sizer_installed();
// The installed parameters in device memory (nullptr: none).  The partitioned
// engine's sketch block runs one query's forward pass on them; its outputs
// are out[0] = log2(groups + 1) / 32, out[1] = logit(sort engine faster).
const float *sizer_device_params();
void sizer_predict_many(const float *feat, uint32_t n, float *out);  // n rows of features, 2n outputs

}  // namespace pmh
