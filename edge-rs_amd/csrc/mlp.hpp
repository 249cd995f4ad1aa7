// Shapes of the warm-start MLP and its device entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

namespace nh {

constexpr int kMlpIn = 12;
constexpr int kMlpHidden = 64;
constexpr int kMlpOut = 2;
constexpr int kMlpBase = 10;  // features 10, 11: stage one's (log sigma^2, log phi)
constexpr size_t kMlpNormFeat = 2 * kMlpIn;  // feature mean, sd
constexpr size_t kMlpNet = (size_t)kMlpHidden * kMlpIn + kMlpHidden +
                           2 * ((size_t)kMlpHidden * kMlpHidden + kMlpHidden) +
                           (size_t)kMlpOut * kMlpHidden + kMlpOut;
constexpr size_t kMlpParams = kMlpNormFeat + kMlpNet + 2 * kMlpOut;  // + target mean, sd

// out[n][2] = (log sigma^2, log phi) predicted from feat[n][12] (BF16 MFMA).
hipError_t launch_mlp_predict(const float* params, const float* feat, int n, float* out,
                              hipStream_t st);
// Feature and target means / sds of a data set into the parameter blob.
void mlp_standardisation(const float* feat, const float* tgt, size_t n, float* params);

}  // namespace nh
