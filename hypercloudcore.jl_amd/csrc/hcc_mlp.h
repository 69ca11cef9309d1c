// Internal declarations of the troubled-node indicator network (the C ABI is in include/hcc.h).
#pragma once
#include "hcc_internal.h"

namespace hcc {
constexpr int kMlpIn = 16, kMlpHidden = 64;
constexpr int kMlpParams = kMlpIn * kMlpHidden + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden + kMlpHidden + 1;
// Logits of n bf16 feature rows (16 per row) on the device, queued on the network's context stream.
void mlp_forward_device(const hcc_mlp* m, const uint16_t* X, int64_t n, float* z);
// Logits z[t] of the feature rows idx[t], t < *n_dev (a device-side count, at most n_max), queued likewise; with
// apply, the indicator's decision too: apply[idx[t]] = 1 where z[t] > thr, and the count of those added to
// *apply_count.
void mlp_forward_list(const hcc_mlp* m, const uint16_t* X, const int32_t* idx, const unsigned long long* n_dev,
                      int64_t n_max, float* z, int8_t* apply = nullptr, float thr = 0.0f,
                      unsigned long long* apply_count = nullptr);
// The context (and stream) the network's kernels run on.
hcc_ctx* mlp_context(const hcc_mlp* m);
}  // namespace hcc
