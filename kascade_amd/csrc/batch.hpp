// Job descriptors of a coalesced batch (device-side table entries).
#pragma once

#include <vector>

#include "common.hpp"

namespace kc {

// Ranks every job of the batch in one pass; all pointers are device pointers.
// tour_build: kc_tour_build for the trees of the batch.
void rank_batch_device(Context* ctx, const std::vector<BatchJob>& jobs, const kc_rank_cfg* cfg,
                       int tour_build, hipStream_t st);

}  // namespace kc
