// Job descriptors of a coalesced batch (device-side table entries).
#pragma once

#include <vector>

#include "common.hpp"

namespace kc {

struct BatchJob {
  uint64_t n;
  const uint64_t* succ;  // list successors / tree parents, local ids (device)
  const int64_t* w;      // list weights (device) or null
  uint64_t* root_out;    // local root ids (device)
  int64_t* rank_out;     // ranks / depths (device)
  uint64_t base;         // set by rank_batch_device: offset in its kind's id space
  uint32_t kind;         // kc_job_kind
  uint32_t pad;
};

// Ranks every job of the batch in one pass; all pointers are device pointers.
// tour_build: kc_tour_build for the trees of the batch.
void rank_batch_device(Context* ctx, const std::vector<BatchJob>& jobs, const kc_rank_cfg* cfg,
                       int tour_build, hipStream_t st);

}  // namespace kc
