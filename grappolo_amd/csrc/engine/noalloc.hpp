// noalloc.hpp -- entry points that allocate no memory (see noalloc.hip).
#pragma once

#include "common.hpp"

namespace glv {

// Scratch memory of the no-allocation entry points on each GPU: 64 KiB of
// flags and partial sums, then 16 bytes per vertex (degrees and community
// degrees).
constexpr size_t kNoAllocScratchBytes = (size_t(64) << 20) + (size_t(64) << 10);
// largest vertex count the scratch serves (4,194,304)
i64 noalloc_max_nv();
// modularity of labels C on g (validated as glv_modularity_dev validates);
// calls that overlap on one GPU share its scratch and race
double modularity_noalloc(const DevGraph &g, const i64 *C, hipStream_t s);

}  // namespace glv
