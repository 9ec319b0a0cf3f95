// H_eff matvec without device allocations (scratch array per device).
#pragma once
#include <cstdint>

#include "common.hpp"

namespace cam {

struct Heff;

// size of the per-device scratch array that holds the matvec intermediates
constexpr int64_t kNoAllocScratchBytes = int64_t(256) << 20;   // 256 MiB

// takes the scratch array's address on `device` (which must be current);
// idempotent per device
void noalloc_setup(int device);
// FP64 matvec on device block-layout vectors with the intermediates in the
// scratch array: no hipMalloc / hipFree and no allocating library call
void heff_apply_noalloc(Heff& H, const cplx* in, cplx* out);

}  // namespace cam
