// Process-lifetime device scratch of the allocation-free interface-flux
// check (noalloc.hip).
#pragma once

#include <cstddef>

namespace mrfg {

// 32768 doubles (256 KiB) of reduction partials -- 8 per 256 interface
// faces, so one side of up to 1,048,576 faces -- plus the two results.
constexpr size_t kFluxScratchDoubles = 32768 + 2;
constexpr int kFluxScratchMaxFaces = (int)((kFluxScratchDoubles - 2) / 8 * 256);

// The device address of `device`'s scratch array (taken once per device with
// hipGetSymbolAddress; `device` must be current on the first call).
double *fluxScratch(int device);

}  // namespace mrfg
