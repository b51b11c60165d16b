// ostc_select.hpp -- small selection helper shared by the air and mixture
// device code.
#pragma once
#include <hip/hip_runtime.h>

namespace ostc {

// a[i] for i = 0..2 by selects (i may be a run-time value; no private array)
__device__ __forceinline__ double pick3(const double* a, int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

}  // namespace ostc
