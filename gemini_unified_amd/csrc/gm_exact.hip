// Bitwise-faithful variant: no fused multiply-add, operation order of the
// reference Fortran (same rules as the CPU oracle).
#include "gm_common.h"
#pragma clang fp contract(off)
#define GM_NS gm_exact
namespace gm_exact { static constexpr bool kFast = false; static constexpr bool kLearnedFused = false; }
#include "gm_device.inc"
#include "gm_kernels.inc"
#include "gm_persist.inc"
#include "gm_launch.inc"
