// Throughput variant: identical algorithm and step control; the compiler may
// fuse multiply-adds (results differ from the exact variant at round-off level).
#include "gm_common.h"
#pragma clang fp contract(fast)
#define GM_NS gm_fast
namespace gm_fast { static constexpr bool kFast = true; static constexpr bool kLearnedFused = false; }
#include "gm_device.inc"
#include "gm_kernels.inc"
#include "gm_persist.inc"
#include "gm_launch.inc"
