// Wave-level prefix sums for CDNA (64 lanes), shared by the build, the walks and the scans.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gz {

// Inclusive prefix sum over the wave with DPP: shifts by 1, 2, 4 and 8 lanes within each row of
// 16 (lanes shifted in from outside the row read 0), then lane 15 of rows 0 and 2 added into rows
// 1 and 3, and lane 31 into rows 2 and 3 (a row outside the row mask keeps its value). Six VALU
// adds and no LDS traffic. Every lane of the wave must be active.
__device__ __forceinline__ int wave_incl_scan_dpp(int v) {
    int x = v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);  // row_shr:1
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);  // row_shr:2
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);  // row_shr:4
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);  // row_shr:8
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false); // row_bcast:15
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false); // row_bcast:31
    return x;
}

__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v) {
    return (uint32_t)wave_incl_scan_dpp((int)v); // two's complement: the same bits
}

// Exclusive prefix sum and the wave's total (a scalar: lane 63's inclusive sum).
__device__ __forceinline__ int wave_excl_scan(int v, int& total) {
    const int x = wave_incl_scan_dpp(v);
    total = __builtin_amdgcn_readlane(x, 63);
    return x - v;
}

} // namespace gz
