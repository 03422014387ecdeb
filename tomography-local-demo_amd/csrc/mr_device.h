// Device helpers shared by the back-projection kernels (mr_bp.hip, mr_corbp.hip).
#pragma once

#include "mr_internal.h"

namespace mr {
namespace {

// In-box mask of recon.py:253-265 with the points built exactly as numpy does:
// pts[p, s] = (mo + t_s * me1) + t_p * me2, t_x = (x - mhalf) * mstep,
// each product and sum rounded separately (no contraction).
__device__ bool in_box(const mr_task& T, int p, int s) {
#pragma clang fp contract(off)
    const double ts = (static_cast<double>(s) - T.mhalf) * T.mstep;
    const double tp = (static_cast<double>(p) - T.mhalf) * T.mstep;
    bool ok = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double a = ts * T.me1[d];
        const double b = tp * T.me2[d];
        const double x = (T.mo[d] + a) + b;
        ok = ok && (fabs(x) <= T.box[d]);
    }
    return ok;
}

// Workgroup IDs are dealt round-robin to the 8 XCDs (ID mod 8), each with its own L2.
// Remapping ID b to tile x * per + min(x, rem) + b / 8 (x = b mod 8) hands every XCD a
// contiguous run of the tile list; neighbouring tiles on a slice then share the
// detector cache lines at their borders inside one L2.  The map is a bijection on
// [0, nb) for any nb.
__device__ inline int xcd_block(int b, int nb) {
    const int x = b & 7, k = b >> 3, per = nb >> 3, rem = nb & 7;
    return x * per + (x < rem ? x : rem) + k;
}

// floor(x) as an integer in one instruction
__device__ inline int cvt_flr(float x) {
    int r;
    asm volatile("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

}  // namespace
}  // namespace mr
