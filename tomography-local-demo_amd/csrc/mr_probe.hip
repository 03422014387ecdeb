// Gather microbenchmark: the rate at which one MI355X serves scattered 16-byte loads (the
// band-ring element load of the back-projection kernel) from its L2 caches, as a function
// of the number of distinct 128-byte lines the 64 lanes of one load instruction touch.
// Measurement only (no counterpart in the reference): it gives the L2 side of the
// back-projection roofline, the time the kernel's gathers would take if every one hit L2.
#include "mr_internal.h"

namespace mr {
namespace {

constexpr int kProbeThreads = 256;
constexpr int kProbeLoads = 8;          // independent loads in flight per thread

// Lane l of instruction i reads the 16-byte element (line_i(x(l)) * 8 + slot(l)) of a
// buffer of nlines 128-byte lines (a power of two, small enough to stay in L2): line_i(x)
// is a hash of (i, x), so consecutive instructions touch unrelated lines.  Group g = 1:
// x(l) = l mod L and slot(l) = (l / L) mod 8 (consecutive lanes on different lines, the
// 64 / L lanes of a line over its eight 16-byte elements); g > 1: x(l) = (l / g) mod L
// and slot(l) = l mod 8 (runs of g consecutive lanes on one line).  Either way the 64
// lanes touch L distinct lines.
__global__ __launch_bounds__(kProbeThreads) void l2_gather_kernel(const uint4* __restrict__ buf,
                                                                   int nlines, int lines_per_wave,
                                                                   int group, int iters,
                                                                   uint4* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * kProbeThreads + threadIdx.x) >> 6;
    const uint32_t mask = static_cast<uint32_t>(nlines - 1);
    const uint32_t x = static_cast<uint32_t>(group == 1 ? lane % lines_per_wave
                                                        : (lane / group) % lines_per_wave);
    const uint32_t slot = static_cast<uint32_t>(group == 1 ? (lane / lines_per_wave) & 7 : lane & 7);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint4*>(buf), 0, nlines * 128, 0x00020000);
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
    uint32_t seed = static_cast<uint32_t>(wave) * 0x9E3779B1u;
    for (int it = 0; it < iters; ++it) {
        uint4 v[kProbeLoads];
#pragma unroll
        for (int j = 0; j < kProbeLoads; ++j) {
            const uint32_t ins = seed + static_cast<uint32_t>(it * kProbeLoads + j) * 0x85EBCA6Bu;
            const uint32_t line = ((ins ^ (ins >> 13)) + x * 0xC2B2AE35u) & mask;
            const int off = static_cast<int>(((line << 3) + slot) << 4);
            v[j] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
        }
#pragma unroll
        for (int j = 0; j < kProbeLoads; ++j) {
            acc.x ^= v[j].x;
            acc.y ^= v[j].y;
            acc.z ^= v[j].z;
            acc.w ^= v[j].w;
        }
    }
    // keep the loads: one store per thread
    out[blockIdx.x * kProbeThreads + threadIdx.x] = acc;
}

}  // namespace
}  // namespace mr

using namespace mr;

struct mr_ctx : Ctx {};

extern "C" int mr_l2_gather_bandwidth(mr_ctx* c, int lines_per_wave, int lane_group,
                                      int64_t buffer_bytes, double* gbps, double* ms) {
    if (!c) return fail(MR_EVALUE, "null context");
    MR_CHECK(gbps != nullptr, MR_EVALUE, "null output");
    MR_CHECK(lines_per_wave >= 1 && lines_per_wave <= 64, MR_EVALUE, "lines per wave in 1..64");
    MR_CHECK((lane_group == 1 || lane_group == 2 || lane_group == 4 || lane_group == 8) &&
                 lines_per_wave * lane_group <= 64,
             MR_EVALUE, "lane group 1, 2, 4 or 8 with lines per wave x group <= 64");
    MR_CHECK(buffer_bytes >= 128 * 64 && buffer_bytes <= (int64_t{1} << 30) &&
                 (buffer_bytes & (buffer_bytes - 1)) == 0,
             MR_EVALUE, "buffer bytes: a power of two in 8 KiB..1 GiB");
    std::lock_guard<std::mutex> lk(c->mu);
    MR_HIP(hipSetDevice(c->device));
    const int nlines = static_cast<int>(buffer_bytes / 128);
    const int blocks = c->n_cus * 8, iters = 64;
    void* d_buf = nullptr;
    void* d_out = nullptr;
    MR_HIP(hipMalloc(&d_buf, static_cast<size_t>(buffer_bytes)));
    hipError_t e = hipMalloc(&d_out, sizeof(uint4) * static_cast<size_t>(blocks) * kProbeThreads);
    if (e != hipSuccess) {
        hipFree(d_buf);
        return fail(MR_ERUNTIME, std::string("probe allocation: ") + hipGetErrorString(e));
    }
    float best = 0.0f;
    e = hipMemsetAsync(d_buf, 0x5a, static_cast<size_t>(buffer_bytes), c->s_req);
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {     // first run warms the L2s
        hipEventRecord(c->ev0, c->s_req);
        hipLaunchKernelGGL(l2_gather_kernel, dim3(blocks), dim3(kProbeThreads), 0, c->s_req,
                           static_cast<const uint4*>(d_buf), nlines, lines_per_wave, lane_group, iters,
                           static_cast<uint4*>(d_out));
        hipEventRecord(c->ev1, c->s_req);
        e = hipStreamSynchronize(c->s_req);
        float t = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, c->ev0, c->ev1);
        if (rep > 0 && (best == 0.0f || t < best)) best = t;
    }
    hipFree(d_out);
    hipFree(d_buf);
    if (e != hipSuccess) return fail(MR_ERUNTIME, std::string("probe: ") + hipGetErrorString(e));
    const double bytes = 16.0 * kProbeThreads * blocks * static_cast<double>(iters) * kProbeLoads;
    *gbps = bytes / (static_cast<double>(best) * 1e-3) / 1e9;
    if (ms) *ms = best;
    return MR_OK;
}
