// Volume statistics on the device, for the threshold that follows a volume
// reconstruction (model.py:18-29 otsu_threshold): exact order statistics of the
// non-zero voxels (np.percentile) and the histogram of the clipped non-zero voxels
// (np.histogram of np.clip), both over the whole volume in HBM.
//
// Order statistics: radix select on monotone 32-bit keys of the float32 values
// (negative values: all bits inverted; others: sign bit set), one byte per pass from
// the top; every pass histograms, for up to 8 target ranks at once, the next byte of
// the keys whose higher bytes equal the target's prefix found so far.
//
// Histogram: numpy's uniform-bin rule in float64 on the value clipped to [lo, hi]:
//   f = ((x - lo) / (hi - lo)) * bins, i = int(f), i == bins -> bins - 1,
//   x < edges[i] -> i - 1;  x >= edges[i + 1] and i != bins - 1 -> i + 1,
// with the caller's edges (np.linspace(lo, hi, bins + 1)).
#include "mr_internal.h"

namespace mr {

namespace {

constexpr int VT = 256;
constexpr int MAX_TARGETS = 8;

__device__ inline uint32_t float_key(float x) {
    const uint32_t u = __float_as_uint(x);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__global__ __launch_bounds__(VT) void radix_hist_kernel(const float* __restrict__ v, int64_t n,
                                                        const uint32_t* __restrict__ prefix,
                                                        int ntarget, int shift,
                                                        unsigned long long* __restrict__ hist) {
    __shared__ unsigned h[MAX_TARGETS][256];
    for (int i = threadIdx.x; i < MAX_TARGETS * 256; i += VT) (&h[0][0])[i] = 0;
    __syncthreads();
    uint32_t pre[MAX_TARGETS];
    for (int t = 0; t < MAX_TARGETS; ++t) pre[t] = t < ntarget ? prefix[t] : 0u;
    const int hs = shift + 8;                       // bits above the current byte
    const int64_t stride = static_cast<int64_t>(gridDim.x) * VT;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) {
        const float x = v[i];
        if (x == 0.0f) continue;                    // vol[vol != 0]
        const uint32_t k = float_key(x);
        const unsigned b = (k >> shift) & 255u;
        if (hs >= 32) {                             // first pass: one histogram
            atomicAdd(&h[0][b], 1u);
            continue;
        }
        for (int t = 0; t < ntarget; ++t)
            if ((k >> hs) == (pre[t] >> hs)) atomicAdd(&h[t][b], 1u);
    }
    __syncthreads();
    const int rows = hs >= 32 ? 1 : ntarget;
    for (int i = threadIdx.x; i < rows * 256; i += VT) {
        const unsigned c = (&h[0][0])[i];
        if (c) atomicAdd(&hist[i], static_cast<unsigned long long>(c));
    }
}

__global__ __launch_bounds__(VT) void clip_hist_kernel(const float* __restrict__ v, int64_t n,
                                                       double lo, double hi, int bins,
                                                       const double* __restrict__ edges,
                                                       unsigned long long* __restrict__ hist) {
#pragma clang fp contract(off)
    extern __shared__ unsigned hsh[];
    for (int i = threadIdx.x; i < bins; i += VT) hsh[i] = 0;
    __syncthreads();
    const double denom = hi - lo;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * VT;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * VT + threadIdx.x; i < n; i += stride) {
        const float xf = v[i];
        if (xf == 0.0f) continue;
        double x = xf;
        x = x < lo ? lo : (x > hi ? hi : x);        // np.clip (NaN stays NaN)
        if (!(x >= lo && x <= hi)) continue;        // np.histogram drops NaN
        int b = static_cast<int>(((x - lo) / denom) * bins);
        if (b >= bins) b = bins - 1;
        if (b < 0) b = 0;
        if (x < edges[b]) b -= 1;
        if (b != bins - 1 && x >= edges[b + 1]) b += 1;
        atomicAdd(&hsh[b], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < bins; i += VT)
        if (hsh[i]) atomicAdd(&hist[i], static_cast<unsigned long long>(hsh[i]));
}

int grid_for(int64_t n) {
    const int64_t g = (n + VT - 1) / VT;
    return static_cast<int>(g < 4096 ? (g > 0 ? g : 1) : 4096);
}

}  // namespace

int launch_radix_hist(const float* d_v, int64_t n, const uint32_t* d_prefix, int ntarget,
                      int shift, unsigned long long* d_hist, hipStream_t s) {
    if (ntarget < 1 || ntarget > MAX_TARGETS) return fail(MR_EVALUE, "1..8 target ranks");
    hipLaunchKernelGGL(radix_hist_kernel, dim3(grid_for(n)), dim3(VT), 0, s, d_v, n, d_prefix,
                       ntarget, shift, d_hist);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int launch_clip_hist(const float* d_v, int64_t n, double lo, double hi, int bins,
                     const double* d_edges, unsigned long long* d_hist, hipStream_t s) {
    hipLaunchKernelGGL(clip_hist_kernel, dim3(grid_for(n)), dim3(VT),
                       sizeof(unsigned) * static_cast<size_t>(bins), s, d_v, n, lo, hi, bins,
                       d_edges, d_hist);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr
