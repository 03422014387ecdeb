// Rotation-centre search scoring on the device (recon.py:498-513).
//
// For each trial image: the score is sum p ln p over np.histogram(img[img != 0],
// bins=128), with numpy's uniform-bin rules reproduced in float64:
//   range = (min, max) of the non-zero values (widened by +-0.5 when equal),
//   f = ((x - first) / (last - first)) * 128, i = int(f), i == 128 -> 127,
//   edges[i] = i * ((last - first) / 128) + first (edges[128] = last),
//   x < edges[i] -> i - 1;  x >= edges[i + 1] and i != 127 -> i + 1.
// One 1024-thread workgroup per image: min/max reduction, LDS histogram, score.
#include "mr_internal.h"

namespace mr {

namespace {

constexpr int CT = 1024;
constexpr int NBINS = 128;

__device__ inline double edge(int i, double first, double last, double step) {
#pragma clang fp contract(off)
    return (i == NBINS) ? last : static_cast<double>(i) * step + first;
}

__global__ __launch_bounds__(CT) void cor_score_kernel(const float* __restrict__ img,
                                                       int64_t npix, double* __restrict__ scores) {
#pragma clang fp contract(off)
    __shared__ float smin[CT], smax[CT];
    __shared__ unsigned hist[NBINS];
    const float* im = img + static_cast<int64_t>(blockIdx.x) * npix;
    float lo = INFINITY, hi = -INFINITY;
    for (int64_t i = threadIdx.x; i < npix; i += CT) {
        const float x = im[i];
        if (x != 0.0f) { lo = fminf(lo, x); hi = fmaxf(hi, x); }
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    if (threadIdx.x < NBINS) hist[threadIdx.x] = 0;
    __syncthreads();
    for (int w = CT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (!(smin[0] <= smax[0])) {           // no non-zero value: empty histogram
        if (threadIdx.x == 0) scores[blockIdx.x] = 0.0;
        return;
    }
    double first = smin[0], last = smax[0];
    if (first == last) { first -= 0.5; last += 0.5; }
    const double denom = last - first;
    const double step = denom / NBINS;
    for (int64_t i = threadIdx.x; i < npix; i += CT) {
        const float xf = im[i];
        if (xf == 0.0f) continue;
        const double x = xf;
        int b = static_cast<int>(((x - first) / denom) * NBINS);
        if (b == NBINS) b -= 1;
        if (x < edge(b, first, last, step)) b -= 1;
        if (b != NBINS - 1 && x >= edge(b + 1, first, last, step)) b += 1;
        atomicAdd(&hist[b], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double total = 0.0;
        for (int b = 0; b < NBINS; ++b) total += hist[b];
        double s = 0.0;
        const double denom_t = total > 1.0 ? total : 1.0;
        for (int b = 0; b < NBINS; ++b) {
            if (hist[b] == 0) continue;
            const double p = hist[b] / denom_t;
            s += p * log(p);
        }
        scores[blockIdx.x] = s;
    }
}

}  // namespace

int launch_cor_scores(Ctx* c, const float* d_img, int nimg, int64_t npix, double* d_scores,
                      hipStream_t s) {
    (void)c;
    if (nimg <= 0) return MR_OK;
    hipLaunchKernelGGL(cor_score_kernel, dim3(nimg), dim3(CT), 0, s, d_img, npix, d_scores);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr
