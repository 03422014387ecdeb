// Rotation-centre search scoring on the device (recon.py:498-513).
//
// For each trial image: the score is sum p ln p over np.histogram(img[img != 0],
// bins=128), with numpy's uniform-bin rules reproduced in float64:
//   range = (min, max) of the non-zero values (widened by +-0.5 when equal),
//   f = ((x - first) / (last - first)) * 128, i = int(f), i == 128 -> 127,
//   edges[i] = i * ((last - first) / 128) + first (edges[128] = last),
//   x < edges[i] -> i - 1;  x >= edges[i + 1] and i != 127 -> i + 1.
//
// Each image is split into chunks, one 256-thread workgroup per (chunk, image), so that
// a search of 25 images fills the device: (1) partial min / max of the non-zero values
// per chunk; (2) the image's range from its partials, the chunk's histogram in LDS (one
// copy per wave, then summed) added to the image's histogram with 128 global atomics;
// (3) one workgroup per image forms the score (a term per bin, summed in bin order).
// The counts are integers and min / max are order-free, so the scores do not depend on
// the chunking.  Images may differ in size (the trial images of several slices scored
// together): each has its own offset, pixel count and chunk count in a descriptor, and
// the grid spans the largest chunk count (surplus workgroups of a smaller image return).
#include "mr_internal.h"

namespace mr {

namespace {

constexpr int CT = 256;
constexpr int NBINS = 128;
constexpr int MAX_CHUNKS = 64;
constexpr int64_t PIX_PER_CHUNK = 16384;

__device__ inline double edge(int i, double first, double last, double step) {
#pragma clang fp contract(off)
    return (i == NBINS) ? last : static_cast<double>(i) * step + first;
}

// Workspace: [nimg][MAX_CHUNKS][2] float partial (min, max), then [nimg][NBINS] counts,
// then [nimg] image descriptors (written by the host before the launches).
__device__ inline float* ws_minmax(void* ws, int img) {
    return static_cast<float*>(ws) + static_cast<int64_t>(img) * MAX_CHUNKS * 2;
}
__device__ inline unsigned* ws_hist(void* ws, int nimg, int img) {
    return reinterpret_cast<unsigned*>(static_cast<float*>(ws) + static_cast<int64_t>(nimg) * MAX_CHUNKS * 2) +
           static_cast<int64_t>(img) * NBINS;
}

__host__ __device__ inline CorImage* ws_images(void* ws, int nimg) {
    return reinterpret_cast<CorImage*>(static_cast<char*>(ws) +
                                       static_cast<size_t>(nimg) * (MAX_CHUNKS * 2 * sizeof(float) + NBINS * sizeof(unsigned)));
}

__device__ inline void chunk_range(int64_t npix, int nchunk, int chunk, int64_t& b, int64_t& e) {
    const int64_t per = (npix + nchunk - 1) / nchunk;
    b = static_cast<int64_t>(chunk) * per;
    e = b + per < npix ? b + per : npix;
}

__global__ __launch_bounds__(CT) void cor_minmax_kernel(const float* __restrict__ img, int nimg,
                                                        void* __restrict__ ws) {
    __shared__ float smin[CT], smax[CT];
    const int im = blockIdx.y, chunk = blockIdx.x, t = threadIdx.x;
    const CorImage d = ws_images(ws, nimg)[im];
    if (chunk >= d.nchunk) return;                                // the whole workgroup
    if (chunk == 0 && t < NBINS) ws_hist(ws, nimg, im)[t] = 0u;    // cleared for pass 2
    int64_t b, e;
    chunk_range(d.npix, d.nchunk, chunk, b, e);
    const float* p = img + d.offset;
    float lo = INFINITY, hi = -INFINITY;
    for (int64_t i = b + t; i < e; i += CT) {
        const float x = p[i];
        if (x != 0.0f) { lo = fminf(lo, x); hi = fmaxf(hi, x); }
    }
    smin[t] = lo;
    smax[t] = hi;
    __syncthreads();
    for (int w = CT / 2; w > 0; w >>= 1) {
        if (t < w) {
            smin[t] = fminf(smin[t], smin[t + w]);
            smax[t] = fmaxf(smax[t], smax[t + w]);
        }
        __syncthreads();
    }
    if (t == 0) {
        float* mm = ws_minmax(ws, im) + 2 * chunk;
        mm[0] = smin[0];
        mm[1] = smax[0];
    }
}

__global__ __launch_bounds__(CT) void cor_hist_kernel(const float* __restrict__ img, int nimg,
                                                      void* __restrict__ ws) {
#pragma clang fp contract(off)
    constexpr int NW = CT / 64;
    __shared__ float smin[MAX_CHUNKS], smax[MAX_CHUNKS];
    __shared__ unsigned hist[NW][NBINS];
    const int im = blockIdx.y, chunk = blockIdx.x, t = threadIdx.x;
    const CorImage d = ws_images(ws, nimg)[im];
    const int nchunk = d.nchunk;
    if (chunk >= nchunk) return;                                  // the whole workgroup
    if (t < nchunk) {
        const float* mm = ws_minmax(ws, im) + 2 * t;
        smin[t] = mm[0];
        smax[t] = mm[1];
    }
    for (int i = t; i < NW * NBINS; i += CT) hist[i / NBINS][i % NBINS] = 0u;
    __syncthreads();
    float lo = INFINITY, hi = -INFINITY;
    for (int i = 0; i < nchunk; ++i) { lo = fminf(lo, smin[i]); hi = fmaxf(hi, smax[i]); }
    if (!(lo <= hi)) return;                 // no non-zero value: empty histogram
    double first = lo, last = hi;
    if (first == last) { first -= 0.5; last += 0.5; }
    const double denom = last - first;
    const double step = denom / NBINS;
    int64_t b0, e0;
    chunk_range(d.npix, nchunk, chunk, b0, e0);
    const float* p = img + d.offset;
    unsigned* h = hist[t >> 6];
    for (int64_t i = b0 + t; i < e0; i += CT) {
        const float xf = p[i];
        if (xf == 0.0f) continue;
        const double x = xf;
        int b = static_cast<int>(((x - first) / denom) * NBINS);
        if (b == NBINS) b -= 1;
        if (x < edge(b, first, last, step)) b -= 1;
        if (b != NBINS - 1 && x >= edge(b + 1, first, last, step)) b += 1;
        atomicAdd(&h[b], 1u);
    }
    __syncthreads();
    if (t < NBINS) {
        unsigned n = 0;
        for (int w = 0; w < NW; ++w) n += hist[w][t];
        if (n) atomicAdd(&ws_hist(ws, nimg, im)[t], n);
    }
}

// One thread per bin forms p ln p; thread 0 sums the terms in bin order.
__global__ __launch_bounds__(NBINS) void cor_entropy_kernel(int nimg, void* __restrict__ ws,
                                                            double* __restrict__ scores) {
    __shared__ double term[NBINS];
    __shared__ double s_total;
    const int im = blockIdx.x, b = threadIdx.x;
    const unsigned* hist = ws_hist(ws, nimg, im);
    if (b == 0) {
        double total = 0.0;
        for (int i = 0; i < NBINS; ++i) total += hist[i];
        s_total = total;
    }
    __syncthreads();
    const double denom_t = s_total > 1.0 ? s_total : 1.0;
    const unsigned h = hist[b];
    double t = 0.0;
    if (h != 0) {
        const double p = h / denom_t;
        t = p * log(p);
    }
    term[b] = t;
    __syncthreads();
    if (b == 0) {
        double s = 0.0;
        for (int i = 0; i < NBINS; ++i)
            if (hist[i] != 0) s += term[i];
        scores[im] = s;
    }
}

int chunks_for(int64_t npix) {
    const int64_t n = (npix + PIX_PER_CHUNK - 1) / PIX_PER_CHUNK;
    return static_cast<int>(n < 1 ? 1 : (n > MAX_CHUNKS ? MAX_CHUNKS : n));
}

}  // namespace

size_t cor_workspace_bytes(int nimg) {
    return static_cast<size_t>(nimg) *
           (MAX_CHUNKS * 2 * sizeof(float) + NBINS * sizeof(unsigned) + sizeof(CorImage));
}

int cor_images(const int64_t* offsets, const int64_t* npix, int nimg, std::vector<CorImage>& out) {
    out.resize(nimg);
    int most = 1;
    for (int i = 0; i < nimg; ++i) {
        out[i].offset = offsets[i];
        out[i].npix = npix[i];
        out[i].nchunk = chunks_for(npix[i]);
        out[i].pad = 0;
        most = std::max(most, out[i].nchunk);
    }
    return most;
}

int write_cor_images(const std::vector<CorImage>& imgs, void* d_ws, hipStream_t s) {
    const int nimg = static_cast<int>(imgs.size());
    if (nimg == 0) return MR_OK;
    MR_HIP(hipMemcpyAsync(ws_images(d_ws, nimg), imgs.data(), sizeof(CorImage) * imgs.size(),
                          hipMemcpyHostToDevice, s));
    // the host vector may go out of scope when the caller returns
    MR_HIP(hipStreamSynchronize(s));
    return MR_OK;
}

int launch_cor_scores(const float* d_img, int nimg, int max_chunks, double* d_scores, void* d_ws,
                      hipStream_t s) {
    if (nimg <= 0) return MR_OK;
    const dim3 grid(std::max(1, std::min(max_chunks, MAX_CHUNKS)), nimg);
    hipLaunchKernelGGL(cor_minmax_kernel, grid, dim3(CT), 0, s, d_img, nimg, d_ws);
    hipLaunchKernelGGL(cor_hist_kernel, grid, dim3(CT), 0, s, d_img, nimg, d_ws);
    hipLaunchKernelGGL(cor_entropy_kernel, dim3(nimg), dim3(NBINS), 0, s, nimg, d_ws, d_scores);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr
