// FDK cosine weighting + zero-padded ramp filter of arriving projections (gfx950).
//
// Replaces fdk_filter_block (recon.py:85-104) and filter_rows (recon.py:58-64):
// irfft(rfft(row, N) * H[:N/2+1], N)[:n].  H is real and even, so filtering is a
// real circular convolution; two detector rows r, r+1 are packed as the real and
// imaginary parts of one complex row z = x_r + i x_{r+1}, transformed with one
// N-point complex FFT, multiplied by H, and transformed back:
//   ifft(Z H) = conj(fft(conj(Z) H)) / N,
// whose real part is row r filtered and imaginary part row r+1 filtered.
//
// The FFT is an in-place LDS Stockham transform with radix-4 passes (one radix-2
// pass first when log2 N is odd).  Each pass reads its butterflies' inputs into
// registers, synchronises, and writes the outputs, so one N-point buffer suffices.
// Twiddles exp(-2 pi i k / N) come from a table computed in float64 on the host.
#include "mr_internal.h"

namespace mr {

namespace {

constexpr int FT = 256;   // threads per workgroup

__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int N, int R>
__device__ void stockham_pass(float2* __restrict__ buf, const float2* __restrict__ tw, int Ns) {
    constexpr int NB = N / R;
    constexpr int PER = (NB + FT - 1) / FT;
    float2 v[PER][R];
    const int tstep_num = N / R;   // twiddle index = r * k * N / (Ns * R)
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int j = threadIdx.x + e * FT;
        if (j < NB) {
            const int k = j & (Ns - 1);
            const int tstep = tstep_num / Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) v[e][r] = buf[j + r * NB];
#pragma unroll
            for (int r = 1; r < R; ++r) v[e][r] = cmul(v[e][r], tw[r * k * tstep]);
            if constexpr (R == 2) {
                const float2 a = v[e][0], b = v[e][1];
                v[e][0] = make_float2(a.x + b.x, a.y + b.y);
                v[e][1] = make_float2(a.x - b.x, a.y - b.y);
            } else {
                const float2 a0 = v[e][0], a1 = v[e][1], a2 = v[e][2], a3 = v[e][3];
                const float2 t0 = make_float2(a0.x + a2.x, a0.y + a2.y);
                const float2 t1 = make_float2(a0.x - a2.x, a0.y - a2.y);
                const float2 t2 = make_float2(a1.x + a3.x, a1.y + a3.y);
                const float2 d = make_float2(a1.x - a3.x, a1.y - a3.y);
                const float2 t3 = make_float2(d.y, -d.x);   // (a1 - a3) * (-i)
                v[e][0] = make_float2(t0.x + t2.x, t0.y + t2.y);
                v[e][1] = make_float2(t1.x + t3.x, t1.y + t3.y);
                v[e][2] = make_float2(t0.x - t2.x, t0.y - t2.y);
                v[e][3] = make_float2(t1.x - t3.x, t1.y - t3.y);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int j = threadIdx.x + e * FT;
        if (j < NB) {
            const int k = j & (Ns - 1);
            const int dst = (j - k) * R + k;   // (j / Ns) * Ns * R + j % Ns
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst + r * Ns] = v[e][r];
        }
    }
    __syncthreads();
}

template <int LOGN>
__device__ void fft_forward(float2* __restrict__ buf, const float2* __restrict__ tw) {
    constexpr int N = 1 << LOGN;
    int Ns = 1;
    if constexpr ((LOGN & 1) != 0) {
        stockham_pass<N, 2>(buf, tw, Ns);
        Ns = 2;
    }
    if constexpr (N >= 4) {
#pragma unroll 1
        for (; Ns < N; Ns *= 4) stockham_pass<N, 4>(buf, tw, Ns);
    }
}

// MODE 0: ring float32; 1: ring bfloat16; 2: float32 staging [k][R][C] plus per-view
// max |q| (for F16 quantisation); 3: plain float32 rows [R][C] (filter_rows hook).
template <int LOGN, int MODE>
__global__ __launch_bounds__(FT) void filter_kernel(
    const float* __restrict__ block, int R, int k, int C, const FilterView* __restrict__ fv,
    const float* __restrict__ H, const float2* __restrict__ tw, void* __restrict__ dst, int pitch,
    int64_t view_stride, unsigned* __restrict__ vmax) {
    constexpr int N = 1 << LOGN;
    __shared__ float2 buf[N];
    const int r0 = 2 * blockIdx.x, r1 = r0 + 1;
    const int v = blockIdx.y;
    const FilterView P = fv[v];
    const bool has1 = r1 < R;
    // MODE 4 (adjoint) reads views [k][R][C]; the others read blocks [R][k][C]
    const float* x0 = MODE == 4 ? block + (static_cast<int64_t>(v) * R + r0) * C
                                : block + (static_cast<int64_t>(r0) * k + v) * C;
    const float* x1 = MODE == 4 ? block + (static_cast<int64_t>(v) * R + (has1 ? r1 : r0)) * C
                                : block + (static_cast<int64_t>(has1 ? r1 : r0) * k + v) * C;
    const float b0 = P.b0 + r0 * P.db, b1 = P.b0 + r1 * P.db;
    for (int c = threadIdx.x; c < N; c += FT) {
        float y0 = 0.0f, y1 = 0.0f;
        if (c < C) {
            y0 = x0[c];
            y1 = has1 ? x1[c] : 0.0f;
            if (MODE != 4 && P.cone) {
                const float a = P.a0 + c * P.da;
                const float aa = fmaf(a, a, P.sdd2);
                y0 *= P.sdd * __frsqrt_rn(fmaf(b0, b0, aa));
                y1 *= P.sdd * __frsqrt_rn(fmaf(b1, b1, aa));
            }
        }
        buf[c] = make_float2(y0, y1);
    }
    __syncthreads();
    fft_forward<LOGN>(buf, tw);
    for (int c = threadIdx.x; c < N; c += FT) {
        const float h = H[c];
        const float2 z = buf[c];
        buf[c] = make_float2(z.x * h, -z.y * h);
    }
    __syncthreads();
    fft_forward<LOGN>(buf, tw);

    const float g = P.gain;
    float lmax = 0.0f;
    for (int c = threadIdx.x; c < C; c += FT) {
        const float2 z = buf[c];
        const float y0 = z.x * g, y1 = -z.y * g;
        if (MODE == 0) {
            float* q = static_cast<float*>(dst) + static_cast<int64_t>(P.slot) * view_stride;
            q[static_cast<int64_t>(r0 + 1) * pitch + c + 1] = y0;
            if (has1) q[static_cast<int64_t>(r1 + 1) * pitch + c + 1] = y1;
        } else if (MODE == 1) {
            uint16_t* q = static_cast<uint16_t*>(dst) + static_cast<int64_t>(P.slot) * view_stride;
            auto bf = [](float f) -> uint16_t {
                uint32_t u = __float_as_uint(f);
                u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
                return static_cast<uint16_t>(u >> 16);
            };
            q[static_cast<int64_t>(r0 + 1) * pitch + c + 1] = bf(y0);
            if (has1) q[static_cast<int64_t>(r1 + 1) * pitch + c + 1] = bf(y1);
        } else if (MODE == 2) {
            float* q = static_cast<float*>(dst) + static_cast<int64_t>(v) * R * C;
            q[static_cast<int64_t>(r0) * C + c] = y0;
            lmax = fmaxf(lmax, fabsf(y0));
            if (has1) {
                q[static_cast<int64_t>(r1) * C + c] = y1;
                lmax = fmaxf(lmax, fabsf(y1));
            }
        } else if (MODE == 3) {
            float* q = static_cast<float*>(dst);
            q[static_cast<int64_t>(r0) * C + c] = y0;
            if (has1) q[static_cast<int64_t>(r1) * C + c] = y1;
        } else {
            // adjoint: the cosine weight after the (self-adjoint) filter, block layout
            float w0 = 1.0f, w1 = 1.0f;
            if (P.cone) {
                const float a = P.a0 + c * P.da;
                const float aa = fmaf(a, a, P.sdd2);
                w0 = P.sdd * __frsqrt_rn(fmaf(b0, b0, aa));
                w1 = P.sdd * __frsqrt_rn(fmaf(b1, b1, aa));
            }
            float* q = static_cast<float*>(dst);
            q[(static_cast<int64_t>(r0) * k + v) * C + c] = y0 * w0;
            if (has1) q[(static_cast<int64_t>(r1) * k + v) * C + c] = y1 * w1;
        }
    }
    if (MODE == 2) {
        // workgroup max through LDS, then one atomic per workgroup
        __shared__ float red[FT];
        red[threadIdx.x] = lmax;
        __syncthreads();
        for (int w = FT / 2; w > 0; w >>= 1) {
            if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicMax(vmax + v, __float_as_uint(red[0]));
    }
}

// F16 storage: power-of-two scale per view so that max |q| / scale <= 2^15, then
// round to nearest half.  The scale is stored per slot and folded into the
// back-projection weight.
__global__ void quantize_f16_kernel(const float* __restrict__ stage, int R, int C,
                                    const FilterView* __restrict__ fv,
                                    const unsigned* __restrict__ vmax, __half* __restrict__ ring,
                                    int pitch, int64_t view_stride, float* __restrict__ vscale) {
    const int v = blockIdx.y, r = blockIdx.x;
    const float mx = __uint_as_float(vmax[v]);
    float s = 1.0f;
    if (mx > 0.0f && isfinite(mx)) {
        int e;
        frexpf(mx, &e);        // mx = f * 2^e, 0.5 <= f < 1  ->  mx < 2^e
        s = ldexpf(1.0f, e - 15);
    }
    const int slot = fv[v].slot;
    if (r == 0 && threadIdx.x == 0) vscale[slot] = s;
    const float inv = 1.0f / s;
    const float* src = stage + (static_cast<int64_t>(v) * R + r) * C;
    __half* dst = ring + static_cast<int64_t>(slot) * view_stride + static_cast<int64_t>(r + 1) * pitch + 1;
    for (int c = threadIdx.x; c < C; c += blockDim.x) dst[c] = __float2half_rn(src[c] * inv);
}

// Store of already-filtered views q [k][R][C] into the ring (F32 / BF16 directly).
template <int MODE>
__global__ void store_kernel(const float* __restrict__ q, int R, int C,
                             const FilterView* __restrict__ fv, void* __restrict__ ring, int pitch,
                             int64_t view_stride) {
    const int v = blockIdx.y, r = blockIdx.x;
    const int slot = fv[v].slot;
    const float* src = q + (static_cast<int64_t>(v) * R + r) * C;
    const int64_t o = static_cast<int64_t>(slot) * view_stride + static_cast<int64_t>(r + 1) * pitch + 1;
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        const float y = src[c];
        if (MODE == 0) {
            static_cast<float*>(ring)[o + c] = y;
        } else {
            uint32_t u = __float_as_uint(y);
            u += 0x7FFFu + ((u >> 16) & 1u);
            static_cast<uint16_t*>(ring)[o + c] = static_cast<uint16_t>(u >> 16);
        }
    }
}

__global__ void zero_u32_kernel(unsigned* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Streaming update, before filtering: copy the staged per-view geometry of a batch into
// its ring slots and advance the device view count by k (graph-replayable: every
// address is fixed, the slots come from the staged metadata).
__global__ void stream_prep_kernel(const FilterView* __restrict__ fv, const ViewGeom* __restrict__ gv,
                                   int k, ViewGeom* __restrict__ geom, float* __restrict__ vscale,
                                   int* __restrict__ nviews, int store_f16) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < k) {
        geom[fv[j].slot] = gv[j];
        if (!store_f16) vscale[fv[j].slot] = 1.0f;
    }
    if (j == 0) atomicAdd(nviews, k);
}

// Per-view max |q| of q [k][R][C] (F16 quantisation of stored views).
__global__ void vmax_kernel(const float* __restrict__ q, int R, int C, unsigned* __restrict__ vmax) {
    const int v = blockIdx.y, r = blockIdx.x;
    const float* src = q + (static_cast<int64_t>(v) * R + r) * C;
    float m = 0.0f;
    for (int c = threadIdx.x; c < C; c += blockDim.x) m = fmaxf(m, fabsf(src[c]));
    __shared__ float red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(vmax + v, __float_as_uint(red[0]));
}

template <int MODE>
int launch_mode(int logn, dim3 grid, hipStream_t s, const float* block, int R, int k, int C,
                const FilterView* fv, const float* H, const float2* tw, void* dst, int pitch,
                int64_t view_stride, unsigned* vmax) {
#define MR_FILTER_CASE(L)                                                                  \
    case L:                                                                                \
        hipLaunchKernelGGL((filter_kernel<L, MODE>), grid, dim3(FT), 0, s, block, R, k, C, \
                           fv, H, tw, dst, pitch, view_stride, vmax);                      \
        break;
    switch (logn) {
        MR_FILTER_CASE(1) MR_FILTER_CASE(2) MR_FILTER_CASE(3) MR_FILTER_CASE(4) MR_FILTER_CASE(5)
        MR_FILTER_CASE(6) MR_FILTER_CASE(7) MR_FILTER_CASE(8) MR_FILTER_CASE(9)
        MR_FILTER_CASE(10) MR_FILTER_CASE(11) MR_FILTER_CASE(12) MR_FILTER_CASE(13)
    default:
        return fail(MR_EVALUE, "FFT length 2^" + std::to_string(logn) +
                                   " outside the supported range 2..8192");
    }
#undef MR_FILTER_CASE
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int log2i(int N) {
    int l = 0;
    while ((1 << l) < N) ++l;
    return l;
}

}  // namespace

int launch_filter(Ctx* c, const float* d_block, int k, int n0, hipStream_t s,
                  const FilterView* fv) {
    (void)n0;
    if (!fv) fv = c->d_fv;
    const int logn = log2i(c->N);
    dim3 grid((c->R + 1) / 2, k);
    if (c->storage == MR_STORE_F32)
        return launch_mode<0>(logn, grid, s, d_block, c->R, k, c->C, fv, c->d_H, c->d_tw,
                              c->ring, c->pitch, c->view_stride, nullptr);
    if (c->storage == MR_STORE_BF16)
        return launch_mode<1>(logn, grid, s, d_block, c->R, k, c->C, fv, c->d_H, c->d_tw,
                              c->ring, c->pitch, c->view_stride, nullptr);
    // F16: filter to float32 staging, then quantise with a per-view scale (the max
    // buffer is cleared with a kernel so that the sequence can be graph-captured)
    hipLaunchKernelGGL(zero_u32_kernel, dim3((k + 255) / 256), dim3(256), 0, s, c->d_vmax, k);
    int rc = launch_mode<2>(logn, grid, s, d_block, c->R, k, c->C, fv, c->d_H, c->d_tw,
                            c->d_fstage, c->pitch, c->view_stride, c->d_vmax);
    if (rc) return rc;
    hipLaunchKernelGGL(quantize_f16_kernel, dim3(c->R, k), dim3(256), 0, s, c->d_fstage, c->R,
                       c->C, fv, c->d_vmax, static_cast<__half*>(c->ring), c->pitch,
                       c->view_stride, c->d_vscale);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int launch_store(Ctx* c, const float* d_q, int k, hipStream_t s) {
    dim3 grid(c->R, k), block(256);
    if (c->storage == MR_STORE_F32) {
        hipLaunchKernelGGL(store_kernel<0>, grid, block, 0, s, d_q, c->R, c->C, c->d_fv, c->ring,
                           c->pitch, c->view_stride);
    } else if (c->storage == MR_STORE_BF16) {
        hipLaunchKernelGGL(store_kernel<1>, grid, block, 0, s, d_q, c->R, c->C, c->d_fv, c->ring,
                           c->pitch, c->view_stride);
    } else {
        hipLaunchKernelGGL(zero_u32_kernel, dim3((k + 255) / 256), dim3(256), 0, s, c->d_vmax, k);
        hipLaunchKernelGGL(vmax_kernel, grid, block, 0, s, d_q, c->R, c->C, c->d_vmax);
        hipLaunchKernelGGL(quantize_f16_kernel, grid, block, 0, s, d_q, c->R, c->C, c->d_fv,
                           c->d_vmax, static_cast<__half*>(c->ring), c->pitch, c->view_stride,
                           c->d_vscale);
    }
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int launch_filter_adjoint(Ctx* c, const float* d_gq, int k, float* d_graw, hipStream_t s) {
    dim3 grid((c->R + 1) / 2, k);
    return launch_mode<4>(log2i(c->N), grid, s, d_gq, c->R, k, c->C, c->d_fv, c->d_H, c->d_tw,
                          d_graw, 0, 0, nullptr);
}

int launch_stream_prep(Ctx* c, const void* d_meta, int k, int* d_nviews, hipStream_t s) {
    const FilterView* fv = static_cast<const FilterView*>(d_meta);
    const ViewGeom* gv = reinterpret_cast<const ViewGeom*>(
        static_cast<const char*>(d_meta) + sizeof(FilterView) * k);
    hipLaunchKernelGGL(stream_prep_kernel, dim3((k + 63) / 64), dim3(64), 0, s, fv, gv, k, c->d_geom,
                       c->d_vscale, d_nviews, c->storage == MR_STORE_F16 ? 1 : 0);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int launch_filter_rows_plain(Ctx* c, const float* d_in, int rows, int n, float gain,
                             const float* d_H, const float2* d_tw, int N, float* d_out,
                             hipStream_t s) {
    // one FilterView with no FDK weight: reuse the context's filter-view buffer
    FilterView fv{};
    fv.gain = gain;
    fv.cone = 0;
    int rc = ensure(reinterpret_cast<void**>(&c->d_fv), &c->fv_len, sizeof(FilterView));
    if (rc) return rc;
    MR_HIP(hipStreamSynchronize(s));
    MR_HIP(hipMemcpy(c->d_fv, &fv, sizeof(fv), hipMemcpyHostToDevice));
    dim3 grid((rows + 1) / 2, 1);
    return launch_mode<3>(log2i(N), grid, s, d_in, rows, 1, n, c->d_fv, d_H, d_tw, d_out, 0, 0,
                          nullptr);
}

}  // namespace mr
