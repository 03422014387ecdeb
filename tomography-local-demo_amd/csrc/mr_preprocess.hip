// Per-projection preprocessing on the device (preprocess.py:8-64).
//
// Blocks are [R][k][C] float32 (rows, views, columns).
//
// rotate_projections (preprocess.py:31-56): every output pixel samples the input at
// the inverse-rotated position (float64 coordinates) with a bilinear whose taps are
// zero outside, summed tap by tap in float32 as the reference does.
//
// gaussian_blur (preprocess.py:8-28): per axis (rows, then columns), the reference
// edge-pads each line by pad = ceil(3 sigma) to length N = n + 2 pad and multiplies
// its rfft by H = exp(-2 (pi f sigma)^2).  That is a circular convolution of the
// padded line with g = irfft(H, N).  The host computes g exactly in float64 and keeps
// the taps m in [-L, L] outside which |g| < 1e-13 g(0) (all N taps for small sigma);
// the device sums them in float64 and rounds each pass to float32 like the reference.
#include <cmath>

#include "mr_internal.h"

namespace mr {

namespace {

__global__ void rotate_kernel(const float* __restrict__ in, float* __restrict__ out, int R, int k,
                              int C, double ct, double st) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    const int v = blockIdx.z;
    if (c >= C) return;
    const double rr = r - (R - 1) / 2.0, cc = c - (C - 1) / 2.0;
    const double rs = ct * rr + st * cc + (R - 1) / 2.0;
    const double cs = -st * rr + ct * cc + (C - 1) / 2.0;
    const double r0f = floor(rs), c0f = floor(cs);
    const int r0 = static_cast<int>(r0f), c0 = static_cast<int>(c0f);
    const double fr = rs - r0f, fc = cs - c0f;
    float acc = 0.0f;
#pragma unroll
    for (int dr = 0; dr < 2; ++dr)
#pragma unroll
        for (int dc = 0; dc < 2; ++dc) {
            const int ri = r0 + dr, ci = c0 + dc;
            const double wgt = (dr ? fr : 1.0 - fr) * (dc ? fc : 1.0 - fc);
            float val = 0.0f;
            if (ri >= 0 && ri < R && ci >= 0 && ci < C)
                val = in[(static_cast<int64_t>(ri) * k + v) * C + ci];
            acc += static_cast<float>(wgt * val);
        }
    out[(static_cast<int64_t>(r) * k + v) * C + c] = acc;
}

// y[i] = sum_t g[t] * xp[(i + pad - (m0 + t)) mod N], xp[j] = x[clamp(j - pad, 0, n - 1)]
// AXIS 0: lines run along r (stride k*C); AXIS 2: along c (stride 1).
template <int AXIS>
__global__ void gauss_kernel(const float* __restrict__ in, float* __restrict__ out, int R, int k,
                             int C, const double* __restrict__ g, int M, int m0, int pad, int N) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    const int v = blockIdx.z;
    if (c >= C) return;
    const int n = AXIS == 0 ? R : C;
    const int i = AXIS == 0 ? r : c;
    const int64_t stride = AXIS == 0 ? static_cast<int64_t>(k) * C : 1;
    const float* line = AXIS == 0 ? in + static_cast<int64_t>(v) * C + c
                                  : in + (static_cast<int64_t>(r) * k + v) * C;
    double acc = 0.0;
    for (int t = 0; t < M; ++t) {
        int j = (i + pad - (m0 + t)) % N;
        if (j < 0) j += N;
        int src = j - pad;
        src = src < 0 ? 0 : (src > n - 1 ? n - 1 : src);
        acc += g[t] * static_cast<double>(line[src * stride]);
    }
    out[(static_cast<int64_t>(r) * k + v) * C + c] = static_cast<float>(acc);
}

// Taps of g = irfft(H, N), H = exp(-2 (pi f sigma)^2), in float64 (exact cosine sum).
// For sigma >~ 2.4 the spectrum is negligible before Nyquist and g decays like a
// Gaussian; for smaller sigma H(1/2) is not small and g keeps an oscillating tail of
// amplitude ~H(1/2)/N, so all N taps are needed.  The window [-L, L] is the smallest
// one outside which every |g| < 1e-13 g(0); otherwise the full circle is used.
void gauss_taps(int n, double sigma, int* pad_out, int* N_out, int* m0_out,
                std::vector<double>& g) {
    const int pad = static_cast<int>(std::ceil(3.0 * sigma));
    const int N = n + 2 * pad;
    const int K = N / 2 + 1;                         // rfft bins
    std::vector<double> H(K);
    for (int q = 0; q < K; ++q) {
        const double f = static_cast<double>(q) / N;  // rfftfreq
        const double a = M_PI * f * sigma;
        H[q] = std::exp(-2.0 * a * a);
    }
    std::vector<double> full(N);
    for (int m = 0; m < N; ++m) {
        double s = H[0];
        for (int q = 1; q < K; ++q) {
            const double w = (N % 2 == 0 && q == N / 2) ? 1.0 : 2.0;
            s += w * H[q] * std::cos(2.0 * M_PI * static_cast<double>(q) * m / N);
        }
        full[m] = s / N;
    }
    int L = N / 2;
    const double thr = 1e-13 * std::fabs(full[0]);
    for (int m = N / 2; m >= 1; --m) {
        if (std::fabs(full[m]) >= thr || std::fabs(full[N - m]) >= thr) { L = m; break; }
        L = m - 1;
    }
    int m0, M;
    if (2 * L + 1 >= N) { m0 = 0; M = N; } else { m0 = -L; M = 2 * L + 1; }
    g.resize(M);
    for (int t = 0; t < M; ++t) {
        int m = (m0 + t) % N;
        if (m < 0) m += N;
        g[t] = full[m];
    }
    *pad_out = pad;
    *N_out = N;
    *m0_out = m0;
}

// Raw detector counts -> line integrals (flexray.py:111-116 + 174-184, hdf5.py:24-31 +
// 77-78).  Frame v of `frames` is a [H][W] uint16 detector image.  Output pixel (r, c)
// of the [R][k][C] block is binned frame pixel (row0 + r, col0 + c), counted after the
// optional vertical flip (flexray stores detector rows bottom-up):
//   x = mean of the sample x sample raw counts (an exact integer sum, one rounding)
//   y = (x - dark) * scale   (op 0, flexray: scale = 1 / max(flat - dark, 1))
//   y = (x - dark) / scale   (op 1, hdf5:    scale = max(flat - dark, 1e-6))
//   out = -log(max(y, 1e-3))  (NaN stays NaN, like np.clip)
template <int FLIP>
__global__ void raw_correct_kernel(const uint16_t* __restrict__ frames, int H, int W, int k,
                                   int R, int C, int sample, int row0, int col0,
                                   const float* __restrict__ dark, const float* __restrict__ scale,
                                   int op, float* __restrict__ out) {
    const int c = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    const int v = blockIdx.z;
    if (c >= C) return;
    int br = row0 + r;
    if (FLIP) br = H / sample - 1 - br;
    const int bc = col0 + c;
    const uint16_t* f = frames + static_cast<int64_t>(v) * H * W;
    uint32_t sum = 0;
    for (int i = 0; i < sample; ++i) {
        const uint16_t* row = f + static_cast<int64_t>(br * sample + i) * W + bc * sample;
        for (int j = 0; j < sample; ++j) sum += row[j];
    }
    float x = static_cast<float>(sum);
    if (sample > 1) x = x / static_cast<float>(sample * sample);
    const float d = dark[r * C + c], g = scale[r * C + c];
    float y = op == 0 ? (x - d) * g : (x - d) / g;
    y = y < 1e-3f ? 1e-3f : y;
    out[(static_cast<int64_t>(r) * k + v) * C + c] = -logf(y);
}

}  // namespace

// d_a holds the block; the result ends in *d_res (d_a or d_b).
int launch_preprocess(Ctx* c, float* d_a, float* d_b, int R, int k, int C, double gaussian,
                      double tilt, float** d_res, hipStream_t s) {
    float* cur = d_a;
    float* oth = d_b;
    dim3 grid((C + 255) / 256, R, k), block(256);
    if (tilt != 0.0 && std::fabs(tilt) >= 1e-9) {
        // apply() rotates by -tilt (preprocess.py:61); rotate_projections uses
        // t = deg2rad(degrees)
        const double t = -tilt * M_PI / 180.0;
        hipLaunchKernelGGL(rotate_kernel, grid, block, 0, s, cur, oth, R, k, C, std::cos(t),
                           std::sin(t));
        MR_HIP(hipGetLastError());
        std::swap(cur, oth);
    }
    if (gaussian > 0.0) {
        for (int axis = 0; axis < 2; ++axis) {
            const int n = axis == 0 ? R : C;
            int pad, N, m0;
            std::vector<double> g;
            gauss_taps(n, gaussian, &pad, &N, &m0, g);
            int rc = ensure(reinterpret_cast<void**>(&c->d_misc), &c->misc_bytes,
                            sizeof(double) * g.size());
            if (rc) return rc;
            MR_HIP(hipMemcpyAsync(c->d_misc, g.data(), sizeof(double) * g.size(),
                                  hipMemcpyHostToDevice, s));
            if (axis == 0)
                hipLaunchKernelGGL(gauss_kernel<0>, grid, block, 0, s, cur, oth, R, k, C, c->d_misc,
                                   static_cast<int>(g.size()), m0, pad, N);
            else
                hipLaunchKernelGGL(gauss_kernel<2>, grid, block, 0, s, cur, oth, R, k, C, c->d_misc,
                                   static_cast<int>(g.size()), m0, pad, N);
            MR_HIP(hipGetLastError());
            // the taps buffer is reused by the next pass: wait for this one
            MR_HIP(hipStreamSynchronize(s));
            std::swap(cur, oth);
        }
    }
    *d_res = cur;
    return MR_OK;
}

int launch_raw_correct(Ctx* c, const uint16_t* d_frames, int k, float* d_out, hipStream_t s) {
    const RawCorr& q = c->raw;
    dim3 grid((c->C + 255) / 256, c->R, k), block(256);
    if (q.flip)
        hipLaunchKernelGGL(raw_correct_kernel<1>, grid, block, 0, s, d_frames, q.H, q.W, k, c->R,
                           c->C, q.sample, q.row0, q.col0, q.d_dark, q.d_scale, q.op, d_out);
    else
        hipLaunchKernelGGL(raw_correct_kernel<0>, grid, block, 0, s, d_frames, q.H, q.W, k, c->R,
                           c->C, q.sample, q.row0, q.col0, q.d_dark, q.d_scale, q.op, d_out);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr
