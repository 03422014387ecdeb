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
// The FFT is a Stockham transform in radix-16 passes (one radix-2/4/8 pass first when
// log2 N is not a multiple of 4): each thread holds the 16 inputs of a butterfly in
// registers, twiddles them, runs a 16-point DFT in registers and writes the outputs, so
// N = 4096 takes 3 passes and 5 barriers for both transforms instead of 12 radix-4 passes.
// The first pass reads the weighted, zero-padded rows straight from global memory, the
// multiplication by H is fused into the first pass of the second transform, and its last
// pass writes the filtered rows out (in natural order, coalesced).  Between passes the
// N-point buffer lives in LDS with the low 4 index bits XORed with the next 4, so that
// the stride-16 writes of the first pass do not conflict in the banks.  Twiddles
// exp(-2 pi i k / N) come from a table computed in float64 on the host.
#include "mr_internal.h"

namespace mr {

namespace {

// Threads per workgroup: 256 (one radix-16 butterfly per thread per pass at N = 4096),
// 512 for N = 8192, 1024 for N = 16384 (128 KB of LDS, one workgroup per CU).
constexpr int ft_of(int logn) { return logn >= 14 ? 1024 : logn >= 13 ? 512 : 256; }

// Longest transform one workgroup holds in LDS; longer rows (up to 16 times that) take the
// four-step path.
constexpr int kLdsMaxLogN = kFilterLdsLogN;
constexpr int kMaxLogN = kFilterMaxLogN;
// rows per view whose max |x| gives the provisional f16 scale (every ceil(R / 96)-th row)
constexpr int kF16SampleRows = 96;
static_assert(kMaxLogN == kLdsMaxLogN + 4, "four-step: 16-point column transforms");

__device__ inline float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// a * exp(-2 pi i e / 16), e = 0..7 (a constant once the DFT loops are unrolled):
// (x + iy)(cos t - i sin t) = (x cos t + y sin t) + i (y cos t - x sin t)
__device__ __forceinline__ float2 w16(float2 a, int e) {
    constexpr float C1 = 0.92387953251128674f;   // cos(pi/8)
    constexpr float S1 = 0.38268343236508977f;   // sin(pi/8)
    constexpr float H2 = 0.70710678118654752f;   // cos(pi/4)
    switch (e) {
    case 0: return a;
    case 1: return make_float2(a.x * C1 + a.y * S1, a.y * C1 - a.x * S1);
    case 2: return make_float2((a.x + a.y) * H2, (a.y - a.x) * H2);
    case 3: return make_float2(a.x * S1 + a.y * C1, a.y * S1 - a.x * C1);
    case 4: return make_float2(a.y, -a.x);
    case 5: return make_float2(a.y * C1 - a.x * S1, -(a.y * S1 + a.x * C1));
    case 6: return make_float2((a.y - a.x) * H2, -(a.x + a.y) * H2);
    default: return make_float2(a.y * S1 - a.x * C1, -(a.y * C1 + a.x * S1));
    }
}

// In-register DFT of R = 2..16 points, decimation in frequency: x[bitrev(q)] = X[q].
template <int R>
__device__ __forceinline__ void dft_dif(float2 (&x)[R]) {
#pragma unroll
    for (int s = R / 2; s >= 1; s >>= 1) {
#pragma unroll
        for (int b = 0; b < R; b += 2 * s) {
#pragma unroll
            for (int i = 0; i < s; ++i) {
                const float2 a = x[b + i], c = x[b + i + s];
                x[b + i] = cadd(a, c);
                x[b + i + s] = w16(csub(a, c), i * (8 / s));   // W_{2s}^i = W_16^{8i/s}
            }
        }
    }
}

constexpr int bitrev(int q, int R) {
    int r = 0;
    for (int b = 1; b < R; b <<= 1) {
        r = (r << 1) | (q & 1);
        q >>= 1;
    }
    return r;
}

// LDS index of buffer element i: the low 4 bits XORed with the next 4 (a bijection).
__device__ __forceinline__ int sw(int i) { return i ^ ((i >> 4) & 15); }

// Passes of an N = 2^LOGN transform: radix 2^(LOGN mod 4) first (if not 1), then 16s;
// REV: the same radices in the reverse order (the small one last).
template <int LOGN, bool REV = false> struct FftPlan {
    static constexpr int REM = LOGN % 4;
    static constexpr int NPASS = LOGN / 4 + (REM ? 1 : 0);
    static constexpr int radix_fwd(int p) { return (REM && p == 0) ? (1 << REM) : 16; }
    static constexpr int radix(int p) { return REV ? radix_fwd(NPASS - 1 - p) : radix_fwd(p); }
    static constexpr int span(int p) {          // product of the radices before pass p
        int n = 1;
        for (int i = 0; i < p; ++i) n *= radix(i);
        return n;
    }
};

// In-register DIF DFT whose inputs R/2 .. R-1 are zero: the first stage only twiddles.
template <int R>
__device__ __forceinline__ void dft_dif_upper_zero(float2 (&x)[R]) {
    if constexpr (R == 1) {
        return;
    } else {
        constexpr int s = R / 2;
#pragma unroll
        for (int i = 0; i < s; ++i) x[i + s] = w16(x[i], i * (8 / s));
#pragma unroll
        for (int t = s / 2; t >= 1; t >>= 1) {
#pragma unroll
            for (int b = 0; b < R; b += 2 * t) {
#pragma unroll
                for (int i = 0; i < t; ++i) {
                    const float2 a = x[b + i], c = x[b + i + t];
                    x[b + i] = cadd(a, c);
                    x[b + i + t] = w16(csub(a, c), i * (8 / t));
                }
            }
        }
    }
}

// One Stockham pass of radix R over span NS (the product of the earlier radices):
// butterfly j (k = j mod NS) takes in(j + r N/R), r < R, twiddles input r by
// exp(-2 pi i r k / (NS R)) and computes its R-point DFT in v (outputs in bit-reversed
// order); pass_store then writes DFT output q to out((j - k) R + k + q NS).  Thread t
// owns butterflies t + e FT, e < PER.  UZ: inputs r >= R/2 are known to be zero (the
// zero padding of the first pass) and are neither read nor added.
template <int N, int R, int FT> struct Pass {
    static constexpr int NB = N / R;
    static constexpr int PER = (NB + FT - 1) / FT;
    using Regs = float2[PER][R];
    // the thread index, opaque to the compiler: index expressions of different passes
    // are not merged (they would stay live across the passes and spill at N = 4096)
    __device__ static int tid() {
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        return t;
    }
    __device__ static bool has(int t, int e) { return NB % FT == 0 || t + e * FT < NB; }
    template <int NS, bool UZ, typename In>
    __device__ static void load_dft(const float2* __restrict__ tw, In in, Regs& v) {
        const int t = tid();
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int j = t + e * FT;
            if (!has(t, e)) continue;
            const int k = j & (NS - 1);
            // the twiddles first, all in flight together, then the inputs
            float2 w[R];
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) w[r] = tw[r * k * (N / (NS * R))];
            }
#pragma unroll
            for (int r = 0; r < (UZ ? R / 2 : R); ++r) v[e][r] = in(j + r * NB);
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) v[e][r] = cmul(v[e][r], w[r]);
            }
            if constexpr (UZ) dft_dif_upper_zero<R>(v[e]);
            else dft_dif<R>(v[e]);
        }
    }
    template <int NS, typename Out>
    __device__ static void store(Out out, const Regs& v) {
        const int t = tid();
#pragma unroll
        for (int e = 0; e < PER; ++e) {
            const int j = t + e * FT;
            if (!has(t, e)) continue;
            const int k = j & (NS - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) out(base + q * NS, v[e][bitrev(q, R)]);
        }
    }
};

// Passes P .. LAST - 1 of plan PL through the LDS buffer, in place: each reads buf,
// waits for every thread's reads, writes buf and waits again.
template <typename PL, int N, int FT, int P, int LAST>
__device__ __forceinline__ void lds_passes(const float2* __restrict__ tw, float2* buf) {
    if constexpr (P < LAST) {
        constexpr int R = PL::radix(P), NS = PL::span(P);
        using PS = Pass<N, R, FT>;
        typename PS::Regs v;
        PS::template load_dft<NS, false>(tw, [buf](int i) { return buf[sw(i)]; }, v);
        __syncthreads();
        PS::template store<NS>([buf](int i, float2 z) { buf[sw(i)] = z; }, v);
        __syncthreads();
        // keep the next pass's twiddle loads below the barrier (hoisting them above it
        // raises the register pressure past 128 VGPRs at N = 4096)
        __builtin_amdgcn_sched_barrier(0);
        lds_passes<PL, N, FT, P + 1, LAST>(tw, buf);
    }
}

// Input of transform 1: rows r0 and r0 + 1 of view v as the real and imaginary parts (the
// second is zero past the last row), cone views cosine-weighted except in MODE 4; columns
// >= C read as zero.  MODE 4 reads views [k][R][C], the others blocks [R][k][C].
template <int MODE> struct RowsIn {
    const float* x0;
    const float* x1;
    FilterView P;
    float b0, b1;
    int C;
    bool has1;
    __device__ RowsIn(const float* block, int R, int k, int C_, const FilterView& P_, int r0, int v)
        : P(P_), C(C_) {
        const int r1 = r0 + 1;
        has1 = r1 < R;
        const int rr = has1 ? r1 : r0;
        x0 = MODE == 4 ? block + (static_cast<int64_t>(v) * R + r0) * C
                       : block + (static_cast<int64_t>(r0) * k + v) * C;
        x1 = MODE == 4 ? block + (static_cast<int64_t>(v) * R + rr) * C
                       : block + (static_cast<int64_t>(rr) * k + v) * C;
        b0 = P.b0 + r0 * P.db;
        b1 = P.b0 + r1 * P.db;
    }
    // Branch-free: the column is clamped into the row (x1 is row r0 itself when there is no
    // second row) and the value masked by a multiply, so that the loads of all the inputs a
    // thread reads are issued together rather than one guarded pair at a time (a select of
    // the loaded value is turned back into a branch around the load).  In-range values are
    // multiplied by exactly 1; a non-finite input already poisons its whole row pair
    // through the transform.
    __device__ float2 operator()(int c) const {
        const bool in = c < C;
        const int cc = in ? c : C - 1;
        float y0 = x0[cc] * (in ? 1.0f : 0.0f);
        float y1 = x1[cc] * (in && has1 ? 1.0f : 0.0f);
        if (MODE != 4) {
            const float a = P.a0 + cc * P.da;
            const float aa = fmaf(a, a, P.sdd2);
            const float w0 = P.sdd * __frsqrt_rn(fmaf(b0, b0, aa));
            const float w1 = P.sdd * __frsqrt_rn(fmaf(b1, b1, aa));
            y0 = P.cone ? y0 * w0 : y0;
            y1 = P.cone ? y1 * w1 : y1;
        }
        return make_float2(y0, y1);
    }
};

// f16 storage: the power-of-two scale 2^(e - 15) of a view whose filtered values are
// bounded by `bound` < 2^e, so |q| / scale < 2^15 (1 for a zero or non-finite bound).
// Any power-of-two scale stores the same products half(q / scale) * scale while q / scale
// stays within the normal half range, so the scale only has to keep the view's largest
// value between 2^6 and the largest half.
__device__ inline float f16_scale(float bound) {
    float sc = 1.0f;
    if (bound > 0.0f && isfinite(bound)) {
        int ex;
        frexpf(bound, &ex);
        sc = ldexpf(1.0f, ex - 15);
    }
    return sc;
}
// The provisional scale of batch view v (MODE 6): the bound max|q| <= max|x| gain l1 from
// the max |x| of sampled rows, kept per part of the sample (sample_max_kernel stores each
// part's bound; rounding is monotone, so their max is the bound of the max).
__device__ inline float f16_prov_scale(const unsigned* vmax, int v) {
    const unsigned* p = vmax + kF16SampleParts * v;
    unsigned m = p[0];
#pragma unroll
    for (int j = 1; j < kF16SampleParts; ++j) m = max(m, p[j]);
    return f16_scale(__uint_as_float(m));
}
// Max |q| slot j of batch view v: wave w of workgroup x of the view raises slot (x waves +
// w) mod kF16QmaxSlots in a MODE 6 pass.  Slots lie in separate 128-byte lines: atomics on
// one line (or one address) per view would serialise its ~2R atomics.
template <typename U> __device__ inline U* f16_qmax_slot(U* vmax, int k, int v, int j) {
    return vmax + kF16SampleParts * k + (kF16QmaxSlots * v + j) * kF16QmaxStride;
}
// The exact max |q| of batch view v gathered by a MODE 6 pass.
__device__ inline float f16_view_qmax(const unsigned* vmax, int k, int v) {
    unsigned m = *f16_qmax_slot(vmax, k, v, 0);
#pragma unroll
    for (int j = 1; j < kF16QmaxSlots; ++j) m = max(m, *f16_qmax_slot(vmax, k, v, j));
    return __uint_as_float(m);
}
// Largest of a wave's non-negative floats (their bits order as unsigned integers): the
// max within each row of 16 lanes by DPP moves (quad permutations, half-row and row
// mirrors), then across the four rows by lane reads.
__device__ inline float wave_max_nonneg(float x) {
    unsigned u = __float_as_uint(x);
    u = max(u, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u), 0xB1, 0xF, 0xF, false)));
    u = max(u, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u), 0x4E, 0xF, 0xF, false)));
    u = max(u, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u), 0x141, 0xF, 0xF, false)));
    u = max(u, static_cast<unsigned>(__builtin_amdgcn_update_dpp(0, static_cast<int>(u), 0x140, 0xF, 0xF, false)));
    const unsigned a = __builtin_amdgcn_readlane(static_cast<int>(u), 0);
    const unsigned b = __builtin_amdgcn_readlane(static_cast<int>(u), 16);
    const unsigned c = __builtin_amdgcn_readlane(static_cast<int>(u), 32);
    const unsigned d = __builtin_amdgcn_readlane(static_cast<int>(u), 48);
    return __uint_as_float(max(max(a, b), max(c, d)));
}
// Whether a view filtered with scale sc and largest |q| qmax is stored as it should be:
// qmax / sc within [2^6, 65504] (no half overflows, the largest values keep their place
// far above the subnormals), or a view of zeros.  NaN fails.
__device__ inline bool f16_scale_ok(float qmax, float sc) {
    const float r = qmax * (1.0f / sc);
    return qmax == 0.0f || (r >= 64.0f && r <= 65504.0f);
}

// Output of transform 2 (its conjugate, times the gain, is the filtered row pair):
// MODE 0 ring float32; 1 ring bfloat16; 2 ring float16 with a power-of-two scale per view
// from the bound max|q| <= max|x| * gain * l1 (vmax: max |x| of the view's raw rows, l1:
// sum_m |h(m)| of the tau = 1 response), stored in vscale by the thread given `first`;
// 6 ring float16 with the provisional scale sc6 (f16_prov_scale: max |x| of sampled rows),
// the largest |q| written kept in *qm; 7 ring float16 with the scale of the exact
// max |q| of the view (f16_view_qmax, gathered by the MODE 6 pass); 3 plain float32 rows
// [R][C]; 4 adjoint (the cosine weight after the self-adjoint filter, blocks [R][k][C]).
// Columns >= C are dropped.
template <int MODE> struct RowsOut {
    static constexpr bool F16 = MODE == 2 || MODE == 6 || MODE == 7;
    float* o0 = nullptr;
    float* o1 = nullptr;
    uint16_t* h0 = nullptr;
    uint16_t* h1 = nullptr;
    float* qm = nullptr;
    FilterView P;
    float b0, b1, g, inv_s = 1.0f;
    int C;
    bool has1;
    __device__ RowsOut(void* dst, int R, int k, int C_, int pitch, int64_t view_stride,
                       const FilterView& P_, int r0, int v, const unsigned* vmax, float l1,
                       float* vscale, bool first, float sc6 = 1.0f)
        : P(P_), C(C_) {
        const int r1 = r0 + 1;
        has1 = r1 < R;
        b0 = P.b0 + r0 * P.db;
        b1 = P.b0 + r1 * P.db;
        g = P.gain;
        if (F16) {
            const float sc = MODE == 2   ? f16_scale(__uint_as_float(vmax[v]) * g * l1)
                             : MODE == 6 ? sc6
                                         : f16_scale(f16_view_qmax(vmax, k, v));
            inv_s = 1.0f / sc;
            if (first) vscale[P.slot] = sc;
        }
        // uniform row bases (one per row of the pair), indexed by the column
        if (MODE == 0) {
            o0 = static_cast<float*>(dst) + static_cast<int64_t>(P.slot) * view_stride +
                 static_cast<int64_t>(r0 + 1) * pitch + 1;
            o1 = o0 + pitch;
        } else if (MODE == 1 || F16) {
            h0 = static_cast<uint16_t*>(dst) + static_cast<int64_t>(P.slot) * view_stride +
                 static_cast<int64_t>(r0 + 1) * pitch + 1;
            h1 = h0 + pitch;
        } else if (MODE == 3) {
            o0 = static_cast<float*>(dst) + static_cast<int64_t>(r0) * C;
            o1 = o0 + C;
        } else {
            o0 = static_cast<float*>(dst) + (static_cast<int64_t>(r0) * k + v) * C;
            o1 = static_cast<float*>(dst) + (static_cast<int64_t>(has1 ? r1 : r0) * k + v) * C;
        }
    }
    __device__ void operator()(int c, float2 z) const {
        if (c >= C) return;
        const float y0 = z.x * g, y1 = -z.y * g;
        if (MODE == 0 || MODE == 3) {
            o0[c] = y0;
            if (has1) o1[c] = y1;
        } else if (MODE == 1) {
            auto bf = [](float f) -> uint16_t {
                uint32_t u = __float_as_uint(f);
                u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
                return static_cast<uint16_t>(u >> 16);
            };
            h0[c] = bf(y0);
            if (has1) h1[c] = bf(y1);
        } else if (F16) {
            h0[c] = __half_as_ushort(__float2half_rn(y0 * inv_s));
            if (has1) h1[c] = __half_as_ushort(__float2half_rn(y1 * inv_s));
            // (a lone last row's partner is the filter of a zero row, rounding noise far
            // below the row's own values: one max3 per pair)
            if (MODE == 6) *qm = fmaxf(*qm, fmaxf(fabsf(y0), fabsf(y1)));
        } else {
            float w0 = 1.0f, w1 = 1.0f;
            if (P.cone) {
                const float a = P.a0 + c * P.da;
                const float aa = fmaf(a, a, P.sdd2);
                w0 = P.sdd * __frsqrt_rn(fmaf(b0, b0, aa));
                w1 = P.sdd * __frsqrt_rn(fmaf(b1, b1, aa));
            }
            o0[c] = y0 * w0;
            if (has1) o1[c] = y1 * w1;
        }
    }
};

// Four-step transforms (rows longer than one workgroup's LDS holds): N = 16 N2, index
// n = n1 N2 + n2, frequency k = k1 + 16 k2.  A slot is the N-point buffer of one row pair
// of one view in the four-step scratch, row k1 of the slot at offset k1 N2.
//
// Middle step (MODE 5 of filter_kernel, N2 points): row k1 of a slot holds
// Y[n2] = W_N^(n2 k1) sum_n1 z[n1 N2 + n2] W_16^(n1 k1); its N2-point transform is
// Z[k1 + 16 k2]; transform 2 takes conj(Z) H at those indices, and its output at m2 times
// W_N^(m2 k1) goes to the same row of the output scratch.
struct SlotIn {
    const float2* src;
    __device__ float2 operator()(int i) const { return src[i]; }
};
struct SlotOut {
    float2* dst;
    const float2* twn;     // N-point twiddles
    int k1;
    __device__ void operator()(int m2, float2 z) const { dst[m2] = cmul(z, twn[m2 * k1]); }
};

// Both transforms are forward Stockham FFTs through the LDS buffer; the second runs its
// passes in the reverse radix order, so that its first pass (radix 16, span 1) reads in
// each thread exactly the indices j + r N/16 the last pass of the first transform
// leaves in that thread's registers: the product with H and the conjugation are done
// there, with no exchange through LDS.  MODE 5: tw is the N2-point table, preceded by the
// 16 N2-point one; H has 16 N2 points.
//
// filter_view: row pair x of view v (MODE 0-4, 6, 7), or row x mod 16 of slot x / 16
// (MODE 5).
template <int LOGN, int MODE>
__device__ __forceinline__ void filter_view(
    const float* __restrict__ block, int R, int k, int C, const FilterView* __restrict__ fv,
    const float* __restrict__ H, const float2* __restrict__ tw, void* __restrict__ dst, int pitch,
    int64_t view_stride, unsigned* __restrict__ vmax, float l1, float* __restrict__ vscale,
    int x, int v) {
    constexpr int N = 1 << LOGN, FT = ft_of(LOGN);
    constexpr bool MID = MODE == 5;
    using P1 = FftPlan<LOGN, false>;
    using P2 = FftPlan<LOGN, true>;
    constexpr int NP = P1::NPASS;
    static_assert(P1::radix(NP - 1) == P2::radix(0), "fused passes share the radix");
    static_assert(!MID || NP > 1, "the four-step middle transform uses the LDS passes");
    // MODE 6: the provisional scale first, its loads in flight with the row loads
    const float sc6 = MODE == 6 ? f16_prov_scale(vmax, v) : 1.0f;
    __shared__ float2 lds[N];
    float2* const buf = lds;                         // captured by the pass lambdas
    const int k1 = MID ? (x & 15) : 0;
    const int64_t slot_row = (static_cast<int64_t>(x >> 4) * 16 + k1) * N;
    auto in_rows = [&] {
        if constexpr (MID)
            return SlotIn{reinterpret_cast<const float2*>(block) + slot_row};
        else
            return RowsIn<MODE>(block, R, k, C, fv[v], 2 * x, v);
    }();
    auto to_lds = [buf](int i, float2 z) { buf[sw(i)] = z; };
    constexpr int RF = P2::radix(0);                 // radix of the fused passes
    using PF = Pass<N, RF, FT>;
    typename PF::Regs vf;
    if constexpr (NP == 1) {
        PF::template load_dft<1, true>(tw, in_rows, vf);
    } else {
        {
            // transform 1 reads the row pair from global memory; its upper half (>= N/2
            // >= C) is zero padding and not read at all, except in the middle step
            constexpr int R0 = P1::radix(0);
            using P0 = Pass<N, R0, FT>;
            typename P0::Regs v0;
            P0::template load_dft<1, !MID>(tw, in_rows, v0);
            P0::template store<1>(to_lds, v0);
            __syncthreads();
            __builtin_amdgcn_sched_barrier(0);
        }
        lds_passes<P1, N, FT, 1, NP - 1>(tw, buf);
        PF::template load_dft<P1::span(NP - 1), false>(tw, [buf](int i) { return buf[sw(i)]; }, vf);
    }
    __builtin_amdgcn_sched_barrier(0);
    // transform 1's output index j + q N/RF is in vf[e][bitrev(q)]: transform 2's first
    // pass takes conj(Z) H at the same indices, r = q
    const int tf = PF::tid();
#pragma unroll
    for (int e = 0; e < PF::PER; ++e) {
        const int j = tf + e * FT;
        if (!PF::has(tf, e)) continue;
        float2 u[RF];
#pragma unroll
        for (int r = 0; r < RF; ++r) {
            const float2 z = vf[e][bitrev(r, RF)];
            const float h = MID ? H[k1 + 16 * (j + r * (N / RF))] : H[j + r * (N / RF)];
            u[r] = make_float2(z.x * h, -z.y * h);
        }
        dft_dif<RF>(u);
#pragma unroll
        for (int r = 0; r < RF; ++r) vf[e][r] = u[r];
    }
    auto out_rows = [&] {
        if constexpr (MID)
            return SlotOut{static_cast<float2*>(dst) + slot_row, tw - 16 * N, k1};
        else
            return RowsOut<MODE>(dst, R, k, C, pitch, view_stride, fv[v], 2 * x, v, vmax, l1,
                                 vscale, x == 0 && threadIdx.x == 0, sc6);
    }();
    float qm = 0.0f;                                 // MODE 6: largest |q| this thread writes
    if constexpr (MODE == 6) out_rows.qm = &qm;
    if constexpr (NP == 1) {
        PF::template store<1>(out_rows, vf);
    } else {
        // the other threads may still read transform 1's last buffer
        __syncthreads();
        PF::template store<1>(to_lds, vf);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        lds_passes<P2, N, FT, 1, NP - 1>(tw, buf);
        constexpr int RL = P2::radix(NP - 1);
        using PL = Pass<N, RL, FT>;
        typename PL::Regs vl;
        PL::template load_dft<P2::span(NP - 1), false>(tw, [buf](int i) { return buf[sw(i)]; }, vl);
        PL::template store<P2::span(NP - 1)>(out_rows, vl);
    }
    if constexpr (MODE == 6) {
        // the wave's largest |q|, one atomic per wave (non-negative floats order as their
        // bits; no barrier for a workgroup-wide reduction)
        qm = wave_max_nonneg(qm);
        if ((threadIdx.x & 63) == 0)
            atomicMax(f16_qmax_slot(vmax, k, v,
                                    (x * (FT / 64) + static_cast<int>(threadIdx.x >> 6)) % kF16QmaxSlots),
                      __float_as_uint(qm));
    }
}

// Grid (rows or slots, views) for MODE 0-6: block (x, y) filters row pair x of view y (or
// row x mod 16 of slot x / 16, MODE 5).  MODE 7 (the fix-up pass after MODE 6): grid (row
// pairs, 1); block x filters row pair x of every batch view whose provisional scale does
// not fit, and skips the others (usually all) after reading their maxima.
template <int LOGN, int MODE>
__global__ __launch_bounds__(ft_of(LOGN))
__attribute__((amdgpu_waves_per_eu(LOGN >= 12 ? 4 : 5))) void filter_kernel(
    const float* __restrict__ block, int R, int k, int C, const FilterView* __restrict__ fv,
    const float* __restrict__ H, const float2* __restrict__ tw, void* __restrict__ dst, int pitch,
    int64_t view_stride, unsigned* __restrict__ vmax, float l1,
    float* __restrict__ vscale) {
    const int x = blockIdx.x;
    if constexpr (MODE == 7) {
        // lane j of every wave checks view v0 + j; the ballot is the same in every wave
        for (int v0 = 0; v0 < k; v0 += 64) {
            const int v = v0 + static_cast<int>(threadIdx.x & 63);
            const bool bad = v < k && !f16_scale_ok(f16_view_qmax(vmax, k, v), f16_prov_scale(vmax, v));
            for (uint64_t m = __ballot(bad); m; m &= m - 1) {
                filter_view<LOGN, MODE>(block, R, k, C, fv, H, tw, dst, pitch, view_stride, vmax,
                                        l1, vscale, x, v0 + __builtin_ctzll(m));
                // the next view reuses the LDS buffer
                __syncthreads();
            }
        }
    } else {
        filter_view<LOGN, MODE>(block, R, k, C, fv, H, tw, dst, pitch, view_stride, vmax, l1,
                                vscale, x, blockIdx.y);
    }
}

// First and last steps of the four-step filter (N = 16 n2len), one thread per column of
// a slot, 16 points in registers.  Block (x, y, z): columns 256 x .., row pair p0 + y of
// view v0 + z, slot z gridDim.y + y.
//   OUT = false: the 16-point DFT over n1 of the input rows at columns n1 N2 + n2 (n1 >= 8
//     is the zero padding: C <= N/2), times W_N^(n2 k1), to row k1 of the slot in fs;
//   OUT = true: the 16-point DFT over k1 of column m2 of the slot in fs, output column
//     m1 N2 + m2 of the row pair.
// tw: the N-point twiddle table.
template <int MODE, bool OUT>
__global__ __launch_bounds__(256) void fourstep_cols_kernel(
    const float* __restrict__ block, int R, int k, int C, const FilterView* __restrict__ fv,
    const float2* __restrict__ tw, int n2len, float2* __restrict__ fs, void* __restrict__ dst,
    int pitch, int64_t view_stride, const unsigned* __restrict__ vmax, float l1,
    float* __restrict__ vscale, int p0, int v0) {
    const int n2 = blockIdx.x * 256 + threadIdx.x;
    const int r0 = 2 * (p0 + static_cast<int>(blockIdx.y));
    const int v = v0 + static_cast<int>(blockIdx.z);
    float2* slot = fs + (static_cast<int64_t>(blockIdx.z) * gridDim.y + blockIdx.y) * 16 * n2len;
    float2 x[16];
    if constexpr (!OUT) {
        const RowsIn<MODE> in(block, R, k, C, fv[v], r0, v);
#pragma unroll
        for (int n1 = 0; n1 < 8; ++n1) x[n1] = in(n1 * n2len + n2);
        dft_dif_upper_zero<16>(x);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int k1 = bitrev(q, 16);
            slot[static_cast<int64_t>(k1) * n2len + n2] = k1 ? cmul(x[q], tw[n2 * k1]) : x[q];
        }
    } else {
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) x[k1] = slot[static_cast<int64_t>(k1) * n2len + n2];
        dft_dif<16>(x);
        const bool first = p0 == 0 && blockIdx.y == 0 && blockIdx.x == 0 && threadIdx.x == 0;
        const RowsOut<MODE> out(dst, R, k, C, pitch, view_stride, fv[v], r0, v, vmax, l1, vscale,
                                first);
#pragma unroll
        for (int q = 0; q < 16; ++q) out(bitrev(q, 16) * n2len + n2, x[q]);
    }
}

// Max |x| over rows x rs, (x + G) rs, (x + 2G) rs, ... (below R) of view v of a raw block
// [R][k][C], reduced over the 256-thread workgroup; the result is in thread 0.  A row is
// read as 16-byte loads from its first 16-byte boundary (two in flight per thread), with
// the elements before it and after the last whole load read singly.
__device__ float rows_max_abs(const float* __restrict__ block, int R, int k, int C, int v, int x,
                              int G, int rs) {
    float m = 0.0f;
    const int ns = (R + rs - 1) / rs;                // rows 0, rs, 2 rs, ...
    // every row starts on a 16-byte boundary: the workgroup's rows as one index space of
    // (row, 16-byte load) pairs, four loads in flight per thread
    const int c4 = C >> 2;
    const int nr = x < ns ? (ns - 1 - x) / G + 1 : 0;
    const bool aligned = (C & 3) == 0 && (reinterpret_cast<uintptr_t>(block) & 15) == 0 &&
                         static_cast<int64_t>(nr) * c4 < (int64_t{1} << 31) - 1024;
    if (aligned) {
        const int total = nr * c4;
        auto at = [&](int j) {
            const int rr = j / c4, i = j - rr * c4;
            const int r = (x + rr * G) * rs;
            return reinterpret_cast<const float4*>(block + (static_cast<int64_t>(r) * k + v) * C)[i];
        };
        auto mx = [](float4 a) {
            return fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w)));
        };
        int j = threadIdx.x;
        for (; j + 768 < total; j += 1024) {
            const float4 a = at(j), b = at(j + 256), c = at(j + 512), d = at(j + 768);
            m = fmaxf(m, fmaxf(fmaxf(mx(a), mx(b)), fmaxf(mx(c), mx(d))));
        }
        for (; j < total; j += 256) m = fmaxf(m, mx(at(j)));
    }
    for (int ir = aligned ? ns : x; ir < ns; ir += G) {
        const int r = ir * rs;
        const float* src = block + (static_cast<int64_t>(r) * k + v) * C;
        const int head = static_cast<int>(((16 - (reinterpret_cast<uintptr_t>(src) & 15)) & 15) >> 2);
        const int h = head < C ? head : C;
        const int nq = (C - h) >> 2;                   // whole 16-byte loads
        const float4* q = reinterpret_cast<const float4*>(src + h);
        int i = threadIdx.x;
        for (; i + 256 < nq; i += 512) {
            const float4 a = q[i], b = q[i + 256];
            m = fmaxf(m, fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))),
                               fmaxf(fmaxf(fabsf(b.x), fabsf(b.y)), fmaxf(fabsf(b.z), fabsf(b.w)))));
        }
        for (; i < nq; i += 256) {
            const float4 a = q[i];
            m = fmaxf(m, fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w))));
        }
        // the elements before the first 16-byte boundary and after the last whole load
        const int t = threadIdx.x;
        if (t < h) m = fmaxf(m, fabsf(src[t]));
        const int tail = h + (nq << 2);
        if (t < C - tail) m = fmaxf(m, fabsf(src[tail + t]));
    }
    m = wave_max_nonneg(m);
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    return fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// Per-view max |x| of raw blocks [R][k][C] (the f16 scale bound of MODE 2) into vmax[v]
// (cleared before).  Workgroup (x, v) takes rows x, x + gridDim.x, ... of view v, so that
// each view gets gridDim.x (at most 128) atomics: one atomic per row would serialise ~R
// atomics on one address.
__global__ __launch_bounds__(256) void input_max_kernel(const float* __restrict__ block, int R,
                                                        int k, int C, unsigned* __restrict__ vmax) {
    const int v = blockIdx.y;
    const float m = rows_max_abs(block, R, k, C, v, blockIdx.x, gridDim.x, 1);
    if (threadIdx.x == 0) atomicMax(vmax + v, __float_as_uint(m));
}

// The sampled bound of the provisional f16 scale (MODE 6): grid (kF16SampleParts, k);
// workgroup (x, v) writes the bound max|x| gain l1 from the max |x| of sampled rows x,
// x + kF16SampleParts, ... (rows 0, rs, 2 rs, ...) of view v to vmax[kF16SampleParts v +
// x], a plain store of every entry (0 for none), and workgroup (0, v) clears the view's
// max |q| slots, which the MODE 6 pass raises.  No clearing launch, no atomics.
__global__ __launch_bounds__(256) void sample_max_kernel(const float* __restrict__ block, int R,
                                                         int k, int C, int rs,
                                                         const FilterView* __restrict__ fv,
                                                         float l1, unsigned* __restrict__ vmax) {
    const int v = blockIdx.y, x = blockIdx.x;
    const float m = rows_max_abs(block, R, k, C, v, x, kF16SampleParts, rs);
    if (threadIdx.x == 0) vmax[kF16SampleParts * v + x] = __float_as_uint(m * fv[v].gain * l1);
    if (x == 0 && threadIdx.x < kF16QmaxSlots) *f16_qmax_slot(vmax, k, v, threadIdx.x) = 0u;
}

// F16 storage of already-filtered views: power-of-two scale per view so that
// max |q| / scale <= 2^15, then round to nearest half.  The scale is stored per slot
// and folded into the back-projection weight.
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

// Four-step filter of rows with N = 2^logn > 2^kLdsMaxLogN points: the row pairs of the
// batch go through the scratch fs (two halves, each a whole number of N-point slots) in
// chunks of slots: column step, middle transform pair, column step.  tw: the N-point
// twiddle table followed by the N/16-point one.
template <int MODE>
int launch_fourstep(int logn, hipStream_t s, const float* block, int R, int k, int C,
                    const FilterView* fv, const float* H, const float2* tw, void* dst, int pitch,
                    int64_t view_stride, unsigned* vmax, float l1, float* vscale,
                    float2* fs, size_t fs_bytes) {
    const int64_t N = int64_t{1} << logn;
    const int n2len = static_cast<int>(N / 16);
    const int64_t slots = static_cast<int64_t>(fs_bytes / 2 / (sizeof(float2) * N));
    MR_CHECK(fs != nullptr && slots >= 1, MR_ERUNTIME, "four-step filter scratch missing");
    const int npairs = (R + 1) / 2;
    const int pc = static_cast<int>(std::min<int64_t>(npairs, slots));
    const int kv = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(k, slots / pc)));
    float2* fa = fs;
    float2* fb = fs + slots * N;
    const float2* tws = tw + N;
    const int lm = logn - 4;
    for (int v0 = 0; v0 < k; v0 += kv) {
        const int nv = std::min(kv, k - v0);
        for (int p0 = 0; p0 < npairs; p0 += pc) {
            const int np = std::min(pc, npairs - p0);
            const dim3 gc(n2len / 256, np, nv);
            hipLaunchKernelGGL((fourstep_cols_kernel<MODE, false>), gc, dim3(256), 0, s, block, R, k,
                               C, fv, tw, n2len, fa, dst, pitch, view_stride, vmax, l1, vscale, p0,
                               v0);
            const dim3 gm(16 * np * nv, 1);
            const float* mid_in = reinterpret_cast<const float*>(fa);
#define MR_MID_CASE(L)                                                                      \
    case L:                                                                                 \
        hipLaunchKernelGGL((filter_kernel<L, 5>), gm, dim3(ft_of(L)), 0, s, mid_in, 0, 0, 0, \
                           fv, H, tws, fb, 0, 0, vmax, 0.0f, vscale);                         \
        break;
            switch (lm) {
                MR_MID_CASE(11) MR_MID_CASE(12) MR_MID_CASE(13) MR_MID_CASE(14)
            default:
                return fail(MR_EVALUE, "four-step FFT length 2^" + std::to_string(logn));
            }
#undef MR_MID_CASE
            hipLaunchKernelGGL((fourstep_cols_kernel<MODE, true>), gc, dim3(256), 0, s, block, R, k,
                               C, fv, tw, n2len, fb, dst, pitch, view_stride, vmax, l1, vscale, p0,
                               v0);
        }
    }
    MR_HIP(hipGetLastError());
    return MR_OK;
}

template <int MODE>
int launch_mode(int logn, dim3 grid, hipStream_t s, const float* block, int R, int k, int C,
                const FilterView* fv, const float* H, const float2* tw, void* dst, int pitch,
                int64_t view_stride, unsigned* vmax = nullptr, float l1 = 0.0f,
                float* vscale = nullptr, float2* fs = nullptr, size_t fs_bytes = 0) {
    if (logn > kLdsMaxLogN && logn <= kMaxLogN) {
        if constexpr (MODE == 6 || MODE == 7)
            return fail(MR_EVALUE, "provisional f16 scales are for one-workgroup rows only");
        else
            return launch_fourstep<MODE>(logn, s, block, R, k, C, fv, H, tw, dst, pitch,
                                         view_stride, vmax, l1, vscale, fs, fs_bytes);
    }
#define MR_FILTER_CASE(L)                                                                  \
    case L:                                                                                \
        hipLaunchKernelGGL((filter_kernel<L, MODE>), grid, dim3(ft_of(L)), 0, s, block, R, k, C, \
                           fv, H, tw, dst, pitch, view_stride, vmax, l1, vscale);          \
        break;
    switch (logn) {
        MR_FILTER_CASE(1) MR_FILTER_CASE(2) MR_FILTER_CASE(3) MR_FILTER_CASE(4) MR_FILTER_CASE(5)
        MR_FILTER_CASE(6) MR_FILTER_CASE(7) MR_FILTER_CASE(8) MR_FILTER_CASE(9)
        MR_FILTER_CASE(10) MR_FILTER_CASE(11) MR_FILTER_CASE(12) MR_FILTER_CASE(13)
        MR_FILTER_CASE(14)
    default:
        return fail(MR_EVALUE, "FFT length 2^" + std::to_string(logn) +
                                   " outside the supported range 2..2^" + std::to_string(kMaxLogN));
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
                  const FilterView* fv, unsigned* vmax) {
    (void)n0;
    if (!fv) fv = c->d_fv;
    if (!vmax) vmax = c->d_vmax;
    const int logn = log2i(c->N);
    dim3 grid((c->R + 1) / 2, k);
    if (c->storage == MR_STORE_F32)
        return launch_mode<0>(logn, grid, s, d_block, c->R, k, c->C, fv, c->d_H, c->d_tw,
                              c->ring, c->pitch, c->view_stride, nullptr, 0.0f, nullptr, c->d_fs,
                              c->fs_bytes);
    if (c->storage == MR_STORE_BF16)
        return launch_mode<1>(logn, grid, s, d_block, c->R, k, c->C, fv, c->d_H, c->d_tw,
                              c->ring, c->pitch, c->view_stride, nullptr, 0.0f, nullptr, c->d_fs,
                              c->fs_bytes);
    // F16: the scale of a view must be fixed before the filter writes its halves.  The max
    // |x| of the raw rows bounds the filtered view (max |q| <= max |x| gain l1, the FDK
    // weight is at most 1).  Rows up to 2^kLdsMaxLogN / 2 columns: the max of about 96
    // sampled rows per view gives a provisional scale (sample_max_kernel, which also clears
    // the max |q| slots), the filter (MODE 6) writes with it and gathers the exact max |q|
    // of each view, and a second launch (MODE 7) filters again, with the scale of that max,
    // only the views whose largest value would overflow a half or sit below 2^6; its
    // workgroups read the maxima of every view and return.  Since any power-of-two scale
    // stores the same products there, the stored views equal those of a scale from the
    // full max, and the full read of the batch for the bound (0.11 ms of a 0.87 ms batch of
    // 40 full-size views) is gone.  Wider rows (four-step transform) read the whole batch
    // for the bound first (MODE 2, vmax[v] cleared by a kernel so that the sequence can be
    // graph-captured).
    const float l1 = static_cast<float>(c->h_l1);
    if (logn > kLdsMaxLogN) {
        hipLaunchKernelGGL(zero_u32_kernel, dim3((k + 255) / 256), dim3(256), 0, s, vmax, k);
        hipLaunchKernelGGL(input_max_kernel, dim3(std::min(c->R, 128), k), dim3(256), 0, s, d_block,
                           c->R, k, c->C, vmax);
        MR_HIP(hipGetLastError());
        return launch_fourstep<2>(logn, s, d_block, c->R, k, c->C, fv, c->d_H, c->d_tw, c->ring,
                                  c->pitch, c->view_stride, vmax, l1, c->d_vscale, c->d_fs,
                                  c->fs_bytes);
    }
    const int rs = (c->R + kF16SampleRows - 1) / kF16SampleRows;
    hipLaunchKernelGGL(sample_max_kernel, dim3(kF16SampleParts, k), dim3(256), 0, s, d_block, c->R,
                       k, c->C, rs, fv, l1, vmax);
    MR_HIP(hipGetLastError());
    int rc = launch_mode<6>(logn, grid, s, d_block, c->R, k, c->C, fv, c->d_H, c->d_tw, c->ring,
                            c->pitch, c->view_stride, vmax, l1, c->d_vscale);
    if (rc) return rc;
    return launch_mode<7>(logn, dim3(grid.x, 1), s, d_block, c->R, k, c->C, fv, c->d_H, c->d_tw,
                          c->ring, c->pitch, c->view_stride, vmax, l1, c->d_vscale);
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
                          d_graw, 0, 0, nullptr, 0.0f, nullptr, c->d_fs, c->fs_bytes);
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
                             float2* d_fs, size_t fs_bytes, hipStream_t s) {
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
                          nullptr, 0.0f, nullptr, d_fs, fs_bytes);
}

// Two halves of whole N-point slots: a slot per row pair, for up to 16 views of a batch,
// at most 256 MiB a half (at least one slot).
size_t fft_scratch_bytes(int N, int pairs) {
    if (N <= (1 << kLdsMaxLogN)) return 0;
    const size_t slot = sizeof(float2) * static_cast<size_t>(N);
    size_t half = slot * static_cast<size_t>(std::max(pairs, 1)) * 16;
    half = std::min(half, size_t{256} << 20);
    half = std::max(slot, half / slot * slot);
    return 2 * half;
}

}  // namespace mr
