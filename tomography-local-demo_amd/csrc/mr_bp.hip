// Voxel-driven FDK / FBP back-projection kernels for gfx950.
//
// Replaces bp_cone_numba (recon.py:193-219) and bp_parallel_numba (recon.py:175-191)
// with the per-tap zero-outside bilinear of _bil_nb (recon.py:156-173).
//
// Ring layout: view slot v (arrival order) holds rows -1..R+1 and columns -1..C+1 of
// its filtered projection at ring[v*view_stride + (r+1)*pitch + (c+1)]; the border
// is zero.  Clamping a detector coordinate to [-1, C] (rows [-1, R]) therefore
// reproduces "each tap outside the detector contributes 0" exactly: a clamped
// coordinate lands on a zero tap with weight 1 and a second tap with weight 0.
//
// Tile kernel: a 256-thread workgroup owns a 32 x 32 tile of an affine point grid
// (slice plane, or one plane of a volume) and accumulates all views in FP32
// registers, 4 points per thread.  Per view, one thread derives in FP64 the tile's
// projective coefficients *relative to an integer detector reference (row, col)
// near the tile centre*, so the FP32 inner loop works with small relative
// coordinates (|col - ref| < ~100 px, ulp < 1e-5 px) instead of absolute columns
// near 1000 px (ulp 6e-5 px), which would dominate the error budget.
#include <type_traits>

#include "mr_device.h"
#include "mr_internal.h"

namespace mr {

namespace {

template <typename ST> struct Store;
template <> struct Store<float> {
    __device__ static inline float ld(const float* __restrict__ p, uint32_t i) { return p[i]; }
};
template <> struct Store<__half> {
    __device__ static inline float ld(const __half* __restrict__ p, uint32_t i) {
        return __half2float(p[i]);
    }
};
template <> struct Store<uint16_t> {   // bfloat16 bits
    __device__ static inline float ld(const uint16_t* __restrict__ p, uint32_t i) {
        return __uint_as_float(static_cast<uint32_t>(p[i]) << 16);
    }
};

__device__ inline double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// FP64 record of one (tile, view): coefficients of uh, vh, wh as affine functions of
// the in-tile offsets (ds along the fast axis b, dp along the slow axis a), shifted
// so that uh/wh, vh/wh are columns/rows relative to the reference (cref, rref);
// clamp bounds of the relative coordinates; per-view storage scale; base index.
struct RecRef {
    double cref, rref;                        // integer detector reference (col, row)
    double U0, Ua, Ub, V0, Va, Vb, W0, Wa, Wb;  // float64 relative coefficients (unscaled)
};

__device__ void make_record(const ViewGeom* __restrict__ gp, float vs, const mr_task& T,
                            const double* xt, int R, int C, int pitch, bool cone,
                            float* __restrict__ rec, RecRef* ref = nullptr, bool band = false) {
    const ViewGeom g = *gp;
    const double sh = T.cor_shift;
    const double W0 = dot3(g.P[2], xt) + g.P[2][3];
    const double Wa = dot3(g.P[2], T.b);
    const double Wb = dot3(g.P[2], T.a);
    double U0 = dot3(g.P[0], xt) + g.P[0][3] + sh * W0;
    double Ua = dot3(g.P[0], T.b) + sh * Wa;
    double Ub = dot3(g.P[0], T.a) + sh * Wb;
    const double svv = sh * g.uvv;
    double V0 = dot3(g.P[1], xt) + g.P[1][3] + svv * W0;
    double Va = dot3(g.P[1], T.b) + svv * Wa;
    double Vb = dot3(g.P[1], T.a) + svv * Wb;
    const double hs = 0.5 * (BP_TS - 1), hp = 0.5 * (BP_TP - 1);
    const double wc = W0 + hs * Wa + hp * Wb;
    const double cc = 0.5 * (C - 1), cr = 0.5 * (R - 1);
    double colc = (U0 + hs * Ua + hp * Ub) / wc + cc;
    double rowc = (V0 + hs * Va + hp * Vb) / wc + cr;
    colc = fmin(fmax(colc, -static_cast<double>(C)), 2.0 * C);   // NaN -> -C
    rowc = fmin(fmax(rowc, -static_cast<double>(R)), 2.0 * R);
    const double cref = floor(colc);
    // band ring: the reference row is the last row before a band start (padded row
    // rref + 1 = 3 k), so a point whose row is r rows past it (r = floor(row)) lies in
    // band k + floor(r / 3) at row r mod 3 of the band
    const double kb = floor((rowc + 1.0) / 3.0);
    const double rref = band ? 3.0 * kb - 1.0 : floor(rowc);
    const double dc = cc - cref, dr = cr - rref;
    U0 += dc * W0; Ua += dc * Wa; Ub += dc * Wb;
    V0 += dr * W0; Va += dr * Wa; Vb += dr * Wb;
    if (ref) *ref = RecRef{cref, rref, U0, Ua, Ub, V0, Va, Vb, W0, Wa, Wb};
    // cone: the weight 1/wh^2 absorbs the view's storage scale
    const double f = cone ? 1.0 / sqrt(static_cast<double>(vs)) : 1.0;
    float4* r4 = reinterpret_cast<float4*>(rec);
    r4[0] = make_float4(static_cast<float>(U0 * f), static_cast<float>(Ua * f),
                        static_cast<float>(Ub * f), static_cast<float>(V0 * f));
    r4[1] = make_float4(static_cast<float>(Va * f), static_cast<float>(Vb * f),
                        static_cast<float>(W0 * f), static_cast<float>(Wa * f));
    // [14]: element offset of the padded reference position; for the band ring the byte
    // offset of (band k, padded reference column)
    const int base = band ? static_cast<int>(kb) * (pitch << 3) + ((static_cast<int>(cref) + 1) << 3)
                          : (static_cast<int>(rref) + 1) * pitch + (static_cast<int>(cref) + 1);
    r4[2] = make_float4(static_cast<float>(Wb * f), static_cast<float>(-1.0 - cref),
                        static_cast<float>(C - cref), static_cast<float>(-1.0 - rref));
    r4[3] = make_float4(static_cast<float>(R - rref), vs, __int_as_float(base),
                        static_cast<float>(g.uvv));
}

// ---------------------------------------------------------------- v2 inner loop
// FP64 records of each (tile, view) in LDS, the FP32 interpolation of every point of the
// tile in registers; the gathers go through a
// buffer resource per view (base = the view's first element, range = the view), so
// the per-point address is one 24-bit multiply-add and one shift-add of absolute
// byte offsets, the second detector row comes from the SGPR offset, the +1 column
// from the immediate offset, and an out-of-range offset can only read zero.

// Band ring addressing.  Element (band b, column c) of a view holds the padded rows
// 3b..3b+3 of column c (8 bytes, element c+1 follows); a band record's [14] is the byte
// offset of (band k, reference column), its reference row placed so that padded row
// ref + 1 = 3k.  A point r rows past the reference (r = floor(row)) and c columns past it
// needs rows r, r+1 of band k + floor(r/3), i.e. rows j, j+1 (j = r mod 3) of element
// (k + floor(r/3), c) and of element c+1: one 16-byte load at the element's 8-byte
// aligned offset, then a 16j-bit shift of each column's 8 bytes (BandP).  With
// q = floor((r + 3K)/3) = floor(r/3) + K (K = 2^20, so r + 3K >= 0 for every clamped
// row), the byte offset is [14] + (q - K) pitchb + 8c, formed
// modulo 2^32 by signed 24-bit multiply-adds (|c| < 2^23, q < 2^23, pitchb < 2^23 as
// the band ring serves pitches up to 2^20 columns), and 16j = 16r - 48q modulo 64,
// the bits a 64-bit shift reads (48K is a multiple of 64).
constexpr int BAND_K = 1 << 20;
// q = floor((r + 3K) / 3): 0x55555556 = (2^32 + 2) / 3, so the high word of the product
// is exact for 0 <= r + 3K < 2^31
__device__ inline int band_q(int r) {
    return static_cast<int>(__umulhi(static_cast<uint32_t>(r + 3 * BAND_K), 0x55555556u));
}
// [14] - K pitchb, the per-view base of band_addr (wrapping 32-bit arithmetic)
__device__ inline int band_base(int ref14, int pitchb) {
    return static_cast<int>(static_cast<uint32_t>(ref14) -
                            static_cast<uint32_t>(BAND_K) * static_cast<uint32_t>(pitchb));
}
// Byte offset of element (k + floor(r/3), c) for q = band_q(r), basek = band_base([14],
// pitchb): two signed 24-bit multiply-adds (the compiler would otherwise form a
// multiply, a shift and a three-input add).
__device__ inline int band_addr(int q, int c, int pitchb, int basek) {
    int x, o;
    asm("v_mad_i32_i24 %0, %1, 8, %2" : "=v"(x) : "v"(c), "s"(basek));
    asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(o) : "v"(q), "s"(pitchb), "v"(x));
    return o;
}
// 16 (r mod 3) modulo 64, the shift that selects rows r, r+1 of the element's four
__device__ inline int band_shift(int q, int r) {
    return __mul24(q, -48) + (r << 4);
}


template <typename ST> struct Taps;
template <> struct Taps<float> {
    static constexpr int LOG_ESZ = 2;
    __device__ static inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int soff1, float& a,
                                       float& b, float& c, float& d) {
        // the builtin returns the dword's bits
        a = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff, 0, 0));
        b = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff + 4, 0, 0));
        c = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff1, 0));
        d = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff + 4, soff1, 0));
    }
};
// 16-bit formats: one dword per detector row holds columns c0 (low half) and c0 + 1
// (high half); it is read at 2-byte alignment (the queues run in unaligned mode).
template <> struct Taps<__half> {
    static constexpr int LOG_ESZ = 1;
    __device__ static inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int soff1, float& a,
                                       float& b, float& c, float& d) {
        const uint32_t w0 = __builtin_amdgcn_raw_buffer_load_b32(rs, voff, 0, 0);
        const uint32_t w1 = __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff1, 0);
        a = __half2float(__ushort_as_half(static_cast<unsigned short>(w0 & 0xffffu)));
        b = __half2float(__ushort_as_half(static_cast<unsigned short>(w0 >> 16)));
        c = __half2float(__ushort_as_half(static_cast<unsigned short>(w1 & 0xffffu)));
        d = __half2float(__ushort_as_half(static_cast<unsigned short>(w1 >> 16)));
    }
};
template <> struct Taps<uint16_t> {
    static constexpr int LOG_ESZ = 1;
    __device__ static inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int soff1, float& a,
                                       float& b, float& c, float& d) {
        const uint32_t w0 = __builtin_amdgcn_raw_buffer_load_b32(rs, voff, 0, 0);
        const uint32_t w1 = __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff1, 0);
        a = __uint_as_float(w0 << 16);
        b = __uint_as_float(w0 & 0xffff0000u);
        c = __uint_as_float(w1 << 16);
        d = __uint_as_float(w1 & 0xffff0000u);
    }
};


// Taps of one point as loaded by kernels 2 and 4: f32 keeps four floats; the 16-bit formats keep
// the two row dwords packed until the lerp, where f16 halves feed v_fma_mix_f32
// directly (no separate conversions).
template <typename ST> struct Pair;
template <> struct Pair<float> {
    float a, b, c, d;
    __device__ inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int soff1) {
        Taps<float>::load(rs, voff, soff1, a, b, c, d);
    }
    __device__ inline void lerp(float fc, float& top, float& bot) const {
        top = fmaf(fc, b - a, a);
        bot = fmaf(fc, d - c, c);
    }
};
__device__ inline float half_lo(uint32_t w) {
    return __half2float(__ushort_as_half(static_cast<unsigned short>(w & 0xffffu)));
}
__device__ inline float half_hi(uint32_t w) {
    return __half2float(__ushort_as_half(static_cast<unsigned short>(w >> 16)));
}
// 16-bit taps at odd columns are not dword-aligned: both rows are read as 8 aligned
// bytes and the dword holding columns c, c+1 is taken by a funnel shift.
__device__ inline void load_pair16(__amdgpu_buffer_rsrc_t rs, int voff, int soff1, uint32_t& w0,
                                   uint32_t& w1) {
    const int a = voff & ~3;
    const uint32_t sh = static_cast<uint32_t>(voff & 2) << 3;
    const uint2 x0 = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(rs, a, 0, 0));
    const uint2 x1 = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(rs, a, soff1, 0));
    w0 = __builtin_amdgcn_alignbit(x0.y, x0.x, sh);
    w1 = __builtin_amdgcn_alignbit(x1.y, x1.x, sh);
}
template <> struct Pair<__half> {
    uint32_t w0, w1;
    __device__ inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int soff1) {
        load_pair16(rs, voff, soff1, w0, w1);
    }
    __device__ inline void lerp(float fc, float& top, float& bot) const {
        top = fmaf(fc, half_hi(w0), fmaf(-fc, half_lo(w0), half_lo(w0)));
        bot = fmaf(fc, half_hi(w1), fmaf(-fc, half_lo(w1), half_lo(w1)));
    }
};
template <> struct Pair<uint16_t> {
    uint32_t w0, w1;
    __device__ inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int soff1) {
        load_pair16(rs, voff, soff1, w0, w1);
    }
    __device__ inline void lerp(float fc, float& top, float& bot) const {
        const float a = __uint_as_float(w0 << 16), b = __uint_as_float(w0 & 0xffff0000u);
        const float c = __uint_as_float(w1 << 16), d = __uint_as_float(w1 & 0xffff0000u);
        top = fmaf(fc, b - a, a);
        bot = fmaf(fc, d - c, c);
    }
};

// Taps from the row-pair ring: one element of pair-row r holds (q(r, c), q(r+1, c)) side
// by side, so the four taps of a point are contiguous — 8 bytes (16-bit forms, one
// buffer_load_dwordx2) or 16 bytes (f32, one buffer_load_dwordx4) — and a gather costs
// one instruction instead of two.
template <typename ST> struct PairP;
template <> struct PairP<float> {
    float a, b, c, d;
    __device__ inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int) {
        const uint4 x = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 0, 0));
        a = __uint_as_float(x.x);      // q(r, c)
        c = __uint_as_float(x.y);      // q(r+1, c)
        b = __uint_as_float(x.z);      // q(r, c+1)
        d = __uint_as_float(x.w);      // q(r+1, c+1)
    }
    __device__ inline void lerp(float fc, float& top, float& bot) const {
        top = fmaf(fc, b - a, a);
        bot = fmaf(fc, d - c, c);
    }
};
template <> struct PairP<__half> {
    uint32_t w0, w1;                   // (q(r, c) | q(r+1, c) << 16), same for c+1
    __device__ inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int) {
        const uint2 x = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, 0, 0));
        w0 = x.x;
        w1 = x.y;
    }
    __device__ inline void lerp(float fc, float& top, float& bot) const {
        top = fmaf(fc, half_lo(w1), fmaf(-fc, half_lo(w0), half_lo(w0)));
        bot = fmaf(fc, half_hi(w1), fmaf(-fc, half_hi(w0), half_hi(w0)));
    }
};
// Bilinear value of f16 row-pair taps in packed half arithmetic: the column lerps of both
// rows in one v_pk_fma_f16 (after one v_pk_add_f16 for the differences), the row
// difference as one f16 subtraction, the row lerp in f32 from the f16 operands
// (v_fma_mix_f32): 4 VALU against 6 with f32 lerps.  The lerps round to f16 (relative
// 2^-11, the order of the storage rounding itself); fc2 is the column fraction as f16 in
// both halves.
typedef _Float16 half2v __attribute__((ext_vector_type(2)));
__device__ inline float pair_bilerp_f16(uint32_t w0, uint32_t w1, half2v fc2, float fr) {
    const half2v x = __builtin_bit_cast(half2v, w0);      // (q(r, c), q(r+1, c))
    const half2v y = __builtin_bit_cast(half2v, w1);      // (q(r, c+1), q(r+1, c+1))
    const half2v tb = __builtin_elementwise_fma(fc2, y - x, x);   // (top, bottom)
    const _Float16 d = tb.y - tb.x;
    return fmaf(fr, static_cast<float>(d), static_cast<float>(tb.x));
}
template <> struct PairP<uint16_t> {
    uint32_t w0, w1;
    __device__ inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int) {
        const uint2 x = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(rs, voff, 0, 0));
        w0 = x.x;
        w1 = x.y;
    }
    __device__ inline void lerp(float fc, float& top, float& bot) const {
        const float a = __uint_as_float(w0 << 16), b = __uint_as_float(w1 << 16);
        const float c = __uint_as_float(w0 & 0xffff0000u), d = __uint_as_float(w1 & 0xffff0000u);
        top = fmaf(fc, b - a, a);
        bot = fmaf(fc, d - c, c);
    }
};

// Taps from the band ring (16-bit forms): an element of band k holds the padded rows
// 3k..3k+3 of one column (8 bytes), so the 2x2 stencil of every point lies in one band
// and one buffer_load_dwordx4 fetches columns c and c+1; the two rows the point needs,
// 3k+j and 3k+j+1, are picked at the lerp by a 16j-bit funnel shift.  (A 12-byte load at
// the 2-byte-aligned offset of the two rows would need no shift, but costs the texture
// path far more: 5.6 against 3.2 ms for the full_cone f16 update.)
template <typename ST> struct BandP {
    uint32_t d0, d1, d2, d3;           // (rows 0,1 | rows 2,3) of column c, then of c+1
    uint32_t sh;                       // 16 j (modulo 64)
    __device__ inline void load(__amdgpu_buffer_rsrc_t rs, int voff, int shift) {
        const uint4 x = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, 0, 0));
        d0 = x.x;
        d1 = x.y;
        d2 = x.z;
        d3 = x.w;
        sh = static_cast<uint32_t>(shift);
    }
    __device__ inline void lerp(float fc, float& top, float& bot) const {
        PairP<ST> p;
        p.w0 = static_cast<uint32_t>(((static_cast<uint64_t>(d1) << 32) | d0) >> (sh & 63u));
        p.w1 = static_cast<uint32_t>(((static_cast<uint64_t>(d3) << 32) | d2) >> (sh & 63u));
        p.lerp(fc, top, bot);
    }
};

// RK: the ring copy read, with view_stride its elements per view.  0: the ring itself
// (two loads per point-view); 1: the row-pair copy (one load of the four taps, PairP);
// 2 (16-bit forms): the band copy, with band records (one 16-byte load, BandP).
// At least 5 waves per SIMD (89-95 VGPRs instead of 93-111, no spills; 6 would spill).
template <bool CONE, typename ST, int RK = 0>
__global__ __launch_bounds__(BP_THREADS, 5) void bp_tiles_kernel2(
    const ST* __restrict__ ring, int64_t view_stride, int pitch, int R, int C,
    const ViewGeom* __restrict__ geom, const float* __restrict__ vscale,
    const mr_task* __restrict__ tasks, const Tile* __restrict__ tiles, int n_views_arg,
    float scale_arg, float* __restrict__ out, int* __restrict__ qhead,
    const int* __restrict__ qcount, const int* __restrict__ n_views_dev, int xcd,
    const int* __restrict__ v_begin_dev, float* __restrict__ accum, int64_t accum_half) {
    __shared__ float4 rec4[BP_VPASS * BP_REC / 4];
    __shared__ int s_ti;
    float* rec = reinterpret_cast<float*>(rec4);
    constexpr int LE = Taps<ST>::LOG_ESZ;
    // persistent mode: tiles come from a device queue (qhead != null) and the view
    // count may come from device memory, so that a captured graph can be replayed
    const int n_views = n_views_dev ? *n_views_dev : n_views_arg;
    const float scale = n_views_dev ? (n_views > 0 ? static_cast<float>(3.141592653589793 / n_views) : 0.0f)
                                    : scale_arg;
    // running sums: resume from accum and add views [v_first, n_views) only (none when
    // v_first >= n_views); the FP32 sums then equal one pass over all views bit for bit,
    // as each point adds its views in the same order either way
    const int v_first = accum ? *v_begin_dev : 0;
    const int ntile = qhead ? *qcount : 0;
    for (int it = 0;; ++it) {
    int ti;
    if (qhead) {
        __syncthreads();
        if (threadIdx.x == 0) s_ti = atomicAdd(qhead, 1);
        __syncthreads();
        ti = s_ti;
        if (ti >= ntile) break;
    } else {
        if (it > 0) break;
        ti = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    }
    const Tile tl = tiles[ti];
    const mr_task& T = tasks[tl.task];
    const int t = threadIdx.x;
    const int tx = t & (BP_TS - 1);
    const int ty = t / BP_TS;
    constexpr int PSTEP = BP_THREADS / BP_TS;
    constexpr int PPT = BP_TP / PSTEP;

    const int p0 = tl.tp * BP_TP, s0 = tl.ts * BP_TS;
    double xt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d)
        xt[d] = T.o[d] + tl.q * T.c[d] + p0 * T.a[d] + s0 * T.b[d];

    const int s = s0 + tx;
    bool valid[PPT], inside[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        const int p = p0 + ty + PSTEP * m;
        valid[m] = (p < T.np) && (s < T.ns);
        inside[m] = valid[m] && (!T.mask_box || in_box(T, p, s));
    }

    const float ds = static_cast<float>(tx);
    const float dp0 = static_cast<float>(ty);
    static_assert(RK != 2 || sizeof(ST) == 2, "band copy of 16-bit storage only");
    constexpr bool BAND = RK == 2;
    constexpr int ESH = RK == 1 ? LE + 1 : LE;              // log2 bytes per ring element
    const int pitchb = BAND ? pitch << 3 : pitch << ESH;    // row (band) pitch in bytes
    const int view_bytes = static_cast<int>(view_stride << LE);
    float acc[PPT], accb[PPT], fr_par[PPT];
    int roff_par[PPT], jsh_par[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        acc[m] = 0.0f; accb[m] = 0.0f; fr_par[m] = 0.0f; roff_par[m] = 0; jsh_par[m] = 0;
    }
    const int64_t obase = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q;
    auto oidx = [&](int m) {
        return obase + static_cast<int64_t>(p0 + ty + PSTEP * m) * T.stride_p +
               static_cast<int64_t>(s) * T.stride_s;
    };
    if (accum) {
#pragma unroll
        for (int m = 0; m < PPT; ++m) {
            if (!valid[m]) continue;
            acc[m] = accum[oidx(m)];
            if (!CONE) accb[m] = accum[accum_half + oidx(m)];
        }
    }
    if (!CONE && n_views > 0) {
        // parallel beam: the detector row of a point is the same for every view; take it
        // from the record of the first view summed here (view 0 when none is)
        const int v0 = v_first < n_views ? v_first : 0;
        __syncthreads();
        if (t == 0) make_record(geom + v0, vscale[v0], T, xt, R, C, pitch, CONE, rec, nullptr, BAND);
        __syncthreads();
        const float4 r0 = rec4[0], r1 = rec4[1], r2 = rec4[2], r3 = rec4[3];
#pragma unroll
        for (int m = 0; m < PPT; ++m) {
            const float dp = dp0 + static_cast<float>(PSTEP * m);
            float row = fmaf(dp, r1.y, fmaf(ds, r1.x, r0.w));
            row = __builtin_amdgcn_fmed3f(row, r2.w, r3.x);
            fr_par[m] = __builtin_amdgcn_fractf(row);
            if (BAND) {
                const int r = cvt_flr(row), q = band_q(r);
                roff_par[m] = __mul24(q, pitchb);
                jsh_par[m] = band_shift(q, r);
            } else {
                roff_par[m] = cvt_flr(row) * pitchb;
            }
        }
    }

    for (int vb = v_first; vb < n_views; vb += BP_VPASS) {
        const int nv = min(BP_VPASS, n_views - vb);
        __syncthreads();
        if (t < nv)
            make_record(geom + vb + t, vscale[vb + t], T, xt, R, C, pitch, CONE, rec + t * BP_REC,
                        nullptr, BAND);
        __syncthreads();
#pragma unroll 2
        for (int v = 0; v < nv; ++v) {
            const float4 r0 = rec4[v * 4 + 0];
            const float4 r1 = rec4[v * 4 + 1];
            const float4 r2 = rec4[v * 4 + 2];
            const float4 r3 = rec4[v * 4 + 3];
            const int baseb = BAND ? band_base(__float_as_int(r3.z), pitchb) : __float_as_int(r3.z) << ESH;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<ST*>(ring + static_cast<int64_t>(vb + v) * view_stride), 0, view_bytes,
                0x00020000);
            if (CONE) {
                const float uh0 = fmaf(dp0, r0.z, fmaf(ds, r0.y, r0.x));
                const float vh0 = fmaf(dp0, r1.y, fmaf(ds, r1.x, r0.w));
                const float wh0 = fmaf(dp0, r2.x, fmaf(ds, r1.w, r1.z));
#pragma unroll
                for (int m = 0; m < PPT; ++m) {
                    const float dm = static_cast<float>(PSTEP * m);
                    const float uh = m ? fmaf(dm, r0.z, uh0) : uh0;
                    const float vh = m ? fmaf(dm, r1.y, vh0) : vh0;
                    const float wh = m ? fmaf(dm, r2.x, wh0) : wh0;
                    const float rw = __builtin_amdgcn_rcpf(wh);
                    const float col = __builtin_amdgcn_fmed3f(uh * rw, r2.y, r2.z);
                    const float row = __builtin_amdgcn_fmed3f(vh * rw, r2.w, r3.x);
                    const float wt = rw * rw;
                    const float fc = __builtin_amdgcn_fractf(col);
                    const float fr = __builtin_amdgcn_fractf(row);
                    float top, bot;
                    if constexpr (BAND) {
                        const int r = cvt_flr(row), q = band_q(r);
                        BandP<ST> tap;
                        tap.load(rs, band_addr(q, cvt_flr(col), pitchb, baseb), band_shift(q, r));
                        tap.lerp(fc, top, bot);
                    } else if constexpr (RK == 1) {
                        PairP<ST> tap;
                        tap.load(rs, __mul24(cvt_flr(row), pitchb) + ((cvt_flr(col) << ESH) + baseb), 0);
                        tap.lerp(fc, top, bot);
                    } else {
                        const int voff = __mul24(cvt_flr(row), pitchb) + ((cvt_flr(col) << LE) + baseb);
                        Pair<ST> tap;              // 16-bit taps: lerps on the packed halves
                        tap.load(rs, voff, pitchb);
                        tap.lerp(fc, top, bot);
                    }
                    acc[m] = fmaf(wt, fmaf(fr, bot - top, top), acc[m]);
                }
            } else {
                const float vs = r3.y;
                const float uh0 = fmaf(dp0, r0.z, fmaf(ds, r0.y, r0.x));
#pragma unroll
                for (int m = 0; m < PPT; ++m) {
                    const float dm = static_cast<float>(PSTEP * m);
                    const float col = __builtin_amdgcn_fmed3f(m ? fmaf(dm, r0.z, uh0) : uh0, r2.y, r2.z);
                    const float fc = __builtin_amdgcn_fractf(col);
                    float top, bot;
                    if constexpr (BAND) {
                        BandP<ST> tap;
                        tap.load(rs, roff_par[m] + ((cvt_flr(col) << 3) + baseb), jsh_par[m]);
                        tap.lerp(fc, top, bot);
                    } else if constexpr (RK == 1) {
                        PairP<ST> tap;
                        tap.load(rs, roff_par[m] + ((cvt_flr(col) << ESH) + baseb), 0);
                        tap.lerp(fc, top, bot);
                    } else {
                        const int voff = roff_par[m] + ((cvt_flr(col) << LE) + baseb);
                        Pair<ST> tap;
                        tap.load(rs, voff, pitchb);
                        tap.lerp(fc, top, bot);
                    }
                    acc[m] = fmaf(vs, top, acc[m]);
                    accb[m] = fmaf(vs, bot, accb[m]);
                }
            }
        }
    }

#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        if (!valid[m]) continue;
        if (accum) {
            accum[oidx(m)] = acc[m];
            if (!CONE) accum[accum_half + oidx(m)] = accb[m];
        }
        float val = acc[m];
        if (!CONE) val = fmaf(fr_par[m], accb[m] - acc[m], acc[m]);
        out[oidx(m)] = inside[m] ? val * scale : 0.0f;
    }
    }   // tile loop
}

// ---------------------------------------------------------------- v4: precomputed records
// The float64 per-(tile, view) records are computed by their own kernel into global
// memory ([tile][view][16] floats), so the back-projection kernel reads them with
// scalar loads (uniform per wave): no LDS, no barriers, no float64 in the hot loop,
// fewer VGPRs and more waves.  The view loop is software-pipelined: the four taps of
// view v+1 are requested before view v is interpolated, doubling the gathers in flight
// per wave.  PPT points per thread (rows ty + STEP*m of a 32 x 32 tile).
__global__ __launch_bounds__(256) void records_kernel(const ViewGeom* __restrict__ geom,
                                                      const float* __restrict__ vscale,
                                                      const mr_task* __restrict__ tasks,
                                                      const Tile* __restrict__ tiles, int ntiles,
                                                      int n_views, int R, int C, int pitch,
                                                      int cone, float* __restrict__ recs, int band) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= static_cast<int64_t>(ntiles) * n_views) return;
    const int ti = static_cast<int>(gid / n_views);
    const int v = static_cast<int>(gid - static_cast<int64_t>(ti) * n_views);
    const Tile tl = tiles[ti];
    const mr_task& T = tasks[tl.task];
    const int p0 = tl.tp * BP_TP, s0 = tl.ts * BP_TS;
    double xt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xt[d] = T.o[d] + tl.q * T.c[d] + p0 * T.a[d] + s0 * T.b[d];
    make_record(geom + v, vscale ? vscale[v] : 1.0f, T, xt, R, C, pitch, cone != 0,
                recs + gid * BP_REC, nullptr, band != 0);
}

// RING: 0 the ring itself, 1 the row-pair ring, 2 the band ring (16-bit forms).
template <typename ST, int PPT, int RING = 0>
struct PointState {
    float fc[PPT], fr[PPT], wt[PPT];
    typename std::conditional<RING == 2, BandP<ST>,
                              typename std::conditional<RING == 1, PairP<ST>, Pair<ST>>::type>::type tap[PPT];
};

// Row-pair ring from the ring: pair-row p (padded rows p and p+1, p = 0..R+1) of view v,
// element (col) = (ring[v][p][col], ring[v][p+1][col]).
template <typename ST>
// Views v0 + z, or with v_end_dev the last gridDim.z views before the device view count.
__global__ __launch_bounds__(256) void pair_convert_kernel(const ST* __restrict__ ring,
                                                          ST* __restrict__ pring,
                                                          int64_t view_stride, int64_t pview_stride,
                                                          int pitch, int v0,
                                                          const int* __restrict__ v_end_dev) {
    const int col = blockIdx.x * 256 + threadIdx.x;
    const int pr = blockIdx.y;
    const int64_t v = (v_end_dev ? *v_end_dev - static_cast<int>(gridDim.z) : v0) +
                      static_cast<int64_t>(blockIdx.z);
    if (col >= pitch) return;
    const ST* src = ring + v * view_stride + static_cast<int64_t>(pr) * pitch + col;
    ST* dst = pring + v * pview_stride + (static_cast<int64_t>(pr) * pitch + col) * 2;
    dst[0] = src[0];
    dst[1] = src[pitch];
}

// Band ring from the ring (16-bit forms): element (band b, col) of view v = padded rows
// 3b..3b+3 of that column (rows past the ring's Rp rows read as zero), 8 bytes.
// Views v0 + z, or with v_end_dev the last gridDim.z views before the device view count
// (a captured streaming graph converts the batch it has just stored).
__global__ __launch_bounds__(256) void band_convert_kernel(const uint16_t* __restrict__ ring,
                                                           uint16_t* __restrict__ bring,
                                                           int64_t view_stride, int64_t bview_stride,
                                                           int pitch, int Rp, int v0,
                                                           const int* __restrict__ v_end_dev) {
    const int col = blockIdx.x * 256 + threadIdx.x;
    const int b = blockIdx.y;
    const int64_t v = (v_end_dev ? *v_end_dev - static_cast<int>(gridDim.z) : v0) +
                      static_cast<int64_t>(blockIdx.z);
    if (col >= pitch) return;
    const uint16_t* src = ring + v * view_stride + col;
    uint32_t q[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = 3 * b + i;
        q[i] = r < Rp ? src[static_cast<int64_t>(r) * pitch] : 0u;
    }
    uint2* dst = reinterpret_cast<uint2*>(bring + v * bview_stride + (static_cast<int64_t>(b) * pitch + col) * 4);
    *dst = make_uint2(q[0] | (q[1] << 16), q[2] | (q[3] << 16));
}

// Lane layouts of v4: which in-tile point (dp along a, ds along b) thread t handles
// as its m-th point, in a workgroup that covers rows [sub * 8 * PPT, (sub + 1) * 8 * PPT)
// of the 32 x 32 tile.  LAY 0: each wave covers 2 rows x 32 columns per point
// (points 8 rows apart); LAY 1: each wave covers a 4 x 16 patch per point.  A gather
// instruction costs one cache-line lookup per distinct line its 64 lanes touch, and the
// better layout depends on how the slice axes project onto the detector.
template <int PPT, int LAY> struct Map4 {
    static constexpr int SUBS = BP_TP / (8 * PPT);
    __device__ static inline void base(int t, int sub, int& p, int& s) {
        if (LAY == 0) {
            s = t & (BP_TS - 1);
            p = (t >> 5) + sub * 8 * PPT;
        } else {
            const int l = t & 63, w = t >> 6;
            s = l & 15;
            if (PPT == 1) { p = sub * 8 + 4 * (w >> 1) + (l >> 4); s += 16 * (w & 1); }
            else if (PPT == 2) p = sub * 16 + 4 * w + (l >> 4);
            else p = 8 * w + (l >> 4);
        }
    }
    static constexpr int dp(int m) { return LAY == 0 ? 8 * m : (PPT == 4 ? 4 * (m >> 1) : 0); }
    static constexpr int ds(int m) { return LAY == 0 ? 0 : (PPT == 4 ? 16 * (m & 1) : 16 * m); }
};

// 16-bit forms with 2 views in flight, and the 2-point band kernel with 3: at least 8
// waves per SIMD (the row-pair 4-point kernel then fits 59 VGPRs without spills instead of
// 72, the 2-point band kernel with 3 views 61 instead of 69, i.e. 7 waves); the others
// keep the compiler's choice (they would spill).
//
// Cone beam, views [vb, ve) of n_views: a launch with vb > 0 resumes from the unscaled
// FP32 sums the previous launch left in `out`, and one with ve < n_views leaves them
// there, so an update split into several launches equals the single launch bit for bit.
// part != nullptr: the grid is `parts` copies of the tile grid; copy k sums views
// [k kv, (k + 1) kv) from zero and leaves the unscaled sums of its tile's points in
// part[k][tile] (32 x 32 floats), which part_sum_kernel adds in part order (parallel
// beam: each part's row-lerped sum, the row of a point being the same in every view).
template <bool CONE, typename ST, int PPT, int LAY, int DEPTH, int RING, bool SYNC = false>
__global__ __launch_bounds__(BP_THREADS, (sizeof(ST) == 2 && ((DEPTH == 2 && (RING < 2 || PPT < 4)) || (RING == 2 && PPT == 2))) ? 8 : 1)
void bp_tiles_kernel4(
    const ST* __restrict__ ring, int64_t view_stride, int pitch,
    const float* __restrict__ recs, const mr_task* __restrict__ tasks,
    const Tile* __restrict__ tiles, int n_views, float scale, float* __restrict__ out, int xcd,
    int vb, int ve, float* __restrict__ part, int parts, int kv) {
    using M = Map4<PPT, LAY>;
    constexpr int LE = Taps<ST>::LOG_ESZ;
    constexpr int SUBS = M::SUBS;                    // workgroups per 32-row tile
    static_assert(SUBS * 8 * PPT == BP_TP, "tile shape");
    int lb = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    int pk = 0;                                      // part of a split launch
    const int nbw = static_cast<int>(gridDim.x) / parts;
    if (part) {
        pk = lb / nbw;
        lb -= pk * nbw;
        vb = pk * kv;
        ve = min(n_views, vb + kv);
    }
    const int tile = lb / SUBS;
    const int sub = lb - tile * SUBS;
    const Tile tl = tiles[tile];
    const mr_task& T = tasks[tl.task];
    const int t = threadIdx.x;
    int pt, st0;
    M::base(t, sub, pt, st0);
    const int p0 = tl.tp * BP_TP, s0 = tl.ts * BP_TS;
    bool valid[PPT], inside[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        const int p = p0 + pt + M::dp(m), s = s0 + st0 + M::ds(m);
        valid[m] = (p < T.np) && (s < T.ns);
        inside[m] = valid[m] && (!T.mask_box || in_box(T, p, s));
    }
    const float ds = static_cast<float>(st0), dp0 = static_cast<float>(pt);
    // RING 1: `ring` is the row-pair ring (two elements per padded position); RING 2: the
    // band ring (four elements per band and column)
    static_assert(RING < 2 || sizeof(ST) == 2, "band ring of 16-bit storage only");
    constexpr bool PAIRED = RING == 1;
    constexpr int SH = LE + (PAIRED ? 1 : 0);
    const int pitchb = RING == 2 ? pitch << 3 : pitch << SH;
    const int view_bytes = static_cast<int>(view_stride << LE);
    const float* __restrict__ rt = recs + static_cast<int64_t>(tile) * n_views * BP_REC;

    float acc[PPT], accb[PPT], frp[PPT];
    int roffp[PPT], jsh[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) { acc[m] = 0.0f; accb[m] = 0.0f; frp[m] = 0.0f; roffp[m] = 0; jsh[m] = 0; }
    const int64_t obase = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q;
    auto oidx = [&](int m) {
        return obase + static_cast<int64_t>(p0 + pt + M::dp(m)) * T.stride_p +
               static_cast<int64_t>(s0 + st0 + M::ds(m)) * T.stride_s;
    };
    if (CONE && vb > 0 && !part) {
#pragma unroll
        for (int m = 0; m < PPT; ++m)
            if (valid[m]) acc[m] = out[oidx(m)];
    }
    if (!CONE && n_views > 0) {
        // parallel beam: the row of a point is the same for every view
#pragma unroll
        for (int m = 0; m < PPT; ++m) {
            const float dp = dp0 + static_cast<float>(M::dp(m));
            const float dsm = ds + static_cast<float>(M::ds(m));
            float row = fmaf(dp, rt[5], fmaf(dsm, rt[4], rt[3]));
            row = __builtin_amdgcn_fmed3f(row, rt[11], rt[12]);
            frp[m] = __builtin_amdgcn_fractf(row);
            if (RING == 2) {
                // band ring: the row part of band_addr and the row shift (the same for
                // every view)
                const int r = cvt_flr(row), q = band_q(r);
                roffp[m] = __mul24(q, pitchb);
                jsh[m] = band_shift(q, r);
            } else {
                roffp[m] = cvt_flr(row) * pitchb;
            }
        }
    }

    // coefficient c(m) = c0 + dp(m) * cp + ds(m) * cs with compile-time offsets
    auto step = [](float c0, int m, float cp, float cs) {
        float c = c0;
        if (M::dp(m)) c = fmaf(static_cast<float>(M::dp(m)), cp, c);
        if (M::ds(m)) c = fmaf(static_cast<float>(M::ds(m)), cs, c);
        return c;
    };
    // request the taps of view v into st
    auto issue = [&](int v, PointState<ST, PPT, RING>& st) {
        const float* r = rt + static_cast<int64_t>(v) * BP_REC;
        const float U0 = r[0], Ua = r[1], Ub = r[2];
        // band ring: [14] is the byte offset of the reference band and column
        const int ref14 = __float_as_int(r[14]);
        const int baseb = RING == 2 ? band_base(ref14, pitchb) : ref14 << SH;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<ST*>(ring + static_cast<int64_t>(v) * view_stride), 0, view_bytes, 0x00020000);
        const float uh0 = fmaf(dp0, Ub, fmaf(ds, Ua, U0));
        if (CONE) {
            const float V0 = r[3], Va = r[4], Vb = r[5], W0 = r[6], Wa = r[7], Wb = r[8];
            const float vh0 = fmaf(dp0, Vb, fmaf(ds, Va, V0));
            const float wh0 = fmaf(dp0, Wb, fmaf(ds, Wa, W0));
#pragma unroll
            for (int m = 0; m < PPT; ++m) {
                const float uh = step(uh0, m, Ub, Ua);
                const float vh = step(vh0, m, Vb, Va);
                const float wh = step(wh0, m, Wb, Wa);
                const float rw = __builtin_amdgcn_rcpf(wh);
                const float col = __builtin_amdgcn_fmed3f(uh * rw, r[9], r[10]);
                const float row = __builtin_amdgcn_fmed3f(vh * rw, r[11], r[12]);
                st.wt[m] = rw * rw;
                st.fc[m] = __builtin_amdgcn_fractf(col);
                st.fr[m] = __builtin_amdgcn_fractf(row);
                if (RING == 2) {
                    const int r = cvt_flr(row), q = band_q(r);
                    st.tap[m].load(rs, band_addr(q, cvt_flr(col), pitchb, baseb), band_shift(q, r));
                } else {
                    const int voff = __mul24(cvt_flr(row), pitchb) + ((cvt_flr(col) << SH) + baseb);
                    st.tap[m].load(rs, voff, pitchb);
                }
            }
        } else {
            st.wt[0] = r[13];
#pragma unroll
            for (int m = 0; m < PPT; ++m) {
                const float col = __builtin_amdgcn_fmed3f(step(uh0, m, Ub, Ua), r[9], r[10]);
                st.fc[m] = __builtin_amdgcn_fractf(col);
                if (RING == 2) {
                    st.tap[m].load(rs, roffp[m] + ((cvt_flr(col) << 3) + baseb), jsh[m]);
                } else {
                    const int voff = roffp[m] + ((cvt_flr(col) << SH) + baseb);
                    st.tap[m].load(rs, voff, pitchb);
                }
            }
        }
    };
    auto consume = [&](const PointState<ST, PPT, RING>& st) {
#pragma unroll
        for (int m = 0; m < PPT; ++m) {
            float top, bot;
            st.tap[m].lerp(st.fc[m], top, bot);
            if (CONE) {
                acc[m] = fmaf(st.wt[m], fmaf(st.fr[m], bot - top, top), acc[m]);
            } else {
                acc[m] = fmaf(st.wt[0], top, acc[m]);
                accb[m] = fmaf(st.wt[0], bot, accb[m]);
            }
        }
    };

    // DEPTH views in flight: the taps of view v + DEPTH - 1 are requested before view v
    // is interpolated.  Requests past the last view re-read it (never consumed), so the
    // loop body is unconditional and every wait counts exactly the loads after it.
    if (ve > vb) {
        PointState<ST, PPT, RING> S[DEPTH];
        const int last = ve - 1;
#pragma unroll
        for (int j = 0; j + 1 < DEPTH; ++j) issue(min(vb + j, last), S[j]);
        int v = vb;
        for (; v + DEPTH <= ve; v += DEPTH) {
#pragma unroll
            for (int j = 0; j < DEPTH; ++j) {
                issue(min(v + j + DEPTH - 1, last), S[(j + DEPTH - 1) % DEPTH]);
                // keep the requests ahead of the interpolation of older views (the
                // scheduler would otherwise sink them below the waits)
                __builtin_amdgcn_sched_barrier(0);
                // SYNC: the four waves (adjacent rows of one tile, so mostly the same
                // detector lines) request each view together and share its L1 fills;
                // an execution barrier only, loads stay in flight across it
                if (SYNC) __builtin_amdgcn_s_barrier();
                consume(S[j]);
            }
        }
#pragma unroll
        for (int j = 0; j + 1 < DEPTH; ++j)
            if (v + j < ve) consume(S[j]);
    }

    if (part) {
        // the point (pt + dp(m), st0 + ds(m)) of the tile, row-major
        float* dst = part + (static_cast<int64_t>(pk) * (nbw / SUBS) + tile) * (BP_TP * BP_TS);
#pragma unroll
        for (int m = 0; m < PPT; ++m)
            if (valid[m])
                dst[(pt + M::dp(m)) * BP_TS + st0 + M::ds(m)] =
                    CONE ? acc[m] : fmaf(frp[m], accb[m] - acc[m], acc[m]);
        return;
    }
    const bool final_views = !CONE || ve >= n_views;
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        if (!valid[m]) continue;
        float val = acc[m];
        if (!CONE) val = fmaf(frp[m], accb[m] - acc[m], acc[m]);
        out[oidx(m)] = final_views ? (inside[m] ? val * scale : 0.0f) : val;
    }
}

// Four taps from an LDS box with rows strb bytes apart: f32 two dwords per row; 16-bit
// formats one dword per row at 2-byte alignment.
template <typename ST> struct LdsPair;
template <> struct LdsPair<float> {
    __device__ static inline Pair<float> read(const unsigned char* __restrict__ p, int strb) {
        Pair<float> t;
        t.a = *reinterpret_cast<const float*>(p);
        t.b = *reinterpret_cast<const float*>(p + 4);
        t.c = *reinterpret_cast<const float*>(p + strb);
        t.d = *reinterpret_cast<const float*>(p + strb + 4);
        return t;
    }
};
template <typename ST> struct LdsPair {       // 16-bit formats: one dword per row
    __device__ static inline Pair<ST> read(const unsigned char* __restrict__ p, int strb) {
        Pair<ST> t;
        __builtin_memcpy(&t.w0, p, 4);            // 2-byte aligned: unaligned LDS access
        __builtin_memcpy(&t.w1, p + strb, 4);
        return t;
    }
};

// ---------------------------------------------------------------- z columns (volumes)
// When neither the detector column nor the depth of a point depends on z (P[0] and P[2]
// have no z term: a circular orbit about the z axis with the detector's v axis along z,
// as circular_cone_vectors builds it), every voxel of a column (x, y) of a volume has, in
// each view, the same 1/w, the same column and so the same column fraction and FDK
// weight; only its row moves, linearly in z.  The z-column kernel gives each thread one
// column and a chunk of ZQ planes: per view it forms the reciprocal, the column, its
// fraction and offset and the weight once, then per plane one FMA for the row, the clamp,
// floor and fraction, the gather and the lerps.  The plane tile kernels redo the
// projective division and the column per point.
//
// Records of a (column tile, plane chunk, view) are built in float64 by threads of the
// workgroup into LDS, one pass of 64 views at a time (as in the v2 loop), and read back as
// VGPRs, so the FMAs of the per-column step take every coefficient from a VGPR.
//   [0] U0 [1] Ua [2] Ub [3] V0 [4] Va [5] Vb [6] W0 [7] Wa [8] Wb: u_h, v_h, w_h at the
//       tile's first column of the chunk's first plane and their steps along b (s) and
//       a (p), relative to the integer reference (cref, rref) of the chunk's centre
//   [9] [10] column clamp, [11] [12] row clamp (relative), [13] Vc: v_h per plane
//   [14] rref + 1, [15] cref + 1: the padded reference position (int bits)
//   [16] [17] [18] [19] (staged kernel) the detector box of the chunk's taps in the
//       ring it reads: first row, rows (0: the box does not fit the LDS buffer and the
//       view is gathered from memory), first column (a multiple of the elements in 16
//       bytes), 16-byte chunks per row (int bits)
// The tile-independent part, per (task, view), comes from zcol_coef_kernel (ZC doubles:
// u_h, v_h, w_h at the task origin and their steps along a and b, v_h along c), so the
// record build holds few float64 values (the full build from the view's 3 x 4 form held
// ~100 VGPRs, against ~70 for the main loop).
// (ZT x ZT columns and ZQ or ZQL planes per tile: mr_tiles.h)
constexpr int ZG = 8;          // planes whose gathers are in flight together
constexpr int ZC = 12;         // doubles per (task, view) coefficient set (10 used)
constexpr int ZREC = 20;       // floats per record
constexpr int ZL_BUF = 16384;  // bytes of one staged box (two per workgroup)
static_assert(ZT * ZT == BP_THREADS, "one column per thread");

// coef[(task * n_views + v) * ZC]: Uo Ua Ub Vo Va Vb Vc Wo Wa Wb of view slot v for task
// `task`, with the task's COR shift (the point's u_h gains shift * w_h, v_h shift * u.vv *
// w_h); a, b, c are the task's axes (p, s, plane), o its origin.
__global__ __launch_bounds__(256) void zcol_coef_kernel(const ViewGeom* __restrict__ geom,
                                                        const mr_task* __restrict__ tasks,
                                                        int ntasks, int n_views,
                                                        double* __restrict__ coef) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= static_cast<int64_t>(ntasks) * n_views) return;
    const int ti = static_cast<int>(gid / n_views);
    const int v = static_cast<int>(gid - static_cast<int64_t>(ti) * n_views);
    const mr_task& T = tasks[ti];
    const ViewGeom& g = geom[v];
    const double sh = T.cor_shift, svv = sh * g.uvv;
    const double Wo = dot3(g.P[2], T.o) + g.P[2][3], Wa = dot3(g.P[2], T.a), Wb = dot3(g.P[2], T.b);
    double* k = coef + gid * ZC;
    k[0] = dot3(g.P[0], T.o) + g.P[0][3] + sh * Wo;
    k[1] = dot3(g.P[0], T.a) + sh * Wa;
    k[2] = dot3(g.P[0], T.b) + sh * Wb;
    k[3] = dot3(g.P[1], T.o) + g.P[1][3] + svv * Wo;
    k[4] = dot3(g.P[1], T.a) + svv * Wa;
    k[5] = dot3(g.P[1], T.b) + svv * Wb;
    k[6] = dot3(g.P[1], T.c);          // P[0].c = P[2].c = 0 for this kernel
    k[7] = Wo;
    k[8] = Wa;
    k[9] = Wb;
}

// Part `part` of the record of the tile whose first column is (p0, s0) on plane q0 (NQ
// planes), from its task's coefficient set k: part 0 the column half (u_h, w_h: [0..2],
// [6..10], [15]), part 1 the row half (v_h: [3..5], [11..14]), part 2 the box ([16..19]).
// Threads of the workgroup build the parts in parallel, so fewer float64 values are live
// at once in the kernel that holds the sums.  The integer reference only needs to lie
// near the chunk's centre (coordinates relative to any integer are exact), so it is
// chosen in float; the coefficients relative to it are formed in float64.  The box is
// conservative: the corners of the column tile x plane chunk, projected in float (within
// ~1e-4 px of the kernel's relative coordinates; margin 1/16 px), clamped as the kernel
// clamps, padded by the second tap; a projective map with w > 0 at every corner sends the
// box of points into the hull of the corner images.
// PAIRED: the box rows are pair-rows of the row-pair copy (one per tap pair); else rows of
// the ring itself (a tap pair needs rows r and r + 1).  LEZ: log2 bytes per element of
// the staged ring (16-byte chunks of 16 >> LEZ elements).
template <int NQ, bool PAIRED, int LEZ>
__device__ void make_zrecord(const double* __restrict__ k, float vs, int part, int q0, int p0,
                             int s0, int R, int C, int pitch, float* __restrict__ rec) {
    const float hc = 0.5f * (ZT - 1), hq = 0.5f * (NQ - 1);
    if (part == 2) {
        const float Wo = static_cast<float>(k[7] + p0 * k[8] + s0 * k[9]);
        const float Uo = static_cast<float>(k[0] + p0 * k[1] + s0 * k[2]);
        const float Vo = static_cast<float>(k[3] + p0 * k[4] + s0 * k[5] + q0 * k[6]);
        const float Wp = static_cast<float>(k[8]), Ws = static_cast<float>(k[9]);
        const float Up = static_cast<float>(k[1]), Us = static_cast<float>(k[2]);
        const float Vp = static_cast<float>(k[4]), Vs = static_cast<float>(k[5]);
        const float Vq = static_cast<float>(k[6]) * (NQ - 1);
        const float cc = 0.5f * (C - 1), cr = 0.5f * (R - 1), e = ZT - 1;
        float cmin = 3.0e38f, cmax = -3.0e38f, rmin = 3.0e38f, rmax = -3.0e38f;
        bool ok = true;
#pragma unroll
        for (int c4 = 0; c4 < 4; ++c4) {
            const float a = (c4 & 1) ? e : 0.0f, b = (c4 & 2) ? e : 0.0f;
            const float w = Wo + a * Wp + b * Ws;
            ok = ok && w > 0.0f;
            const float rw = 1.0f / w;
            const float col = (Uo + a * Up + b * Us) * rw + cc;
            const float v0 = Vo + a * Vp + b * Vs;
            const float r0 = v0 * rw + cr, r1 = (v0 + Vq) * rw + cr;
            cmin = fminf(cmin, col);
            cmax = fmaxf(cmax, col);
            rmin = fminf(rmin, fminf(r0, r1));
            rmax = fmaxf(rmax, fmaxf(r0, r1));
        }
        const float m = 0.0625f;
        cmin = fminf(fmaxf(cmin, -1.0f), static_cast<float>(C));
        cmax = fminf(fmaxf(cmax, -1.0f), static_cast<float>(C));
        rmin = fminf(fmaxf(rmin, -1.0f), static_cast<float>(R));
        rmax = fminf(fmaxf(rmax, -1.0f), static_cast<float>(R));
        // padded: a tap pair at detector column c covers padded columns c + 1, c + 2; the
        // pair-row of detector row r is padded row r + 1 (the ring: rows r + 1, r + 2)
        constexpr int CE = 16 >> LEZ;      // elements per 16-byte chunk
        const int c0 = max(0, static_cast<int>(floorf(cmin - m)) + 1) & ~(CE - 1);
        const int c1 = min(C + 2, static_cast<int>(floorf(cmax + m)) + 2);
        const int pr0 = max(0, static_cast<int>(floorf(rmin - m)) + 1);
        const int pr1 = PAIRED ? min(R + 1, static_cast<int>(floorf(rmax + m)) + 1)
                               : min(R + 2, static_cast<int>(floorf(rmax + m)) + 2);
        const int nch = (c1 - c0) / CE + 1;
        const int nr = pr1 - pr0 + 1;
        const bool fits = ok && nr > 0 && nr * nch * 16 <= ZL_BUF;
        rec[16] = __int_as_float(pr0);
        rec[17] = __int_as_float(fits ? nr : 0);
        rec[18] = __int_as_float(c0);
        rec[19] = __int_as_float(nch);
        return;
    }
    const double W0 = k[7] + p0 * k[8] + s0 * k[9];
    const float rwc = 1.0f / static_cast<float>(W0 + hc * (k[8] + k[9]));
    // the weight 1/w_h^2 absorbs the view's storage scale
    const double f = 1.0 / sqrt(static_cast<double>(vs));
    if (part == 0) {
        const double U0 = k[0] + p0 * k[1] + s0 * k[2];
        float colc = static_cast<float>(U0 + hc * (k[1] + k[2])) * rwc + 0.5f * (C - 1);
        colc = fminf(fmaxf(colc, -static_cast<float>(C)), 2.0f * C);   // NaN -> -C
        const int cref = static_cast<int>(floorf(colc));
        const double dc = 0.5 * (C - 1) - cref;
        rec[0] = static_cast<float>((U0 + dc * W0) * f);
        rec[1] = static_cast<float>((k[2] + dc * k[9]) * f);      // along b (s)
        rec[2] = static_cast<float>((k[1] + dc * k[8]) * f);      // along a (p)
        rec[6] = static_cast<float>(W0 * f);
        rec[7] = static_cast<float>(k[9] * f);
        rec[8] = static_cast<float>(k[8] * f);
        rec[9] = static_cast<float>(-1 - cref);
        rec[10] = static_cast<float>(C - cref);
        rec[15] = __int_as_float(cref + 1);
    } else {
        const double V0 = k[3] + p0 * k[4] + s0 * k[5] + q0 * k[6];
        float rowc = static_cast<float>(V0 + hc * (k[4] + k[5]) + hq * k[6]) * rwc + 0.5f * (R - 1);
        rowc = fminf(fmaxf(rowc, -static_cast<float>(R)), 2.0f * R);
        const int rref = static_cast<int>(floorf(rowc));
        const double dr = 0.5 * (R - 1) - rref;
        rec[3] = static_cast<float>((V0 + dr * W0) * f);
        rec[4] = static_cast<float>((k[5] + dr * k[9]) * f);
        rec[5] = static_cast<float>((k[4] + dr * k[8]) * f);
        rec[11] = static_cast<float>(-1 - rref);
        rec[12] = static_cast<float>(R - rref);
        rec[13] = static_cast<float>(k[6] * f);
        rec[14] = __int_as_float(rref + 1);
    }
}

// Tile tl: columns p = tl.tp * ZT + (t >> 4), s = tl.ts * ZT + (t & 15) of task tl.task,
// planes tl.q .. tl.q + NQ - 1 (the last chunk of a task may have fewer; the sums of the
// planes past the task are formed and dropped).  RING 0: the ring (two row loads per
// point), 1: the row-pair copy (one load of the four taps); view_stride: elements of ST
// per view of that ring.
//
// STAGE (NQ = ZQL): the gathers come from LDS instead of the texture path, whose rate (one
// 64-lane load instruction per ~16-24 cycles per CU) bounds the gather form on volumes.
// Per view the workgroup copies the detector box of its chunk (record [16..19]) from the
// ring (or the 16-bit row-pair copy) into LDS with direct 16-byte buffer-to-LDS loads,
// the next view's box into the other buffer while the current view is interpolated (one
// barrier per view); each point then reads its taps from LDS: one read per row (the ring;
// 16-bit rows at 2-byte alignment, which LDS serves unaligned) or one for both rows (the
// row-pair copy).  A view whose box does not fit ZL_BUF is gathered from memory.
// 16 bytes at byte offset voff of buffer rs to LDS at lds + 16 * lane (lds uniform over
// the wave): buffer_load_dwordx4 ... lds, no VGPRs, counted by vmcnt.
__device__ inline void lds_load16(__amdgpu_buffer_rsrc_t rs, uint32_t* lds, int voff) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, lds, 16, voff, 0, 0, 0);
}

template <typename ST, int RING, bool STAGE, int NQ>
__global__ __launch_bounds__(BP_THREADS) void bp_zcol_kernel(
    const ST* __restrict__ ring, int64_t view_stride, int pitch, int R, int C,
    const double* __restrict__ coef, const float* __restrict__ vscale,
    const mr_task* __restrict__ tasks, const Tile* __restrict__ tiles, int n_views, float scale,
    float* __restrict__ out, int xcd) {
    static_assert(!STAGE || RING == 0 || sizeof(ST) == 2, "staged boxes: the ring or a 16-bit row-pair copy");
    __shared__ float4 rec4[BP_VPASS * ZREC / 4];
    __shared__ __attribute__((aligned(16))) uint32_t sbox[STAGE ? 2 * ZL_BUF / 4 : 1];
    float* rec = reinterpret_cast<float*>(rec4);
    const int lb = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const Tile tl = tiles[lb];
    const mr_task& T = tasks[tl.task];
    const int t = threadIdx.x;
    const int p = tl.tp * ZT + (t >> 4), s = tl.ts * ZT + (t & (ZT - 1));
    const double* __restrict__ kt = coef + static_cast<int64_t>(tl.task) * n_views * ZC;
    const float ds = static_cast<float>(t & (ZT - 1)), dp = static_cast<float>(t >> 4);
    constexpr int LE = Taps<ST>::LOG_ESZ;
    constexpr int SH = LE + (RING == 1 ? 1 : 0);     // log2 bytes per ring element
    const int pitchb = pitch << SH;
    const int view_bytes = static_cast<int>(view_stride << LE);
    using Tap = typename std::conditional<RING == 1, PairP<ST>, Pair<ST>>::type;
    auto rsrc = [&](int v) {
        return __builtin_amdgcn_make_buffer_rsrc(
            const_cast<ST*>(ring + static_cast<int64_t>(v) * view_stride), 0, view_bytes, 0x00020000);
    };
    // the box of view vb + v (record rv) into LDS buffer bsel: 16-byte chunk k of the box
    // (row k / nch, chunk k mod nch) to LDS byte 16 k; each wave copies runs of 64 chunks
    auto stage = [&](int vv, const float* rv, int bsel) {
        if constexpr (STAGE) {
            // the box descriptor is the same for the whole workgroup: scalar registers
            const int pr0 = __builtin_amdgcn_readfirstlane(__float_as_int(rv[16]));
            const int nr = __builtin_amdgcn_readfirstlane(__float_as_int(rv[17]));
            const int c0 = __builtin_amdgcn_readfirstlane(__float_as_int(rv[18]));
            const int nch = __builtin_amdgcn_readfirstlane(__float_as_int(rv[19]));
            if (nr == 0) return;
            const int total = nr * nch;
            const __amdgpu_buffer_rsrc_t rs = rsrc(vv);
            // chunk k = t + 256 j: its row k / nch and chunk cc = k mod nch, the first by one
            // division, then steps of 256 chunks = dq rows and dr < nch chunks, one row more
            // when cc passes nch.  floor((x + 0.5) * rcp(nch)) is exact for x <= 256: the
            // quotient lies at least 1 / (2 nch) from an integer, and the reciprocal's error
            // (1 ulp) moves it by under 2^-14 (a staged box has nch <= ZL_BUF / 16 = 1024)
            const float inv = __builtin_amdgcn_rcpf(static_cast<float>(nch));
            const int row = static_cast<int>((static_cast<float>(t) + 0.5f) * inv);
            int cc = t - row * nch;
            int voff = (((pr0 + row) * pitch + c0) << SH) + (cc << 4);
            const int dq = __builtin_amdgcn_readfirstlane(
                static_cast<int>((static_cast<float>(BP_THREADS) + 0.5f) * inv));
            const int dr = BP_THREADS - dq * nch;
            const int dvoff = dq * pitchb + (dr << 4);
            const int wrap = pitchb - (nch << 4);
            const int k0w = __builtin_amdgcn_readfirstlane(t & ~63);   // the wave's first chunk
            const int lim = total - (t & 63);
            uint32_t* dst = sbox + bsel * (ZL_BUF / 4) + k0w * 4;
            for (int k0 = k0w; k0 < total; k0 += BP_THREADS) {
                if (k0 < lim) lds_load16(rs, dst, voff);
                dst += BP_THREADS * 4;
                cc += dr;
                voff += dvoff;
                if (cc >= nch) {
                    cc -= nch;
                    voff += wrap;
                }
            }
        }
    };
    float acc[NQ];
#pragma unroll
    for (int i = 0; i < NQ; ++i) acc[i] = 0.0f;
    for (int vb = 0; vb < n_views; vb += BP_VPASS) {
        const int nv = min(BP_VPASS, n_views - vb);
        __syncthreads();
        {
            const int part = t >> 6, v = t & 63;
            if (v < nv && part < (STAGE ? 3 : 2))
                make_zrecord<NQ, RING == 1, SH>(kt + static_cast<int64_t>(vb + v) * ZC, vscale[vb + v],
                                                part, tl.q, tl.tp * ZT, tl.ts * ZT, R, C, pitch,
                                                rec + v * ZREC);
        }
        __syncthreads();
        if (STAGE) {
            stage(vb, rec, 0);
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        }
        for (int v = 0; v < nv; ++v) {
            const float* rv = rec + v * ZREC;
            const float4 r0 = rec4[v * (ZREC / 4) + 0], r1 = rec4[v * (ZREC / 4) + 1];
            const float4 r2 = rec4[v * (ZREC / 4) + 2], r3 = rec4[v * (ZREC / 4) + 3];
            if (STAGE && v + 1 < nv) stage(vb + v + 1, rv + ZREC, (v + 1) & 1);
            // once per column and view
            const float uh = fmaf(dp, r0.z, fmaf(ds, r0.y, r0.x));
            const float vh = fmaf(dp, r1.y, fmaf(ds, r1.x, r0.w));
            const float wh = fmaf(dp, r2.x, fmaf(ds, r1.w, r1.z));
            const float rw = __builtin_amdgcn_rcpf(wh);
            const float col = __builtin_amdgcn_fmed3f(uh * rw, r2.y, r2.z);
            const float fc = __builtin_amdgcn_fractf(col);
            const float wt = rw * rw;
            const float row0 = vh * rw, drow = r3.y * rw;
            const int ci = cvt_flr(col);
            const int rref1 = __float_as_int(r3.z), cref1 = __float_as_int(r3.w);
            bool staged = false;
            // staged kernels: each path leaves the planes' interpolated values in val and the
            // sums are updated once after both (so the sums have one definition per view
            // and are not copied where the paths meet)
            float val[STAGE ? NQ : 1];
            if constexpr (STAGE) staged = __builtin_amdgcn_readfirstlane(__float_as_int(rv[17])) != 0;
            if constexpr (STAGE) if (staged) {
                // LDS byte address of (pair-row rref1 + r, column cref1 + ci) in the box:
                // r * SB + cb, SB = 16 nch bytes per box row
                const int pr0 = __builtin_amdgcn_readfirstlane(__float_as_int(rv[16]));
                const int c0 = __builtin_amdgcn_readfirstlane(__float_as_int(rv[18]));
                const int sb = __builtin_amdgcn_readfirstlane(__float_as_int(rv[19])) << 4;
                const int cb = ((cref1 + ci - c0) << SH) + (rref1 - pr0) * sb;
                const char* bx = reinterpret_cast<const char*>(sbox) + (v & 1) * ZL_BUF;
                constexpr bool PK = RING == 1 && std::is_same<ST, __half>::value;
                const _Float16 fch = static_cast<_Float16>(fc);
                const half2v fc2 = {fch, fch};
                // the planes' rows fma(i, drow, row0) are monotone in i, so when the first
                // and the last of every lane lie in the clamp range, all do and the clamp
                // (the identity there) is left out
                const float rl = fmaf(static_cast<float>(NQ - 1), drow, row0);
                const bool in_range = fminf(row0, rl) >= r2.w && fmaxf(row0, rl) <= r3.x;
                auto planes = [&](auto clamp) {
#pragma unroll
                    for (int i = 0; i < NQ; ++i) {
                        float row = fmaf(static_cast<float>(i), drow, row0);
                        if constexpr (decltype(clamp)::value) row = __builtin_amdgcn_fmed3f(row, r2.w, r3.x);
                        const uint32_t* q = reinterpret_cast<const uint32_t*>(bx + __mul24(cvt_flr(row), sb) + cb);
                        const float fr = __builtin_amdgcn_fractf(row);
                        if constexpr (PK) {
                            // f16 row-pair copy: packed column lerps
                            val[i] = pair_bilerp_f16(q[0], q[1], fc2, fr);
                        } else {
                            float top, bot;
                            if constexpr (RING == 1) {
                                PairP<ST> tap;     // (q(r, c) | q(r+1, c)), same for c + 1
                                tap.w0 = q[0];
                                tap.w1 = q[1];
                                tap.lerp(fc, top, bot);
                            } else {
                                // the ring: row r, then row r + 1 one box row further
                                const Pair<ST> tap =
                                    LdsPair<ST>::read(reinterpret_cast<const unsigned char*>(q), sb);
                                tap.lerp(fc, top, bot);
                            }
                            val[i] = fmaf(fr, bot - top, top);
                        }
                    }
                };
                if (__all(in_range)) planes(std::false_type{});
                else planes(std::true_type{});
            }
            if (!staged) {
                const __amdgpu_buffer_rsrc_t rs = rsrc(vb + v);
                const int cofs = (ci + __mul24(rref1, pitch) + cref1) << SH;
                // per plane: the row, then ZG gathers in flight before their lerps.  The
                // group's rows start from an opaque copy of row0 and a scheduling barrier
                // ends the group, so the compiler neither forms every plane's row up front
                // nor hoists the next group's gathers (either holds NQ more VGPRs)
#pragma unroll
                for (int g = 0; g < NQ; g += ZG) {
                    Tap tap[ZG];
                    float fr[ZG];
                    float rb = row0;
                    asm volatile("" : "+v"(rb));
#pragma unroll
                    for (int i = 0; i < ZG; ++i) {
                        const float row = __builtin_amdgcn_fmed3f(
                            fmaf(static_cast<float>(g + i), drow, rb), r2.w, r3.x);
                        fr[i] = __builtin_amdgcn_fractf(row);
                        tap[i].load(rs, __mul24(cvt_flr(row), pitchb) + cofs, pitchb);
                    }
#pragma unroll
                    for (int i = 0; i < ZG; ++i) {
                        float top, bot;
                        tap[i].lerp(fc, top, bot);
                        if constexpr (STAGE) val[g + i] = fmaf(fr[i], bot - top, top);
                        else acc[g + i] = fmaf(wt, fmaf(fr[i], bot - top, top), acc[g + i]);
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
            if constexpr (STAGE) {
#pragma unroll
                for (int i = 0; i < NQ; ++i) acc[i] = fmaf(wt, val[i], acc[i]);
            }
            if (STAGE) {
                // the next view's box is in LDS, and every wave is done with this one
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __syncthreads();
            }
        }
    }
    if (p < T.np && s < T.ns) {
        const int64_t ob = T.out_offset + static_cast<int64_t>(p) * T.stride_p +
                           static_cast<int64_t>(s) * T.stride_s +
                           static_cast<int64_t>(tl.q) * T.stride_q;
        const int nq = T.nq - tl.q;
#pragma unroll
        for (int i = 0; i < NQ; ++i)
            if (i < nq) out[ob + static_cast<int64_t>(i) * T.stride_q] = acc[i] * scale;
    }
}

// ---------------------------------------------------------------- v5: LDS-staged detector boxes
// Direct gathers cost one cache-line lookup per distinct line a wave touches (~16 per
// instruction for an oblique tile), and every line fetched from HBM is waited on
// inside the view loop.  v5 instead stages, per (sub-tile, view), the bounding box of
// the sub-tile's taps into LDS with coalesced 16-byte loads, issued two views ahead
// into registers, and reads the four taps from LDS.
//
// Record: v4's 16 floats, then per sub-tile an 8-int box descriptor:
//   [0] goff  element offset of the box origin in the view slot (padded coordinates)
//   [1] nr    box rows (0: the box does not fit; the view uses direct gathers)
//   [2] nch   16-byte chunks per box row
//   [3] strb  LDS bytes per staged row (nch * 16 + 4: an odd dword stride)
//   [4] koffb LDS byte offset of relative tap (0, 0): a = ri * strb + ci * esz - koffb
//   [5] 1 / nch as float bits, [6] nr * nch
// The box is derived in FP64 from the four corner images of the sub-tile (a
// projective map with w > 0 sends the rectangle into the hull of its corner images),
// clamped like the inner loop, one pixel of margin each side (FP32 coordinates differ
// from FP64 by ~1e-5 px), columns aligned down to 16 bytes.
constexpr int V5_REC = 32;            // floats per (tile, view) record
constexpr int V5_BOX = 8;             // ints per sub-tile box descriptor

template <int LE> struct V5Cfg {
    static constexpr int K = LE == 2 ? 3 : 2;                // 16-byte chunks per thread
    static constexpr int CH = 16 >> LE;                      // elements per chunk
    static constexpr int BUF = K * BP_THREADS * 16 + 1024;   // LDS bytes per stage buffer
};

__device__ void make_box5(const RecRef& q, int p_lo, int p_hi, int R, int C, int pitch, int LE,
                          int kcap, int bufbytes, int* __restrict__ d) {
    double cmin = 1e300, cmax = -1e300, rmin = 1e300, rmax = -1e300;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double ds = (k & 1) ? BP_TS - 1 : 0, dp = (k & 2) ? p_hi : p_lo;
        const double w = q.W0 + ds * q.Wa + dp * q.Wb;
        ok = ok && (w > 0.0);
        const double col = (q.U0 + ds * q.Ua + dp * q.Ub) / w;   // relative to cref
        const double row = (q.V0 + ds * q.Va + dp * q.Vb) / w;   // relative to rref
        cmin = fmin(cmin, col); cmax = fmax(cmax, col);
        rmin = fmin(rmin, row); rmax = fmax(rmax, row);
    }
    ok = ok && fabs(cmin) < 1e7 && fabs(cmax) < 1e7 && fabs(rmin) < 1e7 && fabs(rmax) < 1e7;
    for (int i = 0; i < V5_BOX; ++i) d[i] = 0;
    if (!ok) return;
    const double cl = -1.0 - q.cref, ch = C - q.cref, rl = -1.0 - q.rref, rh = R - q.rref;
    const double c_lo = floor(fmin(fmax(cmin, cl), ch)) - 1.0;
    const double c_hi = floor(fmin(fmax(cmax, cl), ch)) + 2.0;
    const double r_lo = floor(fmin(fmax(rmin, rl), rh)) - 1.0;
    const double r_hi = floor(fmin(fmax(rmax, rl), rh)) + 2.0;
    // padded ring coordinates: columns 0..C+2, rows 0..R+2
    const int pc_lo = max(static_cast<int>(c_lo + q.cref) + 1, 0);
    const int pc_hi = min(static_cast<int>(c_hi + q.cref) + 1, C + 2);
    const int pr_lo = max(static_cast<int>(r_lo + q.rref) + 1, 0);
    const int pr_hi = min(static_cast<int>(r_hi + q.rref) + 1, R + 2);
    const int CH = 16 >> LE;
    const int pc0 = pc_lo & ~(CH - 1);
    const int nch = (pc_hi - pc0) / CH + 1;
    const int nr = pr_hi - pr_lo + 1;
    const int strb = nch * 16 + 4;
    if (nr <= 0 || nch <= 0 || nr * nch > kcap || nr * strb > bufbytes) return;
    const int brel_r = pr_lo - static_cast<int>(q.rref) - 1;
    const int brel_c = pc0 - static_cast<int>(q.cref) - 1;
    d[0] = pr_lo * pitch + pc0;
    d[1] = nr;
    d[2] = nch;
    d[3] = strb;
    d[4] = brel_r * strb + (brel_c << LE);
    d[5] = __float_as_int(1.0f / static_cast<float>(nch));
    d[6] = nr * nch;
}

__global__ __launch_bounds__(256) void records5_kernel(const ViewGeom* __restrict__ geom,
                                                       const float* __restrict__ vscale,
                                                       const mr_task* __restrict__ tasks,
                                                       const Tile* __restrict__ tiles, int ntiles,
                                                       int n_views, int R, int C, int pitch,
                                                       int subs, int le, int kcap, int bufbytes,
                                                       float* __restrict__ recs,
                                                       int* __restrict__ fit) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= static_cast<int64_t>(ntiles) * n_views) return;
    const int ti = static_cast<int>(gid / n_views);
    const int v = static_cast<int>(gid - static_cast<int64_t>(ti) * n_views);
    const Tile tl = tiles[ti];
    const mr_task& T = tasks[tl.task];
    const int p0 = tl.tp * BP_TP, s0 = tl.ts * BP_TS;
    double xt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xt[d] = T.o[d] + tl.q * T.c[d] + p0 * T.a[d] + s0 * T.b[d];
    float* rec = recs + gid * V5_REC;
    RecRef q;
    make_record(geom + v, vscale[v], T, xt, R, C, pitch, true, rec, &q);
    const int rows = BP_TP / subs;
    int* bx = reinterpret_cast<int*>(rec + BP_REC);
    bool all = true;
    for (int sb = 0; sb < 2; ++sb) {
        if (sb < subs) {
            make_box5(q, sb * rows, sb * rows + rows - 1, R, C, pitch, le, kcap, bufbytes,
                      bx + V5_BOX * sb);
            all = all && bx[V5_BOX * sb + 1] > 0;
        } else {
            for (int i = 0; i < V5_BOX; ++i) bx[V5_BOX * sb + i] = 0;
        }
    }
    if (!all) fit[ti] = 0;     // some view of this tile does not fit: direct gathers
}

__global__ void set_int_kernel(int* __restrict__ p, int n, int value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

template <typename ST, int K>
struct BoxStage {
    uint4 val[K];
    int lds[K];          // LDS byte offset of chunk k, -1 when the chunk is outside the box
};

template <typename ST, int PPT>
__global__ __launch_bounds__(BP_THREADS) void bp_tiles_kernel5(
    const ST* __restrict__ ring, int64_t view_stride, int pitch,
    const float* __restrict__ recs, const mr_task* __restrict__ tasks,
    const Tile* __restrict__ tiles, int n_views, float scale, float* __restrict__ out, int xcd,
    const int* __restrict__ fit) {
    constexpr int LE = Taps<ST>::LOG_ESZ;
    using Cfg = V5Cfg<LE>;
    constexpr int K = Cfg::K;
    constexpr int PSTEP = BP_THREADS / BP_TS;
    constexpr int SUBS = BP_TP / (PSTEP * PPT);
    static_assert(SUBS * PSTEP * PPT == BP_TP && SUBS <= 2, "tile shape");
    __shared__ __attribute__((aligned(16))) unsigned char stage[2][Cfg::BUF];
    const int lb = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int tile = lb / SUBS;
    const int sub = lb - tile * SUBS;
    const Tile tl = tiles[tile];
    const mr_task& T = tasks[tl.task];
    const int t = threadIdx.x;
    const int tx = t & (BP_TS - 1);
    const int ty = t / BP_TS + sub * PSTEP * PPT;
    const int p0 = tl.tp * BP_TP, s0 = tl.ts * BP_TS;
    const int s = s0 + tx;
    bool valid[PPT], inside[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        const int p = p0 + ty + PSTEP * m;
        valid[m] = (p < T.np) && (s < T.ns);
        inside[m] = valid[m] && (!T.mask_box || in_box(T, p, s));
    }
    const float ds = static_cast<float>(tx), dp0 = static_cast<float>(ty);
    const int pitchb = pitch << LE;
    const int view_bytes = static_cast<int>(view_stride << LE);
    const float* __restrict__ rt = recs + static_cast<int64_t>(tile) * n_views * V5_REC;
    const int boxo = BP_REC + V5_BOX * sub;
    float acc[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) acc[m] = 0.0f;

    auto rsrc = [&](int v) {
        return __builtin_amdgcn_make_buffer_rsrc(
            const_cast<ST*>(ring + static_cast<int64_t>(v) * view_stride), 0, view_bytes, 0x00020000);
    };
    // request the box of view v into registers
    auto load_box = [&](int v, BoxStage<ST, K>& st) {
        const int* bx = reinterpret_cast<const int*>(rt + static_cast<int64_t>(v) * V5_REC + boxo);
        const int goff = bx[0], nch = bx[2], strb = bx[3], tot = bx[6];
        const float inv = __int_as_float(bx[5]);
        const __amdgpu_buffer_rsrc_t rs = rsrc(v);
        // every lane loads (lanes past the box re-read the box origin): unconditional
        // loads keep the vector-memory counter exact, so the wait before staging view v
        // leaves view v+1's loads in flight
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int e = t + BP_THREADS * k;
            const bool in = e < tot;
            const int ee = in ? e : 0;
            const int row = static_cast<int>((static_cast<float>(ee) + 0.5f) * inv);
            const int cc = ee - row * nch;
            const int off = ((goff + row * pitch) << LE) + cc * 16;
            st.val[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0));
            st.lds[k] = in ? row * strb + cc * 16 : -1;
        }
    };
    auto store_box = [&](const BoxStage<ST, K>& st, unsigned char* __restrict__ buf) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (st.lds[k] >= 0) __builtin_memcpy(buf + st.lds[k], &st.val[k], 16);   // 4-byte aligned
    };
    auto coords = [&](const float* r, int m, float& fc, float& fr, float& wt, int& ri, int& ci) {
        const float dp = dp0 + static_cast<float>(PSTEP * m);
        const float uh = fmaf(dp, r[2], fmaf(ds, r[1], r[0]));
        const float vh = fmaf(dp, r[5], fmaf(ds, r[4], r[3]));
        const float wh = fmaf(dp, r[8], fmaf(ds, r[7], r[6]));
        const float rw = __builtin_amdgcn_rcpf(wh);
        const float col = __builtin_amdgcn_fmed3f(uh * rw, r[9], r[10]);
        const float row = __builtin_amdgcn_fmed3f(vh * rw, r[11], r[12]);
        wt = rw * rw;
        fc = __builtin_amdgcn_fractf(col);
        fr = __builtin_amdgcn_fractf(row);
        ci = cvt_flr(col);
        ri = cvt_flr(row);
    };
    auto compute = [&](int v, const unsigned char* __restrict__ buf) {
        const float* r = rt + static_cast<int64_t>(v) * V5_REC;
        const int* bx = reinterpret_cast<const int*>(r + boxo);
        const int strb = bx[3], koffb = bx[4];
#pragma unroll
        for (int m = 0; m < PPT; ++m) {
            float fc, fr, wt;
            int ri, ci;
            coords(r, m, fc, fr, wt, ri, ci);
            const int a = __mul24(ri, strb) + ci * (1 << LE) - koffb;
            const Pair<ST> tp = LdsPair<ST>::read(buf + a, strb);
            float top, bot;
            tp.lerp(fc, top, bot);
            acc[m] = fmaf(wt, fmaf(fr, bot - top, top), acc[m]);
        }
    };
    // direct gathers from the ring (tiles with a view whose box does not fit)
    auto compute_direct = [&](int v) {
        const float* r = rt + static_cast<int64_t>(v) * V5_REC;
        const __amdgpu_buffer_rsrc_t rs = rsrc(v);
        const int baseb = __float_as_int(r[14]) << LE;
#pragma unroll
        for (int m = 0; m < PPT; ++m) {
            float fc, fr, wt;
            int ri, ci;
            coords(r, m, fc, fr, wt, ri, ci);
            Pair<ST> tp;
            tp.load(rs, __mul24(ri, pitchb) + ((ci << LE) + baseb), pitchb);
            float top, bot;
            tp.lerp(fc, top, bot);
            acc[m] = fmaf(wt, fmaf(fr, bot - top, top), acc[m]);
        }
    };

    if (n_views > 0 && !fit[tile]) {
        for (int v = 0; v < n_views; ++v) compute_direct(v);
    } else if (n_views > 0) {
        BoxStage<ST, K> st0, st1;
        load_box(0, st0);
        load_box(min(1, n_views - 1), st1);
        // the prefetches run unconditionally (past the end they re-read the last view)
        // so that the wait before each staging counts exactly the loads issued after it
        // and the loop body has no conditional half (odd counts end in an epilogue)
        const int last = n_views - 1;
        int v = 0;
        for (; v + 1 < n_views; v += 2) {
            store_box(st0, stage[0]);
            load_box(min(v + 2, last), st0);
            __syncthreads();
            compute(v, stage[0]);
            store_box(st1, stage[1]);
            load_box(min(v + 3, last), st1);
            __syncthreads();
            compute(v + 1, stage[1]);
        }
        if (v < n_views) {
            store_box(st0, stage[0]);
            __syncthreads();
            compute(v, stage[0]);
        }
    }

    const int64_t obase = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q;
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        if (!valid[m]) continue;
        const int p = p0 + ty + PSTEP * m;
        out[obase + static_cast<int64_t>(p) * T.stride_p + static_cast<int64_t>(s) * T.stride_s] =
            inside[m] ? acc[m] * scale : 0.0f;
    }
}

// ---------------------------------------------------------------- adjoint and shift derivative
// Transpose of the tile back-projection with respect to the stored (dequantised) views:
// every point scatters scale * w_k * g_p with the forward's bilinear weights into
// gq[view][R][C]; taps outside the detector are dropped, exactly the transpose of the
// forward's zero taps (a clamped coordinate puts weight 1 on a zero tap and 0 on the
// other).  Records are built with storage scale 1.
template <bool CONE>
__global__ __launch_bounds__(BP_THREADS) void bp_adjoint_kernel(
    int pitch, int R, int C, const ViewGeom* __restrict__ geom, const mr_task* __restrict__ tasks,
    const Tile* __restrict__ tiles, int n_views, int vch, float scale,
    const float* __restrict__ g, float* __restrict__ gq) {
    __shared__ float4 rec4[BP_VPASS * BP_REC / 4];
    float* rec = reinterpret_cast<float*>(rec4);
    const Tile tl = tiles[blockIdx.x];
    const mr_task& T = tasks[tl.task];
    const int t = threadIdx.x, tx = t & (BP_TS - 1), ty = t / BP_TS;
    constexpr int PSTEP = BP_THREADS / BP_TS;
    constexpr int PPT = BP_TP / PSTEP;
    const int p0 = tl.tp * BP_TP, s0 = tl.ts * BP_TS;
    double xt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xt[d] = T.o[d] + tl.q * T.c[d] + p0 * T.a[d] + s0 * T.b[d];
    const int s = s0 + tx;
    const int64_t obase = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q;
    float gp[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        const int p = p0 + ty + PSTEP * m;
        const bool ok = (p < T.np) && (s < T.ns) && (!T.mask_box || in_box(T, p, s));
        gp[m] = ok ? g[obase + static_cast<int64_t>(p) * T.stride_p +
                       static_cast<int64_t>(s) * T.stride_s] * scale
                   : 0.0f;
    }
    const float ds = static_cast<float>(tx), dp0 = static_cast<float>(ty);
    const int64_t vstride = static_cast<int64_t>(R) * C;
    // workgroup (tile, y) takes views [y vch, (y + 1) vch)
    const int v_end = min(n_views, (static_cast<int>(blockIdx.y) + 1) * vch);
    for (int vb = static_cast<int>(blockIdx.y) * vch; vb < v_end; vb += BP_VPASS) {
        const int nv = min(BP_VPASS, v_end - vb);
        __syncthreads();
        if (t < nv) make_record(geom + vb + t, 1.0f, T, xt, R, C, pitch, CONE, rec + t * BP_REC);
        __syncthreads();
        for (int v = 0; v < nv; ++v) {
            const float4 r0 = rec4[v * 4 + 0], r1 = rec4[v * 4 + 1];
            const float4 r2 = rec4[v * 4 + 2], r3 = rec4[v * 4 + 3];
            const int base = __float_as_int(r3.z);
            // absolute detector coordinates of the reference: base = (rref+1)*pitch + cref+1
            const int rref = static_cast<int>(-1.0f - r2.w);        // rlo = -1 - rref
            const int cref = static_cast<int>(-1.0f - r2.y);        // clo = -1 - cref
            (void)base;
            float* gv = gq + static_cast<int64_t>(vb + v) * vstride;
#pragma unroll
            for (int m = 0; m < PPT; ++m) {
                if (gp[m] == 0.0f) continue;
                const float dp = dp0 + static_cast<float>(PSTEP * m);
                const float uh = fmaf(dp, r0.z, fmaf(ds, r0.y, r0.x));
                const float vh = fmaf(dp, r1.y, fmaf(ds, r1.x, r0.w));
                float col, row, wt;
                if (CONE) {
                    const float wh = fmaf(dp, r2.x, fmaf(ds, r1.w, r1.z));
                    const float rw = __builtin_amdgcn_rcpf(wh);
                    col = uh * rw; row = vh * rw; wt = rw * rw;
                } else {
                    col = uh; row = vh; wt = 1.0f;
                }
                col = __builtin_amdgcn_fmed3f(col, r2.y, r2.z);
                row = __builtin_amdgcn_fmed3f(row, r2.w, r3.x);
                const float fc = __builtin_amdgcn_fractf(col), fr = __builtin_amdgcn_fractf(row);
                const int c0 = cvt_flr(col) + cref, rr0 = cvt_flr(row) + rref;   // absolute
                const float sv = wt * gp[m];
                const float w00 = (1.0f - fr) * (1.0f - fc) * sv, w01 = (1.0f - fr) * fc * sv;
                const float w10 = fr * (1.0f - fc) * sv, w11 = fr * fc * sv;
                const bool c0ok = c0 >= 0 && c0 < C, c1ok = c0 + 1 >= 0 && c0 + 1 < C;
                if (rr0 >= 0 && rr0 < R) {
                    float* rowp = gv + static_cast<int64_t>(rr0) * C;
                    if (c0ok) atomicAdd(rowp + c0, w00);
                    if (c1ok) atomicAdd(rowp + c0 + 1, w01);
                }
                if (rr0 + 1 >= 0 && rr0 + 1 < R) {
                    float* rowp = gv + static_cast<int64_t>(rr0 + 1) * C;
                    if (c0ok) atomicAdd(rowp + c0, w10);
                    if (c1ok) atomicAdd(rowp + c0 + 1, w11);
                }
            }
        }
    }
}

// ---------------------------------------------------------------- adjoint as a gather
// The transpose of the tile back-projection without global atomics: a workgroup owns one
// ADJ_TH x ADJ_TW tile of one view's detector, accumulates in LDS the bilinear taps that
// fall into it from every point tile whose detector box meets it, and writes the tile
// once.  Per point the arithmetic is that of the scatter kernel above (the same FP32
// records, weights and clamps).  The sums are fixed point: an LDS float atomic add costs
// ~190 cycles per wave instruction on gfx950 against ~6 for a 64-bit integer add, so each
// tap is rounded to an integer multiple of 2^-k (k from a bound on every contribution the
// workgroup can receive, the contributions below 2^30 in those units) and added as a
// 64-bit integer, exact and independent of order; the rounding is at most 2^-31 of that
// bound per tap.  A workgroup whose bound is not finite (non-finite weights) adds floats.
constexpr int ADJ_TW = 64, ADJ_TH = 64;    // detector tile of a workgroup
constexpr int ADJ_LIST = 1024;             // point tiles tested per round
constexpr int kAdjGatherMaxTiles = 4096;   // larger requests: the scatter kernel

// gm[tile][p][s] (32 x 32 per tile) = g of the point times scale, 0 for points outside
// the task or the box; gmax[tile] = max |gm| over the tile (+inf if any is not finite).
__global__ __launch_bounds__(BP_THREADS) void adj_weights_kernel(
    const mr_task* __restrict__ tasks, const Tile* __restrict__ tiles, float scale,
    const float* __restrict__ g, float* __restrict__ gm, float* __restrict__ gmax) {
    const Tile tl = tiles[blockIdx.x];
    const mr_task& T = tasks[tl.task];
    const int t = threadIdx.x, tx = t & (BP_TS - 1), ty = t / BP_TS;
    constexpr int PSTEP = BP_THREADS / BP_TS;
    constexpr int PPT = BP_TP / PSTEP;
    const int p0 = tl.tp * BP_TP, s0 = tl.ts * BP_TS;
    const int s = s0 + tx;
    const int64_t obase = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q;
    float* out = gm + static_cast<int64_t>(blockIdx.x) * (BP_TP * BP_TS);
    float m = 0.0f;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int p = p0 + ty + PSTEP * k;
        const bool ok = (p < T.np) && (s < T.ns) && (!T.mask_box || in_box(T, p, s));
        const float x =
            ok ? g[obase + static_cast<int64_t>(p) * T.stride_p + static_cast<int64_t>(s) * T.stride_s] * scale
               : 0.0f;
        out[(ty + PSTEP * k) * BP_TS + tx] = x;
        m = isfinite(x) ? fmaxf(m, fabsf(x)) : INFINITY;
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    __shared__ float red[BP_THREADS / 64];
    if ((t & 63) == 0) red[t >> 6] = m;
    __syncthreads();
    if (t == 0) {
        float r = red[0];
        for (int w = 1; w < BP_THREADS / 64; ++w) r = fmaxf(r, red[w]);
        gmax[blockIdx.x] = r;
    }
}

// Absolute detector columns and rows of one point of a tile from its (tile, view) record,
// clamped as the back-projection clamps them; wt the FDK weight.
template <bool CONE>
__device__ __forceinline__ void adj_point(const float4& r0, const float4& r1, const float4& r2,
                                          const float4& r3, float dp, float ds, float& col,
                                          float& row, float& wt) {
    const float uh = fmaf(dp, r0.z, fmaf(ds, r0.y, r0.x));
    const float vh = fmaf(dp, r1.y, fmaf(ds, r1.x, r0.w));
    if (CONE) {
        const float wh = fmaf(dp, r2.x, fmaf(ds, r1.w, r1.z));
        const float rw = __builtin_amdgcn_rcpf(wh);
        col = uh * rw; row = vh * rw; wt = rw * rw;
    } else {
        col = uh; row = vh; wt = 1.0f;
    }
    col = __builtin_amdgcn_fmed3f(col, r2.y, r2.z);
    row = __builtin_amdgcn_fmed3f(row, r2.w, r3.x);
}

// boxes[view][tile] = (c_lo, c_hi, r_lo, r_hi): the detector cells (absolute, inclusive)
// the taps of the tile's points can reach in that view, from the images of its corners
// (a projective map takes the tile to the convex hull of those), one cell of margin for
// rounding; bound[view][tile] >= every tap weight the tile adds in that view (max |gm|
// times the largest FDK weight 1/wh^2, at a corner since wh is affine over the tile).
// The whole detector and an infinite bound when a corner lies at or behind the source.
template <bool CONE>
__global__ __launch_bounds__(256) void adj_boxes_kernel(const float* __restrict__ recs,
                                                        const float* __restrict__ gmax,
                                                        int ntiles, int n_views, int R, int C,
                                                        int4* __restrict__ boxes,
                                                        float* __restrict__ bound) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= static_cast<int64_t>(ntiles) * n_views) return;
    const int ti = static_cast<int>(gid / n_views);
    const int v = static_cast<int>(gid - static_cast<int64_t>(ti) * n_views);
    const float4* r4 = reinterpret_cast<const float4*>(recs + gid * BP_REC);
    const float4 r0 = r4[0], r1 = r4[1], r2 = r4[2], r3 = r4[3];
    const int rref = static_cast<int>(-1.0f - r2.w);
    const int cref = static_cast<int>(-1.0f - r2.y);
    float cmin = 1e30f, cmax = -1e30f, rmin = 1e30f, rmax = -1e30f, whmin = 1e30f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const float dp = (k & 1) ? static_cast<float>(BP_TP - 1) : 0.0f;
        const float ds = (k & 2) ? static_cast<float>(BP_TS - 1) : 0.0f;
        if (CONE) whmin = fminf(whmin, fmaf(dp, r2.x, fmaf(ds, r1.w, r1.z)));
        float col, row, wt;
        adj_point<CONE>(r0, r1, r2, r3, dp, ds, col, row, wt);
        cmin = fminf(cmin, col); cmax = fmaxf(cmax, col);
        rmin = fminf(rmin, row); rmax = fmaxf(rmax, row);
    }
    const bool ok = (!CONE || whmin > 0.0f) && cmin == cmin && cmax == cmax && rmin == rmin &&
                    rmax == rmax;
    int4 b;
    float bw = INFINITY;
    if (ok) {
        b.x = cvt_flr(cmin) + cref - 1;
        b.y = cvt_flr(cmax) + cref + 2;
        b.z = cvt_flr(rmin) + rref - 1;
        b.w = cvt_flr(rmax) + rref + 2;
        const float rw = CONE ? 1.0f / whmin : 1.0f;
        bw = gmax[ti] * (rw * rw) * 1.001f;          // margin for the FP32 rounding of wt
    } else {
        b = make_int4(0, C - 1, 0, R - 1);
    }
    const int64_t o = static_cast<int64_t>(v) * ntiles + ti;
    boxes[o] = b;
    bound[o] = bw;
}

__device__ __forceinline__ bool adj_meets(const int4& b, int X0, int Y0) {
    return (b.x < X0 + ADJ_TW) & (b.y >= X0) & (b.z < Y0 + ADJ_TH) & (b.w >= Y0);
}

// Workgroup (v, y, x): detector cells rows [y ADJ_TH, ..), columns [x ADJ_TW, ..) of view v.
template <bool CONE>
__global__ __launch_bounds__(BP_THREADS) void adj_gather_kernel(
    const float* __restrict__ recs, const int4* __restrict__ boxes,
    const float* __restrict__ bound, const float* __restrict__ gm, int ntiles, int n_views,
    int R, int C, float* __restrict__ gq) {
    __shared__ unsigned long long acc[ADJ_TH * ADJ_TW];   // fixed point, or floats in place
    __shared__ int list[ADJ_LIST];
    __shared__ int nlist;
    __shared__ unsigned bmax;
    float* const accf = reinterpret_cast<float*>(acc);
    const int v = blockIdx.x;
    const int Y0 = blockIdx.y * ADJ_TH, X0 = blockIdx.z * ADJ_TW;
    const int t = threadIdx.x, tx = t & (BP_TS - 1), ty = t / BP_TS;
    constexpr int PSTEP = BP_THREADS / BP_TS;
    constexpr int PPT = BP_TP / PSTEP;
    const int4* bv = boxes + static_cast<int64_t>(v) * ntiles;
    const float* bdv = bound + static_cast<int64_t>(v) * ntiles;
    for (int i = t; i < ADJ_TH * ADJ_TW; i += BP_THREADS) acc[i] = 0ull;
    if (t == 0) bmax = 0u;
    __syncthreads();
    // the largest contribution any tile meeting this detector tile can make
    unsigned mb = 0u;
    for (int i = t; i < ntiles; i += BP_THREADS) {
        const int4 b = bv[i];
        if (adj_meets(b, X0, Y0)) mb = max(mb, __float_as_uint(bdv[i]));   // bounds are >= 0
    }
    atomicMax(&bmax, mb);
    __syncthreads();
    const float B = __uint_as_float(bmax);
    bool fixed = B < INFINITY;                        // false: non-finite bound (or NaN)
    float fs = 1.0f;                                  // fixed-point scale 2^(30 - e), B < 2^e
    if (fixed && B > 0.0f) {
        int e;
        frexpf(B, &e);
        fixed = e >= -90 && e <= 150;                 // the scale stays a normal float
        fs = fixed ? ldexpf(1.0f, 30 - e) : 1.0f;
    }
    for (int base = 0; base < ntiles; base += ADJ_LIST) {
        if (t == 0) nlist = 0;
        __syncthreads();
        const int nb = min(ADJ_LIST, ntiles - base);
        for (int i = t; i < nb; i += BP_THREADS) {
            const int4 b = bv[base + i];
            if (adj_meets(b, X0, Y0)) list[atomicAdd(&nlist, 1)] = base + i;
        }
        __syncthreads();
        const int n = nlist;
        for (int e = 0; e < n; ++e) {
            const int ti = list[e];
            const float4* r4 = reinterpret_cast<const float4*>(
                recs + (static_cast<int64_t>(ti) * n_views + v) * BP_REC);
            const float4 r0 = r4[0], r1 = r4[1], r2 = r4[2], r3 = r4[3];
            const int rref = static_cast<int>(-1.0f - r2.w);
            const int cref = static_cast<int>(-1.0f - r2.y);
            const float* gw = gm + static_cast<int64_t>(ti) * (BP_TP * BP_TS);
            float gp[PPT];
#pragma unroll
            for (int m = 0; m < PPT; ++m) gp[m] = gw[(ty + PSTEP * m) * BP_TS + tx];
#pragma unroll
            for (int m = 0; m < PPT; ++m) {
                if (gp[m] == 0.0f) continue;
                float col, row, wt;
                adj_point<CONE>(r0, r1, r2, r3, static_cast<float>(ty + PSTEP * m),
                                static_cast<float>(tx), col, row, wt);
                const float fc = __builtin_amdgcn_fractf(col), fr = __builtin_amdgcn_fractf(row);
                const int c0 = cvt_flr(col) + cref, rr0 = cvt_flr(row) + rref;   // absolute
                const float sv = wt * gp[m];
                const float w[4] = {(1.0f - fr) * (1.0f - fc) * sv, (1.0f - fr) * fc * sv,
                                    fr * (1.0f - fc) * sv, fr * fc * sv};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int rr = rr0 + (k >> 1), cc = c0 + (k & 1);
                    const int ly = rr - Y0, lx = cc - X0;
                    if (ly >= 0 && ly < ADJ_TH && lx >= 0 && lx < ADJ_TW && rr < R && cc < C) {
                        if (fixed) {
                            const long long q = __float2int_rn(w[k] * fs);
                            atomicAdd(&acc[ly * ADJ_TW + lx], static_cast<unsigned long long>(q));
                        } else {
                            atomicAdd(&accf[ly * ADJ_TW + lx], w[k]);
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    float* gv = gq + static_cast<int64_t>(v) * R * C;
    const double inv = 1.0 / static_cast<double>(fs);
    for (int i = t; i < ADJ_TH * ADJ_TW; i += BP_THREADS) {
        const int y = Y0 + i / ADJ_TW, x = X0 + i % ADJ_TW;
        if (y < R && x < C)
            gv[static_cast<int64_t>(y) * C + x] =
                fixed ? static_cast<float>(static_cast<double>(static_cast<long long>(acc[i])) * inv)
                      : accf[i];
    }
}

// d(out)/d(cor_shift) of the tile back-projection: the shift adds s to the column and
// s * (u . vv) to the row of every tap (recon.py:420), so per view the derivative of
// the interpolant is dbil/dcol + (u . vv) dbil/drow, zero where the coordinate is
// clamped outside the detector.
//
// part == nullptr: one workgroup per tile (grid.x) over all views, scaled sums into out.
// Else workgroup (tile, k = blockIdx.y) sums views [k vper, (k + 1) vper) and leaves the
// unscaled sums of its tile's 32 x 32 points in part[k][tile] for part_sum_kernel: small
// requests (few tiles) then fill the device.
template <bool CONE, typename ST>
__global__ __launch_bounds__(BP_THREADS) void bp_dshift_kernel(
    const ST* __restrict__ ring, int64_t view_stride, int pitch, int R, int C,
    const ViewGeom* __restrict__ geom, const float* __restrict__ vscale,
    const mr_task* __restrict__ tasks, const Tile* __restrict__ tiles, int n_views, float scale,
    float* __restrict__ out, int vper, float* __restrict__ part) {
    __shared__ float4 rec4[BP_VPASS * BP_REC / 4];
    float* rec = reinterpret_cast<float*>(rec4);
    constexpr int LE = Taps<ST>::LOG_ESZ;
    const Tile tl = tiles[blockIdx.x];
    const mr_task& T = tasks[tl.task];
    const int t = threadIdx.x, tx = t & (BP_TS - 1), ty = t / BP_TS;
    constexpr int PSTEP = BP_THREADS / BP_TS;
    constexpr int PPT = BP_TP / PSTEP;
    const int p0 = tl.tp * BP_TP, s0 = tl.ts * BP_TS;
    double xt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) xt[d] = T.o[d] + tl.q * T.c[d] + p0 * T.a[d] + s0 * T.b[d];
    const int s = s0 + tx;
    bool valid[PPT], inside[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        const int p = p0 + ty + PSTEP * m;
        valid[m] = (p < T.np) && (s < T.ns);
        inside[m] = valid[m] && (!T.mask_box || in_box(T, p, s));
    }
    const float ds = static_cast<float>(tx), dp0 = static_cast<float>(ty);
    const int pitchb = pitch << LE;
    const int view_bytes = static_cast<int>(view_stride << LE);
    float acc[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) acc[m] = 0.0f;
    const int v0 = part ? static_cast<int>(blockIdx.y) * vper : 0;
    const int v1 = part ? min(n_views, v0 + vper) : n_views;
    for (int vb = v0; vb < v1; vb += BP_VPASS) {
        const int nv = min(BP_VPASS, v1 - vb);
        __syncthreads();
        if (t < nv)
            make_record(geom + vb + t, vscale[vb + t], T, xt, R, C, pitch, CONE, rec + t * BP_REC);
        __syncthreads();
        for (int v = 0; v < nv; ++v) {
            const float4 r0 = rec4[v * 4 + 0], r1 = rec4[v * 4 + 1];
            const float4 r2 = rec4[v * 4 + 2], r3 = rec4[v * 4 + 3];
            const int baseb = __float_as_int(r3.z) << LE;
            const float uvv = r3.w;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<ST*>(ring + static_cast<int64_t>(vb + v) * view_stride), 0, view_bytes,
                0x00020000);
#pragma unroll
            for (int m = 0; m < PPT; ++m) {
                const float dp = dp0 + static_cast<float>(PSTEP * m);
                const float uh = fmaf(dp, r0.z, fmaf(ds, r0.y, r0.x));
                const float vh = fmaf(dp, r1.y, fmaf(ds, r1.x, r0.w));
                float colr, rowr, wt;
                if (CONE) {
                    const float wh = fmaf(dp, r2.x, fmaf(ds, r1.w, r1.z));
                    const float rw = __builtin_amdgcn_rcpf(wh);
                    colr = uh * rw; rowr = vh * rw; wt = rw * rw;
                } else {
                    colr = uh; rowr = vh; wt = r3.y;
                }
                const float col = __builtin_amdgcn_fmed3f(colr, r2.y, r2.z);
                const float row = __builtin_amdgcn_fmed3f(rowr, r2.w, r3.x);
                const float fc = __builtin_amdgcn_fractf(col), fr = __builtin_amdgcn_fractf(row);
                const int voff = __mul24(cvt_flr(row), pitchb) + ((cvt_flr(col) << LE) + baseb);
                float a, b, c, d;
                Taps<ST>::load(rs, voff, pitchb, a, b, c, d);
                const float ddc = (col == colr) ? fmaf(fr, (d - c) - (b - a), b - a) : 0.0f;
                const float ddr = (row == rowr) ? fmaf(fc, (d - b) - (c - a), c - a) : 0.0f;
                acc[m] = fmaf(wt, fmaf(uvv, ddr, ddc), acc[m]);
            }
        }
    }
    if (part) {
        float* pt = part + (static_cast<int64_t>(blockIdx.y) * gridDim.x + blockIdx.x) * (BP_TP * BP_TS);
#pragma unroll
        for (int m = 0; m < PPT; ++m)
            if (valid[m]) pt[(ty + PSTEP * m) * BP_TS + tx] = acc[m];
        return;
    }
    const int64_t obase = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q;
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        if (!valid[m]) continue;
        const int p = p0 + ty + PSTEP * m;
        out[obase + static_cast<int64_t>(p) * T.stride_p + static_cast<int64_t>(s) * T.stride_s] =
            inside[m] ? acc[m] * scale : 0.0f;
    }
}

// Point-list back-projection with float64 coordinates (reconstruct_points): each
// thread owns one point; per view the projective form is evaluated in FP64, the
// fractional parts and FDK weight are rounded to FP32 for the interpolation.
template <bool CONE, typename ST>
__global__ __launch_bounds__(256) void bp_points_kernel(
    const ST* __restrict__ ring, int64_t view_stride, int pitch, int R, int C,
    const ViewGeom* __restrict__ geom, const float* __restrict__ vscale,
    const double* __restrict__ pts, int64_t M, double shift, int n_views, float scale,
    float* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const bool ok = i < M;
    const double x = ok ? pts[3 * i + 0] : 0.0;
    const double y = ok ? pts[3 * i + 1] : 0.0;
    const double z = ok ? pts[3 * i + 2] : 0.0;
    const double cc = 0.5 * (C - 1), cr = 0.5 * (R - 1);
    float acc = 0.0f;
    for (int k = 0; k < n_views; ++k) {
        const ViewGeom* g = geom + k;
        const double wh = CONE ? g->P[2][0] * x + g->P[2][1] * y + g->P[2][2] * z + g->P[2][3]
                               : 1.0;
        const double uh = g->P[0][0] * x + g->P[0][1] * y + g->P[0][2] * z + g->P[0][3] + shift * wh;
        const double vh = g->P[1][0] * x + g->P[1][1] * y + g->P[1][2] * z + g->P[1][3] +
                          shift * g->uvv * wh;
        const double rw = CONE ? 1.0 / wh : 1.0;
        double col = uh * rw + cc, row = vh * rw + cr;
        col = fmin(fmax(col, -1.0), static_cast<double>(C));
        row = fmin(fmax(row, -1.0), static_cast<double>(R));
        const double cf = floor(col), rf = floor(row);
        const float fc = static_cast<float>(col - cf), fr = static_cast<float>(row - rf);
        const uint32_t idx = static_cast<uint32_t>((static_cast<int>(rf) + 1) * pitch +
                                                   static_cast<int>(cf) + 1);
        const ST* __restrict__ q0 = ring + static_cast<int64_t>(k) * view_stride;
        const ST* __restrict__ q1 = q0 + pitch;
        const float a = Store<ST>::ld(q0, idx), b = Store<ST>::ld(q0, idx + 1);
        const float c = Store<ST>::ld(q1, idx), d = Store<ST>::ld(q1, idx + 1);
        const float top = fmaf(fc, b - a, a);
        const float bot = fmaf(fc, d - c, c);
        const float wt = CONE ? static_cast<float>(rw * rw) * vscale[k] : vscale[k];
        acc = fmaf(wt, fmaf(fr, bot - top, top), acc);
    }
    if (ok) out[i] = acc * scale;
}

// Dequantised read-back of stored views (parity hook): out[count][R][C] float32.
template <typename ST>
__global__ void read_views_kernel(const ST* __restrict__ ring, int64_t view_stride, int pitch,
                                  int R, int C, const float* __restrict__ vscale, int first,
                                  float* __restrict__ out) {
    const int v = blockIdx.y;
    const int r = blockIdx.x;
    const ST* src = ring + static_cast<int64_t>(first + v) * view_stride +
                    static_cast<int64_t>(r + 1) * pitch + 1;
    const float s = vscale[first + v];
    float* dst = out + (static_cast<int64_t>(v) * R + r) * C;
    for (int c = threadIdx.x; c < C; c += blockDim.x) dst[c] = Store<ST>::ld(src, c) * s;
}

template <bool CONE>
int dispatch_tiles(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles, int n_views,
                   float scale, float* d_out, hipStream_t s, const TileQueue& q) {
    c->kern_used = 2;
    // persistent launches size the grid to the resident capacity and pull tiles from q
    dim3 grid(q.head ? q.grid : ntiles), block(BP_THREADS);
    if (c->bp_variant == 2 && c->v2_ring == 1 && c->storage == MR_STORE_F32) {
        hipLaunchKernelGGL((bp_tiles_kernel2<CONE, float, 1>), grid, block, 0, s,
                           static_cast<const float*>(c->pring), pair_view_stride(c), c->pitch, c->R,
                           c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out,
                           q.head, q.count, q.n_views, c->xcd_map != 0, q.v_begin, q.accum,
                           q.accum_half);
        MR_HIP(hipGetLastError());
        return MR_OK;
    }
    if (c->bp_variant == 2 && c->v2_ring == 2) {
        const int64_t bvs = band_view_stride(c);
        if (c->storage == MR_STORE_F16)
            hipLaunchKernelGGL((bp_tiles_kernel2<CONE, __half, 2>), grid, block, 0, s,
                               static_cast<const __half*>(c->bring), bvs, c->pitch, c->R,
                               c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out,
                               q.head, q.count, q.n_views, c->xcd_map != 0, q.v_begin, q.accum,
                               q.accum_half);
        else
            hipLaunchKernelGGL((bp_tiles_kernel2<CONE, uint16_t, 2>), grid, block, 0, s,
                               static_cast<const uint16_t*>(c->bring), bvs, c->pitch, c->R,
                               c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out,
                               q.head, q.count, q.n_views, c->xcd_map != 0, q.v_begin, q.accum,
                               q.accum_half);
        MR_HIP(hipGetLastError());
        return MR_OK;
    }
    // kernel 2 on the ring itself: the v2 loop, and every launch not taken above (no
    // views, one tile per workgroup)
    {
        switch (c->storage) {
        case MR_STORE_F32:
            hipLaunchKernelGGL((bp_tiles_kernel2<CONE, float>), grid, block, 0, s,
                               static_cast<const float*>(c->ring), c->view_stride, c->pitch, c->R,
                               c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out,
                               q.head, q.count, q.n_views, c->xcd_map != 0, q.v_begin, q.accum,
                               q.accum_half);
            break;
        case MR_STORE_F16:
            hipLaunchKernelGGL((bp_tiles_kernel2<CONE, __half>), grid, block, 0, s,
                               static_cast<const __half*>(c->ring), c->view_stride, c->pitch, c->R,
                               c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out,
                               q.head, q.count, q.n_views, c->xcd_map != 0, q.v_begin, q.accum,
                               q.accum_half);
            break;
        default:
            hipLaunchKernelGGL((bp_tiles_kernel2<CONE, uint16_t>), grid, block, 0, s,
                               static_cast<const uint16_t*>(c->ring), c->view_stride, c->pitch,
                               c->R, c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale,
                               d_out, q.head, q.count, q.n_views, c->xcd_map != 0, q.v_begin, q.accum,
                               q.accum_half);
            break;
        }
        MR_HIP(hipGetLastError());
        return MR_OK;
    }
}

template <bool CONE>
int dispatch_points(Ctx* c, const double* d_pts, int64_t M, double shift, int n_views,
                    float scale, float* d_out, hipStream_t s) {
    dim3 grid(static_cast<unsigned>((M + 255) / 256)), block(256);
    switch (c->storage) {
    case MR_STORE_F32:
        hipLaunchKernelGGL((bp_points_kernel<CONE, float>), grid, block, 0, s,
                           static_cast<const float*>(c->ring), c->view_stride, c->pitch, c->R, c->C,
                           c->d_geom, c->d_vscale, d_pts, M, shift, n_views, scale, d_out);
        break;
    case MR_STORE_F16:
        hipLaunchKernelGGL((bp_points_kernel<CONE, __half>), grid, block, 0, s,
                           static_cast<const __half*>(c->ring), c->view_stride, c->pitch, c->R,
                           c->C, c->d_geom, c->d_vscale, d_pts, M, shift, n_views, scale, d_out);
        break;
    default:
        hipLaunchKernelGGL((bp_points_kernel<CONE, uint16_t>), grid, block, 0, s,
                           static_cast<const uint16_t*>(c->ring), c->view_stride, c->pitch, c->R,
                           c->C, c->d_geom, c->d_vscale, d_pts, M, shift, n_views, scale, d_out);
        break;
    }
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace

int launch_bp_tiles_v4(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles, int n_views,
                       float scale, float* d_out, hipStream_t s);
int64_t band_view_stride(const Ctx* c);
int ensure_band_ring(Ctx* c, int n_views, hipStream_t s);

int launch_bp_tiles(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles, int n_views,
                    float scale, float* d_out, hipStream_t s, const TileQueue* q) {
    const TileQueue none{};
    if (!q && ntiles <= 0) return MR_OK;
    if (!q && c->bp_variant >= 4 && n_views > 0) {
        if (!c->bp_explicit && c->bp_variant == 7 && !c->req_slices) {
            // default configuration, volume request (stacks of z-planes): a z-plane tile
            // touches only a few detector rows per view, where the row-pair ring (32
            // columns per 128-byte line) beats the band ring (16): kernel 6, 4 points,
            // 2 views in flight
            const int sv = c->bp_variant, sp = c->bp_ppt, sd = c->bp_depth;
            c->bp_variant = 6;
            c->bp_ppt = 4;
            c->bp_depth = 2;
            const int rc = launch_bp_tiles_v4(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s);
            c->bp_variant = sv;
            c->bp_ppt = sp;
            c->bp_depth = sd;
            return rc;
        }
        if (!c->bp_explicit && (c->bp_variant == 7 || c->bp_variant == 6 || c->bp_variant == 4) && c->req_slices &&
            static_cast<size_t>(ntiles) * n_views * BP_REC * sizeof(float) > c->rec_budget) {
            // default configuration, slice requests whose records exceed one pass of the
            // records buffer (a COR search of 25 trial images of 1024^2 over 1800 views
            // needs 2.9 GB): the v2 loop, which builds each tile's records in LDS and keeps
            // every view's sums in registers, on the band copy of a 16-bit ring, the
            // row-pair copy of an f32 parallel-beam ring or the f32 ring itself, instead
            // of records in chunks and view blocks whose sums make a round trip through
            // the 105 MB output per launch (measured on that search: f16 38 against 54
            // ms, f32 parallel 38 against 82 ms)
            const int rc0 = prepare_persistent_bp(c, n_views, s);
            if (rc0) return rc0;
            const int rk = v2_ring_kind(c);
            const int sv = c->bp_variant;
            c->bp_variant = 2;
            c->v2_ring = rk;
            const TileQueue none{};
            const int rc = c->kind == MR_KIND_CONE
                               ? dispatch_tiles<true>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s, none)
                               : dispatch_tiles<false>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s, none);
            c->bp_variant = sv;
            c->v2_ring = 0;
            return rc;
        }
        return launch_bp_tiles_v4(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s);
    }
    if (q) {                               // persistent launches always run the v2 loop
        const int saved = c->bp_variant;
        c->v2_ring = v2_ring_kind(c);
        c->bp_variant = 2;
        const int rc = c->kind == MR_KIND_CONE
                           ? dispatch_tiles<true>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s, *q)
                           : dispatch_tiles<false>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s, *q);
        c->bp_variant = saved;
        c->v2_ring = 0;
        return rc;
    }
    return c->kind == MR_KIND_CONE
               ? dispatch_tiles<true>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s, q ? *q : none)
               : dispatch_tiles<false>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s, q ? *q : none);
}

__global__ __launch_bounds__(256) void zero_tiles_kernel(const mr_task* __restrict__ tasks,
                                                        const Tile* __restrict__ tiles,
                                                        float* __restrict__ out) {
    const Tile tl = tiles[blockIdx.x];
    const mr_task& T = tasks[tl.task];
    const int s = tl.ts * BP_TS + (threadIdx.x & (BP_TS - 1));
    const int64_t obase = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q;
    for (int r = threadIdx.x / BP_TS; r < BP_TP; r += 256 / BP_TS) {
        const int p = tl.tp * BP_TP + r;
        if (p < T.np && s < T.ns)
            out[obase + static_cast<int64_t>(p) * T.stride_p + static_cast<int64_t>(s) * T.stride_s] =
                0.0f;
    }
}

int launch_zero_tiles(const mr_task* d_tasks, const Tile* d_tiles, int ntiles, float* d_out,
                      hipStream_t s) {
    if (ntiles <= 0) return MR_OK;
    hipLaunchKernelGGL(zero_tiles_kernel, dim3(ntiles), dim3(256), 0, s, d_tasks, d_tiles, d_out);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

// Device work queue: one thread per candidate tile of every task (prefix offsets per
// task), the conservative box test of the host path, block-ordered appends.
__global__ __launch_bounds__(256) void enumerate_tiles_kernel(const mr_task* __restrict__ tasks,
                                                              const int64_t* __restrict__ toff,
                                                              int ntasks, Tile* __restrict__ queue,
                                                              int* __restrict__ count) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int64_t total = toff[ntasks];
    bool hit = false;
    Tile tl{0, 0, 0, 0};
    if (gid < total) {
        int lo = 0, hi = ntasks - 1;                // last task with toff[task] <= gid
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (toff[mid] <= gid) lo = mid; else hi = mid - 1;
        }
        const mr_task& T = tasks[lo];
        const int ntp = (T.np + BP_TP - 1) / BP_TP, nts = (T.ns + BP_TS - 1) / BP_TS;
        const int64_t loc = gid - toff[lo];
        const int q = static_cast<int>(loc / (static_cast<int64_t>(ntp) * nts));
        const int rem = static_cast<int>(loc - static_cast<int64_t>(q) * ntp * nts);
        const int tp = rem / nts, ts = rem - tp * nts;
        tl = Tile{lo, q, tp, ts};
        hit = true;
        if (T.mask_box) {
            const int p0 = tp * BP_TP, s0 = ts * BP_TS;
            const int p1 = min(p0 + BP_TP, T.np) - 1, s1 = min(s0 + BP_TS, T.ns) - 1;
            for (int d = 0; d < 3 && hit; ++d) {
                const double lim = T.box[d] * (1.0 + 1e-9) + 1e-9;
                bool all_hi = true, all_lo = true;
                for (int cp = 0; cp < 2; ++cp)
                    for (int cs = 0; cs < 2; ++cs) {
                        const int p = cp ? p1 : p0, sx = cs ? s1 : s0;
                        const double x = T.o[d] + q * T.c[d] + p * T.a[d] + sx * T.b[d];
                        all_hi = all_hi && (x > lim);
                        all_lo = all_lo && (x < -lim);
                    }
                if (all_hi || all_lo) hit = false;
            }
        }
    }
    // block-level prefix of the hits, one atomic per block
    __shared__ int s_pos[256];
    __shared__ int s_base;
    s_pos[threadIdx.x] = hit ? 1 : 0;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = threadIdx.x >= off ? s_pos[threadIdx.x - off] : 0;
        __syncthreads();
        s_pos[threadIdx.x] += v;
        __syncthreads();
    }
    if (threadIdx.x == 255) s_base = atomicAdd(count, s_pos[255]);
    __syncthreads();
    if (hit) queue[s_base + s_pos[threadIdx.x] - 1] = tl;
}

__global__ void fill_int_kernel(int* __restrict__ p, int n, int value) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

int launch_enumerate_tiles(const mr_task* d_tasks, const int64_t* d_toff, int ntasks,
                           int64_t total, Tile* d_queue, int* d_counters, hipStream_t s) {
    // counters: [0] queue length, [1] head; cleared with a kernel (graph-safe)
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, d_counters, 2, 0);
    if (total > 0)
        hipLaunchKernelGGL(enumerate_tiles_kernel, dim3(static_cast<unsigned>((total + 255) / 256)),
                           dim3(256), 0, s, d_tasks, d_toff, ntasks, d_queue, d_counters);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int persistent_grid(Ctx* c) {
    int cus = 0, per = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, bp_tiles_kernel2<true, float>,
                                                     BP_THREADS, 0) != hipSuccess || per < 1)
        per = 4;
    return cus * per;
}

int launch_reset_head(int* d_counters, hipStream_t s) {
    hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, s, d_counters + 1, 1, 0);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

int launch_bp_points(Ctx* c, const double* d_pts, int64_t M, double cor_shift, int n_views,
                     float scale, float* d_out, hipStream_t s) {
    if (M <= 0) return MR_OK;
    return c->kind == MR_KIND_CONE
               ? dispatch_points<true>(c, d_pts, M, cor_shift, n_views, scale, d_out, s)
               : dispatch_points<false>(c, d_pts, M, cor_shift, n_views, scale, d_out, s);
}

// Sums of the view parts of a split launch, in part order: out = scale * sum over k of
// part[k][tile] for every point of every tile (0 outside a masked plane's box).  A thread
// adds 4 consecutive points of a tile row with 16-byte loads, four parts' loads issued
// before their additions (the additions stay in part order).
__global__ __launch_bounds__(256) void part_sum_kernel(const mr_task* __restrict__ tasks,
                                                      const Tile* __restrict__ tiles,
                                                      const float* __restrict__ part, int nparts,
                                                      float scale, float* __restrict__ out) {
    static_assert(BP_TP * BP_TS == 4 * 256 && BP_TS % 4 == 0, "4 points per thread");
    const Tile tl = tiles[blockIdx.x];
    const mr_task& T = tasks[tl.task];
    const int i = 4 * static_cast<int>(threadIdx.x);
    const int p = tl.tp * BP_TP + i / BP_TS, s0 = tl.ts * BP_TS + i % BP_TS;
    if (p >= T.np || s0 >= T.ns) return;
    const float4* __restrict__ src =
        reinterpret_cast<const float4*>(part + static_cast<int64_t>(blockIdx.x) * (BP_TP * BP_TS) + i);
    const int64_t ps = static_cast<int64_t>(gridDim.x) * (BP_TP * BP_TS / 4);   // float4s per part
    float4 sum = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    auto add = [&](const float4 a) {
        sum.x += a.x;
        sum.y += a.y;
        sum.z += a.z;
        sum.w += a.w;
    };
    int k = 0;
    for (; k + 4 <= nparts; k += 4) {
        const float4 a = src[k * ps], b = src[(k + 1) * ps], c = src[(k + 2) * ps], d = src[(k + 3) * ps];
        add(a);
        add(b);
        add(c);
        add(d);
    }
    for (; k < nparts; ++k) add(src[k * ps]);
    const float v[4] = {sum.x, sum.y, sum.z, sum.w};
    const int64_t ob = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q + static_cast<int64_t>(p) * T.stride_p;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int s = s0 + j;
        if (s < T.ns) {
            const bool inside = !T.mask_box || in_box(T, p, s);
            out[ob + static_cast<int64_t>(s) * T.stride_s] = inside ? v[j] * scale : 0.0f;
        }
    }
}

// Split launches keep their part sums within this many bytes (else one part, or for the
// tile kernels the view-block launches).
constexpr size_t kPartBudget = size_t{1} << 30;

template <bool CONE, typename ST, int PPT, int LAY, int DEPTH, int RING, bool SYNC>
static void launch_k4(Ctx* c, const ST* ring, int64_t vs, const float* recs, const mr_task* d_tasks,
                      const Tile* d_tiles, int nt, int n_views, float scale, float* d_out,
                      hipStream_t s) {
    const int nb = Map4<PPT, LAY>::SUBS * nt;
    // one launch of ceil(n_views / bp_vchunk) copies of the grid, copy k
    // summing views [k bp_vchunk, (k + 1) bp_vchunk) of every tile, then the part sums
    // added in part order.  Small requests fill the device this way, and large ones run
    // faster than one launch per view block with the sums carried through `out` (full_cone
    // 3 x 1024^2 at 1200 views, f16: 2.9 against 3.4 ms; 3 x 128^2: 0.88 against 1.34 ms).
    // Every request whose part sums fit kPartBudget uses the same partition of the views,
    // so a plane gets the same values alone or within any batch.
    if (c->bp_vchunk > 0 && !c->vchunk_any && n_views >= 2 * c->bp_vchunk) {
        const int kv = c->bp_vchunk;
        const int parts = (n_views + kv - 1) / kv;
        const size_t need = sizeof(float) * BP_TP * BP_TS * static_cast<size_t>(nt) * parts;
        if (need <= kPartBudget && ensure(&c->d_part, &c->part_bytes, need) == MR_OK) {
            float* part = static_cast<float*>(c->d_part);
            hipLaunchKernelGGL((bp_tiles_kernel4<CONE, ST, PPT, LAY, DEPTH, RING, SYNC>), dim3(nb * parts),
                               dim3(BP_THREADS), 0, s, ring, vs, c->pitch, recs, d_tasks, d_tiles,
                               n_views, scale, d_out, c->xcd_map != 0 ? 1 : 0, 0, n_views, part,
                               parts, kv);
            hipLaunchKernelGGL(part_sum_kernel, dim3(nt), dim3(256), 0, s, d_tasks, d_tiles, part,
                               parts, scale, d_out);
            ++c->tile_launches;
            return;
        }
    }
    // cone beam, requests beyond the part budget (or MINIRECAST_BP_VCHUNK_ANY): for slice
    // requests whose grid fills the device, one launch per block of bp_vchunk views, so
    // that all tiles read the same few views at a time and share detector lines in the L2,
    // at the cost of a round trip of the sums through `out` per launch.  Smaller grids lose
    // more to the launches than they gain; volume requests (stacks of planes, whose tiles
    // share lines plane to plane already) would only pay the round trips.
    const bool split = CONE && c->bp_vchunk > 0 &&
                       (c->vchunk_any || (c->req_slices && nb >= 8 * c->n_cus));
    const int k = split ? c->bp_vchunk : std::max(n_views, 1);
    int vb = 0;
    do {
        const int ve = std::min(n_views, vb + k);
        hipLaunchKernelGGL((bp_tiles_kernel4<CONE, ST, PPT, LAY, DEPTH, RING, SYNC>), dim3(nb),
                           dim3(BP_THREADS), 0, s, ring, vs, c->pitch, recs, d_tasks, d_tiles,
                           n_views, scale, d_out, c->xcd_map != 0 ? 1 : 0, vb, ve, nullptr, 1, 0);
        ++c->tile_launches;
        vb = ve;
    } while (vb < n_views);
}

// Kernels 4, 6 and 7: the storage-independent switches (ring layout, lockstep waves, views
// in flight) as template arguments.  Kernel 7 reads the band ring for 16-bit storage and
// the row-pair ring for f32.
template <bool CONE, typename ST, int PPT, int LAY>
static void launch_v4_one(Ctx* c, const float* recs, const mr_task* d_tasks, const Tile* d_tiles,
                          int nt, int n_views, float scale, float* d_out, hipStream_t s) {
    constexpr bool B16 = sizeof(ST) == 2;
    const int ring_kind = (B16 && c->bp_variant == 7) ? 2 : c->bp_variant >= 6 ? 1 : 0;
    const ST* ring = static_cast<const ST*>(ring_kind == 2 ? c->bring : ring_kind == 1 ? c->pring : c->ring);
    const int64_t vs = ring_kind == 2 ? band_view_stride(c)
                     : ring_kind == 1 ? 2 * static_cast<int64_t>(c->R + 2) * c->pitch : c->view_stride;
    const bool d3 = c->bp_depth >= 3;
#define MR_K4(D, G, S) launch_k4<CONE, ST, PPT, LAY, D, G, S>(c, ring, vs, recs, d_tasks, d_tiles, nt, n_views, scale, d_out, s)
    if (ring_kind == 2) {
        if constexpr (B16) {
            if (c->bp_sync) d3 ? MR_K4(3, 2, true) : MR_K4(2, 2, true);
            else d3 ? MR_K4(3, 2, false) : MR_K4(2, 2, false);
        }
    } else if (ring_kind == 1 && c->bp_sync) d3 ? MR_K4(3, 1, true) : MR_K4(2, 1, true);
    else if (ring_kind == 1) d3 ? MR_K4(3, 1, false) : MR_K4(2, 1, false);
    else if (c->bp_sync) d3 ? MR_K4(3, 0, true) : MR_K4(2, 0, true);
    else d3 ? MR_K4(3, 0, false) : MR_K4(2, 0, false);
#undef MR_K4
}

template <bool CONE, typename ST>
static void launch_v4_kernel(Ctx* c, const float* recs, const mr_task* d_tasks, const Tile* d_tiles,
                             int nt, int n_views, float scale, float* d_out, hipStream_t s) {
    const int lay = c->bp_layout;
    if (c->bp_ppt == 1)
        lay ? launch_v4_one<CONE, ST, 1, 1>(c, recs, d_tasks, d_tiles, nt, n_views, scale, d_out, s)
            : launch_v4_one<CONE, ST, 1, 0>(c, recs, d_tasks, d_tiles, nt, n_views, scale, d_out, s);
    else if (c->bp_ppt == 2)
        lay ? launch_v4_one<CONE, ST, 2, 1>(c, recs, d_tasks, d_tiles, nt, n_views, scale, d_out, s)
            : launch_v4_one<CONE, ST, 2, 0>(c, recs, d_tasks, d_tiles, nt, n_views, scale, d_out, s);
    else
        lay ? launch_v4_one<CONE, ST, 4, 1>(c, recs, d_tasks, d_tiles, nt, n_views, scale, d_out, s)
            : launch_v4_one<CONE, ST, 4, 0>(c, recs, d_tasks, d_tiles, nt, n_views, scale, d_out, s);
}

template <typename ST>
static void launch_v5_kernel(Ctx* c, const float* recs, const int* fit, const mr_task* d_tasks,
                             const Tile* d_tiles, int nt, int n_views, float scale, float* d_out,
                             hipStream_t s) {
    if (c->bp_ppt == 2)
        hipLaunchKernelGGL((bp_tiles_kernel5<ST, 2>), dim3(2 * nt), dim3(BP_THREADS), 0, s,
                           static_cast<const ST*>(c->ring), c->view_stride, c->pitch, recs, d_tasks,
                           d_tiles, n_views, scale, d_out, c->xcd_map != 0, fit);
    else
        hipLaunchKernelGGL((bp_tiles_kernel5<ST, 4>), dim3(nt), dim3(BP_THREADS), 0, s,
                           static_cast<const ST*>(c->ring), c->view_stride, c->pitch, recs, d_tasks,
                           d_tiles, n_views, scale, d_out, c->xcd_map != 0, fit);
}

// Ring copies.  The row-pair copy (kernel 6, twice the ring) and the band copy (kernel 7,
// 4/3 of a 16-bit ring) are caches of the ring, built on first use and extended to the
// views stored since (slots below pring_n / bring_n never change until the views are
// reset).  A copy is built only if the device keeps copy_reserve bytes free after it and
// the context's copies stay within copy_budget; when only freeing the other copy (not
// read by the request at hand) makes room, that copy is dropped first, and when nothing
// does the request runs without it (kernel 6 -> 7 -> 4 on the ring, the v2 loop on the
// ring).  A drop bumps copies_epoch, so graphs captured over the dropped copy's address
// are captured again.
constexpr int kNoRoom = 1;   // ensure_*_ring: not built for lack of memory (not an error)

size_t copies_held(const Ctx* c) {
    return (c->pring ? c->pring_bytes : 0) + (c->bring ? c->bring_bytes : 0);
}

namespace {

// 0: room for `bytes` now; 1: room once the `freeable` bytes of the other copy are
// freed; -1: no room
int copy_room(const Ctx* c, size_t bytes, size_t freeable) {
    int r = 0;
    if (c->copy_budget >= 0) {
        const size_t held = copies_held(c), budget = static_cast<size_t>(c->copy_budget);
        if (held + bytes > budget) r = held - freeable + bytes <= budget ? 1 : -1;
    }
    size_t fr = 0, total = 0;
    if (r >= 0 && hipMemGetInfo(&fr, &total) == hipSuccess) {
        const size_t need = bytes + c->copy_reserve;
        if (fr < need) r = fr + freeable >= need ? 1 : -1;
    }
    return r;
}

void drop_pair_ring(Ctx* c, hipStream_t s) {
    if (!c->pring) return;
    (void)hipStreamSynchronize(s);
    (void)hipFree(c->pring);
    c->pring = nullptr;
    c->pring_bytes = 0;
    c->pring_n = 0;
    ++c->copies_epoch;
}

void drop_band_ring(Ctx* c, hipStream_t s) {
    if (!c->bring) return;
    (void)hipStreamSynchronize(s);
    (void)hipFree(c->bring);
    c->bring = nullptr;
    c->bring_bytes = 0;
    c->bring_n = 0;
    ++c->copies_epoch;
}

// Allocates one copy of `bytes` at *dst (other: the other copy, dropped if that makes
// room).  MR_OK, kNoRoom, or an error.
int alloc_copy(Ctx* c, void** dst, size_t bytes, bool other_is_band, bool* no_room,
               hipStream_t s) {
    const size_t freeable = other_is_band ? (c->bring ? c->bring_bytes : 0)
                                          : (c->pring ? c->pring_bytes : 0);
    const int room = copy_room(c, bytes, freeable);
    if (room < 0) {
        *no_room = true;
        return kNoRoom;
    }
    if (room == 1) other_is_band ? drop_band_ring(c, s) : drop_pair_ring(c, s);
    const hipError_t e = hipMalloc(dst, bytes);
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        *dst = nullptr;
        *no_room = true;
        return kNoRoom;
    }
    MR_HIP(e);
    *no_room = false;
    return MR_OK;
}

}  // namespace

void trim_ring_copies(Ctx* c, hipStream_t s) {
    if (c->copy_budget < 0) return;
    if (copies_held(c) > static_cast<size_t>(c->copy_budget)) drop_pair_ring(c, s);
    if (copies_held(c) > static_cast<size_t>(c->copy_budget)) drop_band_ring(c, s);
}

int ensure_pair_ring(Ctx* c, int n_views, hipStream_t s) {
    const size_t esz = c->storage == MR_STORE_F32 ? 4 : 2;
    const int64_t pvs = 2 * static_cast<int64_t>(c->R + 2) * c->pitch;
    if (!c->pring) {
        const size_t bytes = esz * static_cast<size_t>(pvs) * c->n_total;
        const int rc = alloc_copy(c, &c->pring, bytes, true, &c->pair_no_room, s);
        if (rc) return rc;
        c->pring_bytes = bytes;
        c->pring_n = 0;
    }
    c->pair_no_room = false;
    const int v1 = n_views;
    for (int v0 = c->pring_n; v0 < v1; v0 += 4096) {
        const int nv = std::min(4096, v1 - v0);
        const dim3 grid((c->pitch + 255) / 256, c->R + 2, nv);
        if (esz == 4)
            hipLaunchKernelGGL(pair_convert_kernel<float>, grid, dim3(256), 0, s,
                               static_cast<const float*>(c->ring), static_cast<float*>(c->pring),
                               c->view_stride, pvs, c->pitch, v0, nullptr);
        else
            hipLaunchKernelGGL(pair_convert_kernel<uint16_t>, grid, dim3(256), 0, s,
                               static_cast<const uint16_t*>(c->ring), static_cast<uint16_t*>(c->pring),
                               c->view_stride, pvs, c->pitch, v0, nullptr);
        MR_HIP(hipGetLastError());
    }
    c->pring_n = std::max(c->pring_n, v1);
    return MR_OK;
}

// The band copy: bands of 4 padded rows stepping 3 rows.
int64_t band_view_stride(const Ctx* c) {
    return static_cast<int64_t>(band_count(c->R)) * c->pitch * 4;
}

int ensure_band_ring(Ctx* c, int n_views, hipStream_t s) {
    const int64_t bvs = band_view_stride(c);
    if (!c->bring) {
        const size_t bytes = sizeof(uint16_t) * static_cast<size_t>(bvs) * c->n_total;
        const int rc = alloc_copy(c, &c->bring, bytes, false, &c->band_no_room, s);
        if (rc) return rc;
        c->bring_bytes = bytes;
        c->bring_n = 0;
    }
    c->band_no_room = false;
    for (int v0 = c->bring_n; v0 < n_views; v0 += 4096) {
        const int nv = std::min(4096, n_views - v0);
        const dim3 grid((c->pitch + 255) / 256, band_count(c->R), nv);
        hipLaunchKernelGGL(band_convert_kernel, grid, dim3(256), 0, s,
                           static_cast<const uint16_t*>(c->ring), static_cast<uint16_t*>(c->bring),
                           c->view_stride, bvs, c->pitch, c->Rp, v0, nullptr);
        MR_HIP(hipGetLastError());
    }
    c->bring_n = std::max(c->bring_n, n_views);
    return MR_OK;
}

// The band copy of 16-bit storage fits kernel 7's 32-bit byte offsets and 24-bit pitch
// products (a record's reference band can lie up to a view beyond the last).
static bool band_fits(const Ctx* c) {
    return c->storage != MR_STORE_F32 &&
           band_view_stride(c) * static_cast<int64_t>(sizeof(uint16_t)) < (int64_t{1} << 30) &&
           (static_cast<int64_t>(c->pitch) << 3) < (int64_t{1} << 23);
}

int64_t pair_view_stride(const Ctx* c) { return 2 * static_cast<int64_t>(c->R + 2) * c->pitch; }

// The ring copy the v2 loop reads in this context: the band copy for 16-bit storage in a
// kernel-7 context, or a kernel-6 one of parallel beam (the 16-bit default there); the
// row-pair copy for f32 parallel beam (its points keep one detector row, so one 16-byte
// load brings both rows; measured on the parallel COR search: 38 against 73 ms; oblique
// cone slices lose, 7.4 against 5.8 ms on kernel 4); else the ring itself.
// A copy the last prepare_persistent_bp found no room for is not chosen.
int v2_ring_kind(const Ctx* c) {
    const bool par = c->kind == MR_KIND_PARALLEL;
    if ((c->bp_variant == 7 || (par && c->bp_variant == 6)) && band_fits(c))
        return c->band_no_room ? 0 : 2;
    if (c->storage == MR_STORE_F32 && par && (c->bp_variant == 4 || c->bp_variant >= 6))
        return c->pair_no_room ? 0 : 1;
    return 0;
}

int prepare_persistent_bp(Ctx* c, int n_views, hipStream_t s) {
    c->band_no_room = c->pair_no_room = false;       // memory may have been freed since
    const int rk = v2_ring_kind(c);
    const int rc = rk == 2 ? ensure_band_ring(c, n_views, s)
                           : rk == 1 ? ensure_pair_ring(c, n_views, s) : MR_OK;
    return rc == kNoRoom ? MR_OK : rc;               // v2_ring_kind now answers 0
}

// In a captured streaming graph: row-pair copy of the k views just stored, ending at the
// device view count.
int launch_pair_convert_tail(Ctx* c, const int* d_n, int k, hipStream_t s) {
    if (!c->pring || k <= 0 || c->storage != MR_STORE_F32) return fail(MR_ERUNTIME, "pair copy not allocated");
    const dim3 grid((c->pitch + 255) / 256, c->R + 2, k);
    hipLaunchKernelGGL(pair_convert_kernel<float>, grid, dim3(256), 0, s,
                       static_cast<const float*>(c->ring), static_cast<float*>(c->pring),
                       c->view_stride, pair_view_stride(c), c->pitch, 0, d_n);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

// In a captured streaming graph: band copy of the k views just stored, ending at the
// device view count.
int launch_band_convert_tail(Ctx* c, const int* d_n, int k, hipStream_t s) {
    const int64_t bvs = band_view_stride(c);
    if (!c->bring || k <= 0) return fail(MR_ERUNTIME, "band copy not allocated");
    const dim3 grid((c->pitch + 255) / 256, band_count(c->R), k);
    hipLaunchKernelGGL(band_convert_kernel, grid, dim3(256), 0, s,
                       static_cast<const uint16_t*>(c->ring), static_cast<uint16_t*>(c->bring),
                       c->view_stride, bvs, c->pitch, c->Rp, 0, d_n);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

// v4 (and v5 for cone beam): per chunk of tiles, the record kernel then the tile kernel.
int launch_bp_tiles_v4(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles, int n_views,
                       float scale, float* d_out, hipStream_t s) {
    if (ntiles <= 0 || n_views <= 0) {
        // no views: every point is zero; run with one view count of 0 through v2
        return launch_bp_tiles(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s);
    }
    const bool cone = c->kind == MR_KIND_CONE;
    const bool v5 = cone && c->bp_variant == 5;
    // the kernel the ring copies in memory allow: without room for the row-pair copy,
    // kernel 7 (16-bit, band copy) or 4 (the ring); without room for the band copy, 4
    auto run_as = [&](int variant) {
        const int sv = c->bp_variant;
        c->bp_variant = variant;
        const int rc = launch_bp_tiles_v4(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s);
        c->bp_variant = sv;
        return rc;
    };
    if (c->bp_variant == 6 || (c->bp_variant == 7 && c->storage == MR_STORE_F32)) {
        const int rc = ensure_pair_ring(c, n_views, s);
        if (rc == kNoRoom) return run_as(c->storage != MR_STORE_F32 && band_fits(c) ? 7 : 4);
        if (rc) return rc;
    } else if (c->bp_variant == 7) {
        if (!band_fits(c)) return run_as(6);
        const int rc = ensure_band_ring(c, n_views, s);
        if (rc == kNoRoom) return run_as(4);
        if (rc) return rc;
    }
    const bool band = c->bp_variant == 7 && c->storage != MR_STORE_F32;   // band-ring records
    c->kern_used = c->bp_variant;
    const int rec_floats = v5 ? V5_REC : BP_REC;
    const size_t per_tile = sizeof(float) * rec_floats * static_cast<size_t>(n_views);
    int chunk = static_cast<int>(std::max<size_t>(1, c->rec_budget / per_tile));
    chunk = std::min(chunk, ntiles);
    // v5: one "every box fits" flag per tile after the records
    const size_t flag_off = (per_tile * chunk + 255) / 256 * 256;
    int rc = ensure(reinterpret_cast<void**>(&c->d_recs), &c->recs_bytes,
                    flag_off + sizeof(int) * chunk);
    if (rc) return rc;
    int* d_fit = reinterpret_cast<int*>(reinterpret_cast<char*>(c->d_recs) + flag_off);
    const int le = c->storage == MR_STORE_F32 ? 2 : 1;
    const int subs = c->bp_ppt == 2 ? 2 : 1;
    const int kcap = (le == 2 ? V5Cfg<2>::K : V5Cfg<1>::K) * BP_THREADS;
    const int bufb = le == 2 ? V5Cfg<2>::BUF : V5Cfg<1>::BUF;
    for (int t0 = 0; t0 < ntiles; t0 += chunk) {
        const int nt = std::min(chunk, ntiles - t0);
        const int64_t nrec = static_cast<int64_t>(nt) * n_views;
        const dim3 rgrid(static_cast<unsigned>((nrec + 255) / 256));
        const bool last_chunk = t0 + chunk >= ntiles;
        if (v5) {
            hipLaunchKernelGGL(set_int_kernel, dim3((nt + 255) / 256), dim3(256), 0, s, d_fit, nt, 1);
            hipLaunchKernelGGL(records5_kernel, rgrid, dim3(256), 0, s, c->d_geom, c->d_vscale,
                               d_tasks, d_tiles + t0, nt, n_views, c->R, c->C, c->pitch, subs, le,
                               kcap, bufb, c->d_recs, d_fit);
            if (t0 == 0) MR_HIP(hipEventRecord(c->evk0, s));
            switch (c->storage) {
            case MR_STORE_F32:
                launch_v5_kernel<float>(c, c->d_recs, d_fit, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s);
                break;
            case MR_STORE_F16:
                launch_v5_kernel<__half>(c, c->d_recs, d_fit, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s);
                break;
            default:
                launch_v5_kernel<uint16_t>(c, c->d_recs, d_fit, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s);
                break;
            }
            MR_HIP(hipGetLastError());
            if (last_chunk) {
                MR_HIP(hipEventRecord(c->evk1, s));
                c->tile_timed = true;
            }
            continue;
        }
        hipLaunchKernelGGL(records_kernel, rgrid, dim3(256), 0, s, c->d_geom, c->d_vscale, d_tasks,
                           d_tiles + t0, nt, n_views, c->R, c->C, c->pitch, cone ? 1 : 0, c->d_recs,
                           band ? 1 : 0);
        if (t0 == 0) MR_HIP(hipEventRecord(c->evk0, s));
        switch (c->storage) {
        case MR_STORE_F32:
            cone ? launch_v4_kernel<true, float>(c, c->d_recs, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s)
                 : launch_v4_kernel<false, float>(c, c->d_recs, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s);
            break;
        case MR_STORE_F16:
            cone ? launch_v4_kernel<true, __half>(c, c->d_recs, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s)
                 : launch_v4_kernel<false, __half>(c, c->d_recs, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s);
            break;
        default:
            cone ? launch_v4_kernel<true, uint16_t>(c, c->d_recs, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s)
                 : launch_v4_kernel<false, uint16_t>(c, c->d_recs, d_tasks, d_tiles + t0, nt, n_views, scale, d_out, s);
            break;
        }
        MR_HIP(hipGetLastError());
        if (last_chunk) {
            MR_HIP(hipEventRecord(c->evk1, s));
            c->tile_timed = true;
        }
    }
    return MR_OK;
}

// The form of the z-column kernel a volume request of n_views runs: ZCOL_STAGED (boxes of
// ZQL-plane chunks through LDS) for f32 rings (staged from the ring) and for 16-bit rings
// with room for the row-pair copy (staged from it); else ZCOL_RING (ZQ-plane chunks
// gathered from the ring).  Builds or extends the copy.
int zcol_form(Ctx* c, int n_views, hipStream_t s, int* form) {
    *form = ZCOL_STAGED;
    if (c->storage == MR_STORE_F32) return MR_OK;      // staged from the ring itself
    *form = ZCOL_RING;
    const int rc = ensure_pair_ring(c, n_views, s);
    if (rc == kNoRoom) return MR_OK;
    if (rc) return rc;
    *form = ZCOL_STAGED;
    return MR_OK;
}

// The z-column kernel over the tiles of build_zcol_tiles (chunks of ZQL planes for
// ZCOL_STAGED, ZQ for ZCOL_RING, x ZT x ZT columns).  n_views > 0.
int launch_bp_zcol(Ctx* c, int form, const mr_task* d_tasks, int ntasks, const Tile* d_tiles,
                   int ntiles, int n_views, float scale, float* d_out, hipStream_t s) {
    if (ntiles <= 0 || n_views <= 0) return MR_OK;
    const int64_t pvs = pair_view_stride(c);
    const int xcd = c->xcd_map != 0 ? 1 : 0;
    // per (task, view) coefficients in the records buffer
    const int64_t ncoef = static_cast<int64_t>(ntasks) * n_views;
    int rc = ensure(reinterpret_cast<void**>(&c->d_recs), &c->recs_bytes,
                    sizeof(double) * ZC * static_cast<size_t>(ncoef));
    if (rc) return rc;
    const double* d_coef = reinterpret_cast<const double*>(c->d_recs);
    hipLaunchKernelGGL(zcol_coef_kernel, dim3(static_cast<unsigned>((ncoef + 255) / 256)), dim3(256),
                       0, s, c->d_geom, d_tasks, ntasks, n_views, reinterpret_cast<double*>(c->d_recs));
    MR_HIP(hipEventRecord(c->evk0, s));
#define MR_ZC(ST, RK, STG, NQ)                                                                 \
    hipLaunchKernelGGL((bp_zcol_kernel<ST, RK, STG, NQ>), dim3(ntiles), dim3(BP_THREADS), 0, s,  \
                       static_cast<const ST*>(RK ? c->pring : c->ring), RK ? pvs : c->view_stride, \
                       c->pitch, c->R, c->C, d_coef, c->d_vscale, d_tasks, d_tiles, n_views,      \
                       scale, d_out, xcd)
    const bool staged = form == ZCOL_STAGED;
    switch (c->storage) {
    case MR_STORE_F32:
        if (staged) MR_ZC(float, 0, true, ZQL);
        else MR_ZC(float, 0, false, ZQ);
        break;
    case MR_STORE_F16:
        if (staged) MR_ZC(__half, 1, true, ZQL);
        else MR_ZC(__half, 0, false, ZQ);
        break;
    default:
        if (staged) MR_ZC(uint16_t, 1, true, ZQL);
        else MR_ZC(uint16_t, 0, false, ZQ);
        break;
    }
#undef MR_ZC
    MR_HIP(hipGetLastError());
    MR_HIP(hipEventRecord(c->evk1, s));
    c->tile_timed = true;
    c->tile_launches = 1;
    c->kern_used = staged ? 9 : 8;
    return MR_OK;
}

int launch_bp_adjoint(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles,
                      int n_views, float scale, const float* d_g, float* d_gq, hipStream_t s) {
    if (ntiles <= 0 || n_views <= 0) return MR_OK;
    const bool cone = c->kind == MR_KIND_CONE;
    const size_t rec_bytes = sizeof(float) * BP_REC * static_cast<size_t>(ntiles) * n_views;
    if (ntiles <= kAdjGatherMaxTiles && rec_bytes <= c->rec_budget) {
        // the gather: records of every (tile, view), the masked weights, the detector box
        // of every (view, tile), then one workgroup per detector tile and view
        int rc = ensure(reinterpret_cast<void**>(&c->d_recs), &c->recs_bytes, rec_bytes);
        if (rc) return rc;
        auto up = [](size_t x) { return (x + 255) / 256 * 256; };
        const size_t gm_bytes = up(sizeof(float) * BP_TP * BP_TS * static_cast<size_t>(ntiles));
        const size_t gmax_bytes = up(sizeof(float) * static_cast<size_t>(ntiles));
        const size_t box_bytes = up(sizeof(int4) * static_cast<size_t>(ntiles) * n_views);
        const size_t bnd_bytes = up(sizeof(float) * static_cast<size_t>(ntiles) * n_views);
        rc = ensure(&c->d_adj, &c->adj_bytes, gm_bytes + gmax_bytes + box_bytes + bnd_bytes);
        if (rc) return rc;
        char* base = static_cast<char*>(c->d_adj);
        float* gm = reinterpret_cast<float*>(base);
        float* gmax = reinterpret_cast<float*>(base + gm_bytes);
        int4* boxes = reinterpret_cast<int4*>(base + gm_bytes + gmax_bytes);
        float* bound = reinterpret_cast<float*>(base + gm_bytes + gmax_bytes + box_bytes);
        const int64_t nrec = static_cast<int64_t>(ntiles) * n_views;
        const dim3 rgrid(static_cast<unsigned>((nrec + 255) / 256));
        hipLaunchKernelGGL(records_kernel, rgrid, dim3(256), 0, s, c->d_geom, nullptr, d_tasks,
                           d_tiles, ntiles, n_views, c->R, c->C, c->pitch, cone ? 1 : 0, c->d_recs, 0);
        hipLaunchKernelGGL(adj_weights_kernel, dim3(ntiles), dim3(BP_THREADS), 0, s, d_tasks, d_tiles,
                           scale, d_g, gm, gmax);
        const dim3 ggrid(n_views, (c->R + ADJ_TH - 1) / ADJ_TH, (c->C + ADJ_TW - 1) / ADJ_TW);
        if (cone) {
            hipLaunchKernelGGL(adj_boxes_kernel<true>, rgrid, dim3(256), 0, s, c->d_recs, gmax,
                               ntiles, n_views, c->R, c->C, boxes, bound);
            hipLaunchKernelGGL(adj_gather_kernel<true>, ggrid, dim3(BP_THREADS), 0, s, c->d_recs,
                               boxes, bound, gm, ntiles, n_views, c->R, c->C, d_gq);
        } else {
            hipLaunchKernelGGL(adj_boxes_kernel<false>, rgrid, dim3(256), 0, s, c->d_recs, gmax,
                               ntiles, n_views, c->R, c->C, boxes, bound);
            hipLaunchKernelGGL(adj_gather_kernel<false>, ggrid, dim3(BP_THREADS), 0, s, c->d_recs,
                               boxes, bound, gm, ntiles, n_views, c->R, c->C, d_gq);
        }
        MR_HIP(hipGetLastError());
        return MR_OK;
    }
    // many tiles (every detector workgroup would test every tile's box) or records beyond
    // the buffer: the scatter, with the views split over workgroups too (about 16
    // workgroups per CU): the scatters of different views go to different projections,
    // so their atomics rarely meet
    const int chunks = std::max(1, std::min(n_views, (16 * c->n_cus + ntiles - 1) / ntiles));
    const int vch = (n_views + chunks - 1) / chunks;
    dim3 grid(ntiles, (n_views + vch - 1) / vch), block(BP_THREADS);
    if (c->kind == MR_KIND_CONE)
        hipLaunchKernelGGL(bp_adjoint_kernel<true>, grid, block, 0, s, c->pitch, c->R, c->C,
                           c->d_geom, d_tasks, d_tiles, n_views, vch, scale, d_g, d_gq);
    else
        hipLaunchKernelGGL(bp_adjoint_kernel<false>, grid, block, 0, s, c->pitch, c->R, c->C,
                           c->d_geom, d_tasks, d_tiles, n_views, vch, scale, d_g, d_gq);
    MR_HIP(hipGetLastError());
    return MR_OK;
}

template <bool CONE>
static int dshift_dispatch(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles,
                           int n_views, float scale, float* d_out, hipStream_t s) {
    // a request of fewer tiles than the device holds workgroups splits its views into parts
    // of at least 16 views, up to about one workgroup per resident slot
    int parts = 1;
    const int slots = 8 * c->n_cus;
    if (ntiles < slots && n_views >= 32) {
        parts = std::max(1, std::min((slots + ntiles - 1) / ntiles, n_views / 16));
        const size_t need = sizeof(float) * BP_TP * BP_TS * static_cast<size_t>(ntiles) * parts;
        if (need > kPartBudget) parts = 1;
        if (parts > 1) {
            const int rc = ensure(&c->d_part, &c->part_bytes, need);
            if (rc) return rc;
        }
    }
    int vper = n_views;
    if (parts > 1) {
        vper = (n_views + parts - 1) / parts;
        parts = (n_views + vper - 1) / vper;             // no empty part
    }
    float* part = parts > 1 ? static_cast<float*>(c->d_part) : nullptr;
    dim3 grid(ntiles, parts), block(BP_THREADS);
    switch (c->storage) {
    case MR_STORE_F32:
        hipLaunchKernelGGL((bp_dshift_kernel<CONE, float>), grid, block, 0, s,
                           static_cast<const float*>(c->ring), c->view_stride, c->pitch, c->R, c->C,
                           c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out, vper, part);
        break;
    case MR_STORE_F16:
        hipLaunchKernelGGL((bp_dshift_kernel<CONE, __half>), grid, block, 0, s,
                           static_cast<const __half*>(c->ring), c->view_stride, c->pitch, c->R,
                           c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out,
                           vper, part);
        break;
    default:
        hipLaunchKernelGGL((bp_dshift_kernel<CONE, uint16_t>), grid, block, 0, s,
                           static_cast<const uint16_t*>(c->ring), c->view_stride, c->pitch, c->R,
                           c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out,
                           vper, part);
        break;
    }
    MR_HIP(hipGetLastError());
    if (part) {
        hipLaunchKernelGGL(part_sum_kernel, dim3(ntiles), dim3(256), 0, s, d_tasks, d_tiles, part,
                           parts, scale, d_out);
        MR_HIP(hipGetLastError());
    }
    return MR_OK;
}

int launch_bp_dshift(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles, int n_views,
                     float scale, float* d_out, hipStream_t s) {
    if (ntiles <= 0) return MR_OK;
    return c->kind == MR_KIND_CONE
               ? dshift_dispatch<true>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s)
               : dshift_dispatch<false>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s);
}

int launch_read_views(Ctx* c, int first, int count, float* d_out, hipStream_t s) {
    if (count <= 0) return MR_OK;
    dim3 grid(c->R, count), block(256);
    switch (c->storage) {
    case MR_STORE_F32:
        hipLaunchKernelGGL(read_views_kernel<float>, grid, block, 0, s,
                           static_cast<const float*>(c->ring), c->view_stride, c->pitch, c->R, c->C,
                           c->d_vscale, first, d_out);
        break;
    case MR_STORE_F16:
        hipLaunchKernelGGL(read_views_kernel<__half>, grid, block, 0, s,
                           static_cast<const __half*>(c->ring), c->view_stride, c->pitch, c->R,
                           c->C, c->d_vscale, first, d_out);
        break;
    default:
        hipLaunchKernelGGL(read_views_kernel<uint16_t>, grid, block, 0, s,
                           static_cast<const uint16_t*>(c->ring), c->view_stride, c->pitch, c->R,
                           c->C, c->d_vscale, first, d_out);
        break;
    }
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr
