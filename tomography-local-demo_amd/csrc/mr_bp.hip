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

// FP64 record of one (tile, view): coefficients of uh, vh, wh as affine functions of
// the in-tile offsets (ds along the fast axis b, dp along the slow axis a), shifted
// so that uh/wh, vh/wh are columns/rows relative to the reference (cref, rref);
// clamp bounds of the relative coordinates; per-view storage scale; base index.
__device__ void make_record(const ViewGeom* __restrict__ gp, float vs, const mr_task& T,
                            const double* xt, int R, int C, int pitch, bool cone,
                            float* __restrict__ rec) {
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
    const double cref = floor(colc), rref = floor(rowc);
    const double dc = cc - cref, dr = cr - rref;
    U0 += dc * W0; Ua += dc * Wa; Ub += dc * Wb;
    V0 += dr * W0; Va += dr * Wa; Vb += dr * Wb;
    // cone: the weight 1/wh^2 absorbs the view's storage scale
    const double f = cone ? 1.0 / sqrt(static_cast<double>(vs)) : 1.0;
    float4* r4 = reinterpret_cast<float4*>(rec);
    r4[0] = make_float4(static_cast<float>(U0 * f), static_cast<float>(Ua * f),
                        static_cast<float>(Ub * f), static_cast<float>(V0 * f));
    r4[1] = make_float4(static_cast<float>(Va * f), static_cast<float>(Vb * f),
                        static_cast<float>(W0 * f), static_cast<float>(Wa * f));
    const int base = (static_cast<int>(rref) + 1) * pitch + (static_cast<int>(cref) + 1);
    r4[2] = make_float4(static_cast<float>(Wb * f), static_cast<float>(-1.0 - cref),
                        static_cast<float>(C - cref), static_cast<float>(-1.0 - rref));
    r4[3] = make_float4(static_cast<float>(R - rref), vs, __int_as_float(base), 0.0f);
}

template <bool CONE, typename ST>
__global__ __launch_bounds__(BP_THREADS) void bp_tiles_kernel(
    const ST* __restrict__ ring, int64_t view_stride, int pitch, int R, int C,
    const ViewGeom* __restrict__ geom, const float* __restrict__ vscale,
    const mr_task* __restrict__ tasks, const Tile* __restrict__ tiles, int n_views,
    float scale, float* __restrict__ out) {
    __shared__ float4 rec4[BP_VPASS * BP_REC / 4];
    float* rec = reinterpret_cast<float*>(rec4);

    const Tile tl = tiles[blockIdx.x];
    const mr_task& T = tasks[tl.task];
    const int t = threadIdx.x;
    const int tx = t & (BP_TS - 1);
    const int ty = t / BP_TS;                  // 0..7
    constexpr int PSTEP = BP_THREADS / BP_TS;  // 8
    constexpr int PPT = BP_TP / PSTEP;         // 4 points per thread

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
    float acc[PPT], accb[PPT];
    float fr_par[PPT];
    int roff_par[PPT];
#pragma unroll
    for (int m = 0; m < PPT; ++m) { acc[m] = 0.0f; accb[m] = 0.0f; fr_par[m] = 0.0f; roff_par[m] = 0; }

    for (int vb = 0; vb < n_views; vb += BP_VPASS) {
        const int nv = min(BP_VPASS, n_views - vb);
        __syncthreads();
        if (t < nv)
            make_record(geom + vb + t, vscale[vb + t], T, xt, R, C, pitch, CONE, rec + t * BP_REC);
        __syncthreads();
        if (!CONE && vb == 0) {
            // parallel beam: the detector row of a point is the same for every view
            const float4 r1 = rec4[0 * 4 + 1];
            const float4 r0 = rec4[0 * 4 + 0];
            const float4 r2 = rec4[0 * 4 + 2];
            const float4 r3 = rec4[0 * 4 + 3];
            const float V0 = r0.w, Va = r1.x, Vb = r1.y;
            const float rlo = r2.w, rhi = r3.x;
#pragma unroll
            for (int m = 0; m < PPT; ++m) {
                const float dp = dp0 + static_cast<float>(PSTEP * m);
                float row = fmaf(dp, Vb, fmaf(ds, Va, V0));
                row = __builtin_amdgcn_fmed3f(row, rlo, rhi);
                const float rf = floorf(row);
                fr_par[m] = row - rf;
                roff_par[m] = static_cast<int>(rf) * pitch;
            }
        }
        for (int v = 0; v < nv; ++v) {
            const float4 r0 = rec4[v * 4 + 0];
            const float4 r1 = rec4[v * 4 + 1];
            const float4 r2 = rec4[v * 4 + 2];
            const float4 r3 = rec4[v * 4 + 3];
            const float U0 = r0.x, Ua = r0.y, Ub = r0.z;
            const float clo = r2.y, chi = r2.z;
            const int base = __float_as_int(r3.z);
            const ST* __restrict__ q0 = ring + static_cast<int64_t>(vb + v) * view_stride;
            const ST* __restrict__ q1 = q0 + pitch;
            if (CONE) {
                const float V0 = r0.w, Va = r1.x, Vb = r1.y;
                const float W0 = r1.z, Wa = r1.w, Wb = r2.x;
                const float rlo = r2.w, rhi = r3.x;
                const float uh0 = fmaf(dp0, Ub, fmaf(ds, Ua, U0));
                const float vh0 = fmaf(dp0, Vb, fmaf(ds, Va, V0));
                const float wh0 = fmaf(dp0, Wb, fmaf(ds, Wa, W0));
#pragma unroll
                for (int m = 0; m < PPT; ++m) {
                    const float dm = static_cast<float>(PSTEP * m);
                    const float uh = fmaf(dm, Ub, uh0);
                    const float vh = fmaf(dm, Vb, vh0);
                    const float wh = fmaf(dm, Wb, wh0);
                    const float rw = __builtin_amdgcn_rcpf(wh);
                    float col = uh * rw, row = vh * rw;
                    const float wt = rw * rw;
                    col = __builtin_amdgcn_fmed3f(col, clo, chi);
                    row = __builtin_amdgcn_fmed3f(row, rlo, rhi);
                    const float cf = floorf(col), rf = floorf(row);
                    const float fc = col - cf, fr = row - rf;
                    const uint32_t idx = static_cast<uint32_t>(
                        base + static_cast<int>(rf) * pitch + static_cast<int>(cf));
                    const float a = Store<ST>::ld(q0, idx), b = Store<ST>::ld(q0, idx + 1);
                    const float c = Store<ST>::ld(q1, idx), d = Store<ST>::ld(q1, idx + 1);
                    const float top = fmaf(fc, b - a, a);
                    const float bot = fmaf(fc, d - c, c);
                    acc[m] = fmaf(wt, fmaf(fr, bot - top, top), acc[m]);
                }
            } else {
                const float vs = r3.y;
                const float uh0 = fmaf(dp0, Ub, fmaf(ds, Ua, U0));
#pragma unroll
                for (int m = 0; m < PPT; ++m) {
                    const float dm = static_cast<float>(PSTEP * m);
                    float col = fmaf(dm, Ub, uh0);
                    col = __builtin_amdgcn_fmed3f(col, clo, chi);
                    const float cf = floorf(col);
                    const float fc = col - cf;
                    const uint32_t idx =
                        static_cast<uint32_t>(base + roff_par[m] + static_cast<int>(cf));
                    const float a = Store<ST>::ld(q0, idx), b = Store<ST>::ld(q0, idx + 1);
                    const float c = Store<ST>::ld(q1, idx), d = Store<ST>::ld(q1, idx + 1);
                    acc[m] = fmaf(vs, fmaf(fc, b - a, a), acc[m]);
                    accb[m] = fmaf(vs, fmaf(fc, d - c, c), accb[m]);
                }
            }
        }
    }

    const int64_t obase = T.out_offset + static_cast<int64_t>(tl.q) * T.stride_q;
#pragma unroll
    for (int m = 0; m < PPT; ++m) {
        if (!valid[m]) continue;
        const int p = p0 + ty + PSTEP * m;
        float val = acc[m];
        if (!CONE) val = fmaf(fr_par[m], accb[m] - acc[m], acc[m]);
        out[obase + static_cast<int64_t>(p) * T.stride_p + static_cast<int64_t>(s) * T.stride_s] =
            inside[m] ? val * scale : 0.0f;
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
                   float scale, float* d_out, hipStream_t s) {
    dim3 grid(ntiles), block(BP_THREADS);
    switch (c->storage) {
    case MR_STORE_F32:
        hipLaunchKernelGGL((bp_tiles_kernel<CONE, float>), grid, block, 0, s,
                           static_cast<const float*>(c->ring), c->view_stride, c->pitch, c->R, c->C,
                           c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out);
        break;
    case MR_STORE_F16:
        hipLaunchKernelGGL((bp_tiles_kernel<CONE, __half>), grid, block, 0, s,
                           static_cast<const __half*>(c->ring), c->view_stride, c->pitch, c->R,
                           c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out);
        break;
    default:
        hipLaunchKernelGGL((bp_tiles_kernel<CONE, uint16_t>), grid, block, 0, s,
                           static_cast<const uint16_t*>(c->ring), c->view_stride, c->pitch, c->R,
                           c->C, c->d_geom, c->d_vscale, d_tasks, d_tiles, n_views, scale, d_out);
        break;
    }
    MR_HIP(hipGetLastError());
    return MR_OK;
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

int launch_bp_tiles(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles, int n_views,
                    float scale, float* d_out, hipStream_t s) {
    if (ntiles <= 0) return MR_OK;
    return c->kind == MR_KIND_CONE
               ? dispatch_tiles<true>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s)
               : dispatch_tiles<false>(c, d_tasks, d_tiles, ntiles, n_views, scale, d_out, s);
}

int launch_bp_points(Ctx* c, const double* d_pts, int64_t M, double cor_shift, int n_views,
                     float scale, float* d_out, hipStream_t s) {
    if (M <= 0) return MR_OK;
    return c->kind == MR_KIND_CONE
               ? dispatch_points<true>(c, d_pts, M, cor_shift, n_views, scale, d_out, s)
               : dispatch_points<false>(c, d_pts, M, cor_shift, n_views, scale, d_out, s);
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
