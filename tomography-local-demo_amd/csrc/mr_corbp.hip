// Shift-group back-projection of rotation-centre searches (parallel beam).
//
// estimate_cor (recon.py:498-512) reconstructs one axial slice for every shift of a grid
// (by default -3..3 px in 0.25 px steps).  bp_parallel_numba (recon.py:175-191) samples
// view t of a point at row z / pv + (R - 1) / 2 and column (-x sin t + y cos t) / pu +
// (C - 1) / 2 + shift with the per-tap zero-outside bilinear of _bil_nb
// (recon.py:156-173).  On an axial slice the row is the same for every point and view,
// and the trial images differ only by the shift added to the column.
//
// cor_rows_kernel: for every set of tasks equal but for their shifts (CorRowSet) and every
// view, the detector row of the set interpolated between its two ring rows, times the
// view's storage scale, as floats with kCorPad zero columns on each side (every tap
// outside the detector reads 0, as the clamped ring lookups do), and the view's column
// coefficients in float64.
//
// cor_bp_kernel<Q>: a group (CorGroup) holds the tasks of a set whose shifts are s0 + k / Q,
// k < 6 Q + 1.  Per view a thread forms its point's column c = col + s0 once, loads the
// kCorWin = 8 row values from floor(c) on, and adds each shift's lerp: shift k = Q t + u
// lies at c + t + u / Q: with f = fract(c) and x_u = f + u / Q in [0, 2), its value is
// w[t] + min(x_u, 1) d[t] + max(x_u - 1, 0) d[t + 1] (d[j] = w[j + 1] - w[j]), three VALU
// per shift.  No gather per shift: one 8-value load per point and view serves up to 25
// shifts.
#include <type_traits>

#include "mr_device.h"
#include "mr_internal.h"

namespace mr {

namespace {

constexpr int kCorThreads = 256;            // a 32 x 8 quarter of a 32 x 32 tile
constexpr int kCorSubs = BP_TP / (kCorThreads / BP_TS);   // workgroups per tile

__device__ inline double dot3d(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Grid (ceil(lw / 256), views, sets).  rows[set][v][lw]: row values at detector column
// x - kCorPad; coef[set][v] = (column of the set's task origin + kCorPad, column step
// along p, along s), float64.  Views v >= the device view count are skipped.
template <typename ST>
__global__ __launch_bounds__(256) void cor_rows_kernel(
    const ST* __restrict__ ring, int64_t view_stride, int pitch, int R, int C,
    const float* __restrict__ vscale, const ViewGeom* __restrict__ geom,
    const mr_task* __restrict__ tasks, const CorRowSet* __restrict__ sets,
    const int* __restrict__ n_views_dev, int n_views_arg, int lw, int nvt,
    float* __restrict__ rows, double* __restrict__ coef) {
    const int v = blockIdx.y, set = blockIdx.z;
    const int n_views = n_views_dev ? *n_views_dev : n_views_arg;
    if (v >= n_views) return;
    const mr_task& T = tasks[sets[set].task];
    const ViewGeom& g = geom[v];
    // the set's detector row (the same for every point: a[2] = b[2] = 0), clamped to
    // [-1, R] as the ring lookups clamp it
    double row = dot3d(g.P[1], T.o) + g.P[1][3] + 0.5 * (R - 1);
    row = fmin(fmax(row, -1.0), static_cast<double>(R));          // NaN -> -1
    const double rf = floor(row);
    const int r = static_cast<int>(rf);
    const float fr = static_cast<float>(row - rf);
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x < lw) {
        const int c = x - kCorPad;
        float val = 0.0f;
        if (c >= 0 && c < C) {
            // padded rows r + 1 and r + 2 (rows -1 .. R + 1 are stored, the border zero)
            const ST* p = ring + static_cast<int64_t>(v) * view_stride +
                          static_cast<int64_t>(r + 1) * pitch + (c + 1);
            float a, b;
            if constexpr (sizeof(ST) == 4) {
                a = p[0];
                b = p[pitch];
            } else if constexpr (std::is_same<ST, __half>::value) {
                a = __half2float(p[0]);
                b = __half2float(p[pitch]);
            } else {
                a = __uint_as_float(static_cast<uint32_t>(p[0]) << 16);
                b = __uint_as_float(static_cast<uint32_t>(p[pitch]) << 16);
            }
            val = vscale[v] * fmaf(fr, b - a, a);
        }
        rows[(static_cast<int64_t>(set) * nvt + v) * lw + x] = val;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double* k = coef + (static_cast<int64_t>(set) * nvt + v) * 4;
        k[0] = dot3d(g.P[0], T.o) + g.P[0][3] + 0.5 * (C - 1) + kCorPad;
        k[1] = dot3d(g.P[0], T.a);
        k[2] = dot3d(g.P[0], T.b);
        k[3] = 0.0;
    }
}

// Grid: kCorSubs workgroups per tile (tile = block / kCorSubs after the XCD map).  Thread
// t handles point (p, s) = (32 tp + 8 sub + t / 32, 32 ts + t mod 32) of the group's task
// and its ns shifts.  rows/coef as written by cor_rows_kernel.
template <int Q>
__global__ __launch_bounds__(kCorThreads) void cor_bp_kernel(
    const mr_task* __restrict__ tasks, const CorRowSet* __restrict__ sets,
    const CorGroup* __restrict__ groups, const CorTile* __restrict__ tiles,
    const float* __restrict__ rows, const double* __restrict__ coef, int lw, int nvt, int C,
    const int* __restrict__ n_views_dev, int n_views_arg, float* __restrict__ out, int xcd) {
    constexpr int NK = 6 * Q + 1;                  // shifts of a full group
    static_assert(NK <= kCorMaxShifts, "group size");
    const int b = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const CorTile tl = tiles[b / kCorSubs];
    const int sub = b - (b / kCorSubs) * kCorSubs;
    const CorGroup& G = groups[tl.group];
    const int set = G.set;
    const mr_task& T = tasks[sets[set].task];
    const int t = threadIdx.x;
    const int ds = t & (BP_TS - 1), dp = t / BP_TS;
    const int pb = tl.tp * BP_TP + sub * (kCorThreads / BP_TS), sb = tl.ts * BP_TS;
    const int p = pb + dp, s = sb + ds;
    const bool valid = p < T.np && s < T.ns;
    const bool inside = valid && (!T.mask_box || in_box(T, p, s));
    const int n_views = n_views_dev ? *n_views_dev : n_views_arg;
    const int ns = G.ns;
    float acc[NK];
#pragma unroll
    for (int k = 0; k < NK; ++k) acc[k] = 0.0f;

    if (!tl.zero && n_views > 0) {
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(rows + static_cast<int64_t>(set) * nvt * lw), 0,
            static_cast<int>(static_cast<int64_t>(nvt) * lw * 4), 0x00020000);
        const double* kc = coef + static_cast<int64_t>(set) * nvt * 4;
        const float fdp = static_cast<float>(dp), fds = static_cast<float>(ds);
        const double s0 = G.s0, dpb = static_cast<double>(pb), dsb = static_cast<double>(sb);
        // byte offset of the first of the 8 row values of view v, and the fraction there
        auto locate = [&](int v, float& f) {
            const double c0 = kc[4 * v] + s0 + dpb * kc[4 * v + 1] + dsb * kc[4 * v + 2];
            const double cf = floor(c0);
            const float rel = static_cast<float>(c0 - cf);
            const float c = fmaf(fdp, static_cast<float>(kc[4 * v + 1]),
                                 fmaf(fds, static_cast<float>(kc[4 * v + 2]), rel));
            f = __builtin_amdgcn_fractf(c);
            // columns past either end read the zero padding: [pad - 9, C + pad] keeps all
            // 8 values inside the row, and a clamped start only happens when every tap of
            // every shift lies off the detector
            int i = static_cast<int>(fmin(fmax(cf, -1e6), 1e6)) + cvt_flr(c);
            i = min(max(i, kCorPad - kCorWin - 1), C + kCorPad);
            return (v * lw + i) * 4;
        };
        auto load8 = [&](int voff, float (&w)[kCorWin]) {
#pragma unroll
            for (int j = 0; j < kCorWin; ++j)
                w[j] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff + 4 * j, 0, 0));
        };
        // shift k = Q t + u lies at t + x_u after the window start, x_u = f + u / Q in
        // [0, 2): the linear interpolant there is w[t] + min(x_u, 1) d[t] + max(x_u - 1, 0)
        // d[t + 1] (d[j] = w[j + 1] - w[j]), three VALU per shift and two for u = 0
        auto sum_view = [&](const float (&w)[kCorWin], float f) {
            float d[kCorWin - 1];
#pragma unroll
            for (int j = 0; j + 1 < kCorWin; ++j) d[j] = w[j + 1] - w[j];
#pragma unroll
            for (int tt = 0; Q * tt < NK; ++tt) acc[Q * tt] = fmaf(f, d[tt], acc[Q * tt] + w[tt]);
#pragma unroll
            for (int u = 1; u < Q; ++u) {
                const float x = f + static_cast<float>(u) / Q;
                const float al = fminf(x, 1.0f), be = fmaxf(x - 1.0f, 0.0f);
#pragma unroll
                for (int tt = 0; Q * tt + u < NK; ++tt) {
                    const int k = Q * tt + u;
                    acc[k] = fmaf(be, d[tt + 1], fmaf(al, d[tt], acc[k] + w[tt]));
                }
            }
        };
        // two windows in turn: the next view's values are in flight while one is summed
        // (a request past the last view re-reads the last one and is never summed)
        float wa[kCorWin], wb[kCorWin], fa, fb;
        load8(locate(0, fa), wa);
        int v = 0;
        for (; v + 1 < n_views; v += 2) {
            load8(locate(v + 1, fb), wb);
            __builtin_amdgcn_sched_barrier(0);
            sum_view(wa, fa);
            load8(locate(v + 2 < n_views ? v + 2 : v + 1, fa), wa);
            __builtin_amdgcn_sched_barrier(0);
            sum_view(wb, fb);
        }
        if (v < n_views) sum_view(wa, fa);
    }
    if (!valid) return;
    const float scale = n_views > 0 ? static_cast<float>(M_PI / n_views) : 0.0f;
    const int64_t po = static_cast<int64_t>(p) * T.stride_p + static_cast<int64_t>(s) * T.stride_s;
#pragma unroll
    for (int k = 0; k < NK; ++k) {
        if (k >= ns || G.out[k] < 0) continue;
        out[G.out[k] + po] = inside ? acc[k] * scale : 0.0f;
    }
}

}  // namespace

int launch_cor_groups(Ctx* c, const CorLaunch& L, hipStream_t s) {
    MR_CHECK(L.q == 1 || L.q == 2 || L.q == 4, MR_EVALUE, "shift grid 1/q with q = 1, 2 or 4");
    MR_CHECK(static_cast<int64_t>(L.nvt) * L.lw * 4 < (int64_t{1} << 31), MR_EVALUE,
             "rows of a COR shift set exceed 2 GiB");
    const int nv_grid = L.d_nviews ? L.nvt : L.n_views;
    if (nv_grid > 0 && L.nsets > 0) {
        const dim3 grid((L.lw + 255) / 256, nv_grid, L.nsets);
#define MR_COR_ROWS(ST)                                                                        \
    hipLaunchKernelGGL(cor_rows_kernel<ST>, grid, dim3(256), 0, s, static_cast<const ST*>(c->ring), \
                       c->view_stride, c->pitch, c->R, c->C, c->d_vscale, c->d_geom, L.d_tasks,     \
                       L.d_sets, L.d_nviews, L.n_views, L.lw, L.nvt, L.d_rows, L.d_coef)
        if (c->storage == MR_STORE_F32) MR_COR_ROWS(float);
        else if (c->storage == MR_STORE_F16) MR_COR_ROWS(__half);
        else MR_COR_ROWS(uint16_t);
#undef MR_COR_ROWS
    }
    if (L.ntiles > 0) {
        const dim3 grid(L.ntiles * kCorSubs);
#define MR_COR_BP(QQ)                                                                            \
    hipLaunchKernelGGL(cor_bp_kernel<QQ>, grid, dim3(kCorThreads), 0, s, L.d_tasks, L.d_sets,     \
                       L.d_groups, L.d_tiles, L.d_rows, L.d_coef, L.lw, L.nvt, c->C, L.d_nviews, \
                       L.n_views, L.d_out, c->xcd_map ? 1 : 0)
        if (L.q == 1) MR_COR_BP(1);
        else if (L.q == 2) MR_COR_BP(2);
        else MR_COR_BP(4);
#undef MR_COR_BP
    }
    MR_HIP(hipGetLastError());
    return MR_OK;
}

}  // namespace mr
