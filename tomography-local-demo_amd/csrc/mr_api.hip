// C ABI of libminirecast_hip.so: contexts, float64 host-side geometry preparation,
// ingest into the filtered-projection ring, and back-projection requests.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>

#include "mr_internal.h"

namespace mr {

namespace {
thread_local std::string g_err;
constexpr int kAbiVersion = 2;

int fft_length(int n) {
    // 2 ** ceil(log2(2 n)) (recon.py:37): the smallest power of two >= 2n
    int N = 1;
    while (N < 2 * n) N <<= 1;
    return N;
}

// Widest detector row the filter transforms (FFT length 2^kFilterMaxLogN).
constexpr int kMaxFilterCols = (1 << kFilterMaxLogN) / 2;

int check_filter_cols(int n) {
    MR_CHECK(n <= kMaxFilterCols, MR_EVALUE,
             "detector rows of " + std::to_string(n) + " columns: at most " +
                 std::to_string(kMaxFilterCols) + " are supported");
    return MR_OK;
}

// In-place iterative radix-2 complex FFT in float64 (forward, e^{-2 pi i k n / N}).
void fft64(std::vector<std::complex<double>>& a) {
    const size_t N = a.size();
    for (size_t i = 1, j = 0; i < N; ++i) {
        size_t bit = N >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= N; len <<= 1) {
        const size_t half = len / 2;
        for (size_t k = 0; k < half; ++k) {
            const double ang = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(len);
            const std::complex<double> w(std::cos(ang), std::sin(ang));
            for (size_t i = 0; i < N; i += len) {
                const std::complex<double> u = a[i + k], v = a[i + k + half] * w;
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
        }
    }
}

double window_at(int window, double f) {
    switch (window) {
    case MR_WINDOW_RAMLAK: return 1.0;
    case MR_WINDOW_SHEPP_LOGAN: {
        const double y = M_PI * (f == 0.0 ? 1.0e-20 : f);
        return std::sin(y) / y;
    }
    case MR_WINDOW_HANN: return 0.5 * (1.0 + std::cos(2.0 * M_PI * f));
    case MR_WINDOW_COSINE: return std::cos(M_PI * f);
    default: return 0.0;
    }
}

bool valid_window(int w) { return w >= MR_WINDOW_RAMLAK && w <= MR_WINDOW_COSINE; }
bool valid_storage(int s) { return s >= MR_STORE_F32 && s <= MR_STORE_BF16; }

struct V3 {
    double x, y, z;
};
V3 v3(const double* p) { return {p[0], p[1], p[2]}; }
V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double norm(V3 a) { return std::sqrt(dot(a, a)); }

void set_row(double* P, V3 a, double a3) {
    P[0] = a.x; P[1] = a.y; P[2] = a.z; P[3] = a3;
}

size_t elem_size(int storage) { return storage == MR_STORE_F32 ? 4 : 2; }

void free_geometry(Ctx* c) {
    if (c->ring) hipFree(c->ring);
    if (c->d_geom) hipFree(c->d_geom);
    if (c->d_vscale) hipFree(c->d_vscale);
    if (c->d_H) hipFree(c->d_H);
    if (c->d_tw) hipFree(c->d_tw);
    if (c->d_fs) hipFree(c->d_fs);
    c->ring = nullptr; c->d_geom = nullptr; c->d_vscale = nullptr; c->d_H = nullptr; c->d_tw = nullptr;
    c->d_fs = nullptr;
    c->fs_bytes = 0;
    if (c->pring) hipFree(c->pring);
    c->pring = nullptr;
    c->pring_bytes = 0;
    c->pring_n = 0;
    if (c->bring) hipFree(c->bring);
    c->bring = nullptr;
    c->bring_bytes = 0;
    c->bring_n = 0;
    c->band_no_room = c->pair_no_room = false;
    ++c->copies_epoch;
    // the raw-count correction is per detector frame: a new geometry drops it
    if (c->raw.d_dark) hipFree(c->raw.d_dark);
    if (c->raw.d_scale) hipFree(c->raw.d_scale);
    c->raw = RawCorr{};
    c->kind = MR_KIND_NONE;
    c->n = 0;
    ++c->geom_epoch;
    ++c->views_epoch;
}

// float32 device tables of the ramp response (tau = 1) and FFT twiddles for length N
// (followed, for the four-step path, by those for length N / 16); l1 (optional) =
// sum_m |h(m)|, h = the unnormalised inverse DFT of the response (real and even, so a
// forward DFT gives it), in float64
int make_tables(int N, int window, float** d_H, float2** d_tw, double* l1 = nullptr) {
    MR_CHECK(N >= 2 && N <= (1 << kFilterMaxLogN), MR_EVALUE,
             "rows of at most " + std::to_string((1 << kFilterMaxLogN) / 2) +
                 " detector columns are supported (FFT length " + std::to_string(N) + ")");
    std::vector<double> H;
    int rc = ramp_response_host(N / 2, 1.0, window, H);
    if (rc) return rc;
    if (l1) {
        std::vector<std::complex<double>> h(H.begin(), H.end());
        fft64(h);
        double sum = 0.0;
        for (const auto& z : h) sum += std::abs(z.real());
        *l1 = sum;
    }
    std::vector<float> Hf(N);
    const int n_small = N > (1 << kFilterLdsLogN) ? N / 16 : 0;
    std::vector<float2> tw(N + n_small);
    auto twiddle = [](int k, int n) {
        const double ang = -2.0 * M_PI * static_cast<double>(k) / static_cast<double>(n);
        return make_float2(static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang)));
    };
    for (int k = 0; k < N; ++k) {
        Hf[k] = static_cast<float>(H[k]);
        tw[k] = twiddle(k, N);
    }
    for (int k = 0; k < n_small; ++k) tw[N + k] = twiddle(k, n_small);
    MR_HIP(hipMalloc(reinterpret_cast<void**>(d_H), sizeof(float) * N));
    MR_HIP(hipMalloc(reinterpret_cast<void**>(d_tw), sizeof(float2) * tw.size()));
    MR_HIP(hipMemcpy(*d_H, Hf.data(), sizeof(float) * N, hipMemcpyHostToDevice));
    MR_HIP(hipMemcpy(*d_tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
    return MR_OK;
}

void apply_bp_defaults(Ctx* c, bool parallel);

int alloc_ring(Ctx* c, int n_total, int rows, int cols, int window, int storage) {
    c->n_total = n_total;
    c->R = rows;
    c->C = cols;
    c->window = window;
    c->storage = storage;
    c->N = fft_length(cols);
    if (!c->bp_explicit) apply_bp_defaults(c, false);
    // measured: cone updates that fill the device run fastest in launches of 32 (f32) or
    // 64 (16-bit) views
    if (!c->vchunk_explicit) c->bp_vchunk = storage == MR_STORE_F32 ? 32 : 64;
    c->pitch = (cols + 3 + 63) / 64 * 64;
    c->Rp = rows + 3;
    c->view_stride = static_cast<int64_t>(c->Rp) * c->pitch;
    const size_t bytes = static_cast<size_t>(n_total) * c->view_stride * elem_size(storage);
    MR_HIP(hipMalloc(&c->ring, bytes));
    MR_HIP(hipMemsetAsync(c->ring, 0, bytes, c->s_ingest));
    MR_HIP(hipMalloc(reinterpret_cast<void**>(&c->d_geom), sizeof(ViewGeom) * n_total));
    MR_HIP(hipMalloc(reinterpret_cast<void**>(&c->d_vscale), sizeof(float) * n_total));
    std::vector<float> ones(n_total, 1.0f);
    MR_HIP(hipMemcpy(c->d_vscale, ones.data(), sizeof(float) * n_total, hipMemcpyHostToDevice));
    int rc = make_tables(c->N, window, &c->d_H, &c->d_tw, &c->h_l1);
    if (rc) return rc;
    // rows longer than one workgroup's LDS: the four-step scratch, fixed for the
    // geometry's lifetime (captured stream graphs hold its address)
    c->fs_bytes = fft_scratch_bytes(c->N, (rows + 1) / 2);
    if (c->fs_bytes) MR_HIP(hipMalloc(reinterpret_cast<void**>(&c->d_fs), c->fs_bytes));
    MR_HIP(hipStreamSynchronize(c->s_ingest));
    c->received.assign(n_total, 0);
    c->n = 0;
    return MR_OK;
}

// The measured default back-projection configuration of a storage form and beam.
// Cone: 16-bit forms read the band ring (kernel 7, 2 points per thread, 3 views in
// flight, waves in lockstep per view); f32 reads the ring directly (kernel 4, 2 points,
// 2 views).  Parallel: kernel 6, the row-pair copy (a point keeps one detector row pair
// in every view), 4 points per thread, 2 views in flight; a COR search of 75 images of
// up to 1024^2 over 1800 views with its records in memory: 38.7 ms (f32 and f16) against
// 42.4 ms with 2 points, 74.5 ms for f32 on the ring (kernel 4) and 47 ms for f16 on the
// band copy (kernel 7).
void apply_bp_defaults(Ctx* c, bool parallel) {
    const bool f32 = c->storage == MR_STORE_F32;
    c->bp_layout = 0;
    if (parallel) {
        c->bp_variant = 6;
        c->bp_ppt = 4;
        c->bp_depth = 2;
        c->bp_sync = 0;
    } else {
        c->bp_variant = f32 ? 4 : 7;
        c->bp_ppt = 2;
        c->bp_depth = f32 ? 2 : 3;
        c->bp_sync = f32 ? 0 : 1;
    }
}

int validate_tasks(const mr_task* tasks, int ntasks, int64_t out_len) {
    const std::string m = validate_tasks_msg(tasks, ntasks, out_len);
    return m.empty() ? MR_OK : fail(MR_EVALUE, m);
}

enum TaskOp { OP_BP = 0, OP_ADJOINT = 1, OP_DSHIFT = 2 };

// Copies the tasks and a tile list to the device request buffer (through the pinned
// staging buffer, on the request stream).
int upload_request(Ctx* c, const mr_task* tasks, int ntasks, const std::vector<Tile>& tiles,
                   const mr_task** d_tasks, const Tile** d_tiles) {
    const size_t tb = sizeof(mr_task) * ntasks;
    const size_t tl = sizeof(Tile) * tiles.size();
    const size_t off_tiles = (tb + 255) / 256 * 256;
    const size_t need = off_tiles + tl + 256;
    int rc = ensure(&c->d_req, &c->req_bytes, need);
    if (rc) return rc;
    rc = ensure_host(&c->h_req, &c->h_req_bytes, need);
    if (rc) return rc;
    std::memcpy(c->h_req, tasks, tb);
    if (tl) std::memcpy(static_cast<char*>(c->h_req) + off_tiles, tiles.data(), tl);
    MR_HIP(hipMemcpyAsync(c->d_req, c->h_req, off_tiles + tl, hipMemcpyHostToDevice, c->s_req));
    *d_tasks = static_cast<const mr_task*>(c->d_req);
    *d_tiles = reinterpret_cast<const Tile*>(static_cast<char*>(c->d_req) + off_tiles);
    return MR_OK;
}

// Volume requests of a cone geometry whose views have no z term in P[0] and P[2]
// (Ctx::zsep), in the default configuration, run the z-column kernel: every task a stack
// of at least half a chunk of planes along z (c = (0, 0, h), no in-box mask), enough
// column tiles to fill the device, and byte offsets of the ring (or its row-pair copy)
// within the kernel's 32-bit and 24-bit products.
bool zcol_request(const Ctx* c, const mr_task* tasks, int ntasks, int n_views) {
    if (c->kind != MR_KIND_CONE || !c->zsep || c->bp_explicit || n_views <= 0) return false;
    if (4 * static_cast<int64_t>(c->R + 3) * c->pitch >= (int64_t{1} << 31) ||
        (static_cast<int64_t>(c->pitch) << 3) >= (int64_t{1} << 23))
        return false;
    int64_t ntiles = 0;
    for (int i = 0; i < ntasks; ++i) {
        const mr_task& T = tasks[i];
        if (T.nq < ZQ / 2 || T.mask_box || T.c[0] != 0.0 || T.c[1] != 0.0 || T.c[2] == 0.0)
            return false;
        ntiles += static_cast<int64_t>((T.nq + ZQ - 1) / ZQ) * ((T.np + ZT - 1) / ZT) *
                  ((T.ns + ZT - 1) / ZT);
    }
    return ntiles >= 4 * c->n_cus && ntiles < (int64_t{1} << 31);
}

// The z-column path of a volume request (zcol_request).
int run_zcol(Ctx* c, const mr_task* tasks, int ntasks, int n_views, float* d_out) {
    int form = ZCOL_RING;
    int rc = zcol_form(c, n_views, c->s_req, &form);
    if (rc) return rc;
    std::vector<Tile> tiles;
    build_zcol_tiles(tasks, ntasks, form == ZCOL_STAGED ? ZQL : ZQ, tiles);
    const mr_task* d_tasks = nullptr;
    const Tile* d_tiles = nullptr;
    rc = upload_request(c, tasks, ntasks, tiles, &d_tasks, &d_tiles);
    if (rc) return rc;
    const float scale = static_cast<float>(M_PI / n_views);
    c->tile_timed = false;
    c->tile_launches = 0;
    c->req_slices = false;
    MR_HIP(hipEventRecord(c->ev0, c->s_req));
    rc = launch_bp_zcol(c, form, d_tasks, ntasks, d_tiles, static_cast<int>(tiles.size()), n_views,
                        scale, d_out, c->s_req);
    if (rc) return rc;
    MR_HIP(hipEventRecord(c->ev1, c->s_req));
    return MR_OK;
}

// Requests of a parallel geometry in the default configuration whose tasks are axial
// planes in groups differing only in their rotation-centre shift, the shifts on a grid of
// 1/q px (q = 1, 2, 4) with at least 4 per group on average (a COR search), run the
// shift-group kernel; its detector rows must stay within the 32-bit offsets of the kernel.
bool cor_request(const Ctx* c, const mr_task* tasks, int ntasks, int n_views, int nvt,
                 CorPlan& plan) {
    if (c->kind != MR_KIND_PARALLEL || c->bp_explicit || n_views <= 0) return false;
    if (static_cast<int64_t>(nvt) * cor_row_floats(c->C) * 4 >= (int64_t{1} << 31)) return false;
    return cor_shift_groups(tasks, ntasks, 4.0, plan);
}

// Device layout of a shift-group request in one buffer: rows [nsets][nvt][lw] floats,
// coefficients [nsets][nvt][4] doubles, sets, groups, tiles (256-byte aligned parts).
struct CorLayout {
    size_t rows = 0, coef = 0, sets = 0, groups = 0, tiles = 0, total = 0;
};
CorLayout cor_layout(const CorPlan& P, int nvt, int C) {
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    CorLayout L;
    const size_t nsets = P.sets.size();
    L.rows = 0;
    L.coef = up(sizeof(float) * nsets * nvt * cor_row_floats(C));
    L.sets = L.coef + up(sizeof(double) * 4 * nsets * nvt);
    L.groups = L.sets + up(sizeof(CorRowSet) * nsets);
    L.tiles = L.groups + up(sizeof(CorGroup) * P.groups.size());
    L.total = L.tiles + up(sizeof(CorTile) * P.tiles.size());
    return L;
}

// Copies the tables of P into `buf` (layout L) on stream s through the pinned request
// staging buffer and fills a launch description (d_tasks: the tasks on the device).
int cor_prepare(Ctx* c, const CorPlan& P, const CorLayout& L, void* buf, const mr_task* d_tasks,
                int nvt, CorLaunch& out, hipStream_t s) {
    const size_t nt = L.total - L.sets;
    int rc = ensure_host(&c->h_req, &c->h_req_bytes, nt);
    if (rc) return rc;
    char* h = static_cast<char*>(c->h_req);
    std::memcpy(h, P.sets.data(), sizeof(CorRowSet) * P.sets.size());
    std::memcpy(h + (L.groups - L.sets), P.groups.data(), sizeof(CorGroup) * P.groups.size());
    std::memcpy(h + (L.tiles - L.sets), P.tiles.data(), sizeof(CorTile) * P.tiles.size());
    char* d = static_cast<char*>(buf);
    MR_HIP(hipMemcpyAsync(d + L.sets, h, nt, hipMemcpyHostToDevice, s));
    // the staging buffer is reused by the next request: wait for the copy
    MR_HIP(hipStreamSynchronize(s));
    out.q = P.q;
    out.nsets = static_cast<int>(P.sets.size());
    out.ntiles = static_cast<int>(P.tiles.size());
    out.d_tasks = d_tasks;
    out.d_sets = reinterpret_cast<const CorRowSet*>(d + L.sets);
    out.d_groups = reinterpret_cast<const CorGroup*>(d + L.groups);
    out.d_tiles = reinterpret_cast<const CorTile*>(d + L.tiles);
    out.d_rows = reinterpret_cast<float*>(d + L.rows);
    out.d_coef = reinterpret_cast<double*>(d + L.coef);
    out.lw = cor_row_floats(c->C);
    out.nvt = nvt;
    return MR_OK;
}

// The shift-group path of a request (cor_request).
int run_cor(Ctx* c, const mr_task* tasks, int ntasks, int n_views, const CorPlan& P, float* d_out) {
    const CorLayout L = cor_layout(P, n_views, c->C);
    int rc = ensure(&c->d_cor, &c->cor_bytes, L.total);
    if (rc) return rc;
    const size_t tb = sizeof(mr_task) * ntasks;
    rc = ensure(&c->d_req, &c->req_bytes, tb);
    if (rc) return rc;
    rc = ensure_host(&c->h_req, &c->h_req_bytes, tb);
    if (rc) return rc;
    std::memcpy(c->h_req, tasks, tb);
    MR_HIP(hipMemcpyAsync(c->d_req, c->h_req, tb, hipMemcpyHostToDevice, c->s_req));
    MR_HIP(hipStreamSynchronize(c->s_req));
    CorLaunch CL;
    rc = cor_prepare(c, P, L, c->d_cor, static_cast<const mr_task*>(c->d_req), n_views, CL, c->s_req);
    if (rc) return rc;
    CL.n_views = n_views;
    CL.d_out = d_out;
    c->tile_timed = false;
    c->tile_launches = 1;
    c->req_slices = true;
    MR_HIP(hipEventRecord(c->ev0, c->s_req));
    rc = launch_cor_groups(c, CL, c->s_req);
    if (rc) return rc;
    MR_HIP(hipEventRecord(c->ev1, c->s_req));
    c->kern_used = 10;
    return MR_OK;
}

// Uploads tasks + tiles and launches the tile kernel of `op` (device buffers), timing it.
// For OP_BP and OP_DSHIFT every output element of every task is written: the tiles
// culled by the box test are filled with zeros by their own small launch.
int run_tasks(Ctx* c, const mr_task* tasks, int ntasks, int n_views, float* d_out,
              int op = OP_BP, const float* d_g = nullptr) {
    if (op == OP_BP && zcol_request(c, tasks, ntasks, n_views))
        return run_zcol(c, tasks, ntasks, n_views, d_out);
    if (op == OP_BP) {
        CorPlan P;
        if (cor_request(c, tasks, ntasks, n_views, n_views, P))
            return run_cor(c, tasks, ntasks, n_views, P, d_out);
    }
    // the tile list depends only on the tasks (the culled-tile list and the block order
    // on the operation and the XCD map): a request repeating one of the last few (a
    // viewer re-requesting its slices as views arrive, one call per plane) reuses it
    // instead of enumerating and sorting again (~0.1 ms on the host)
    const int key_flags = (op == OP_ADJOINT ? 1 : 0) | (c->xcd_map << 1);
    const TileCache& entry = cached_tiles(c->tile_cache, tasks, ntasks, key_flags,
                                          op != OP_ADJOINT, c->xcd_map == 2);
    const std::vector<Tile>& tiles = entry.tiles;
    const size_t ne = entry.ne;
    const mr_task* d_tasks = nullptr;
    const Tile* d_tiles = nullptr;
    int rc = upload_request(c, tasks, ntasks, tiles, &d_tasks, &d_tiles);
    if (rc) return rc;
    const float scale = n_views > 0 ? static_cast<float>(M_PI / n_views) : 0.0f;
    const int nt = static_cast<int>(tiles.size() - ne);
    if (ne) {
        rc = launch_zero_tiles(d_tasks, d_tiles + nt, static_cast<int>(ne), d_out, c->s_req);
        if (rc) return rc;
    }
    c->tile_timed = false;
    c->tile_launches = 0;
    c->req_slices = true;
    for (int i = 0; i < ntasks; ++i) c->req_slices = c->req_slices && tasks[i].nq <= 1;
    MR_HIP(hipEventRecord(c->ev0, c->s_req));
    if (op == OP_BP)
        rc = launch_bp_tiles(c, d_tasks, d_tiles, nt, n_views, scale, d_out, c->s_req);
    else if (op == OP_ADJOINT)
        rc = launch_bp_adjoint(c, d_tasks, d_tiles, nt, n_views, scale, d_g, d_out, c->s_req);
    else
        rc = launch_bp_dshift(c, d_tasks, d_tiles, nt, n_views, scale, d_out, c->s_req);
    if (rc) return rc;
    MR_HIP(hipEventRecord(c->ev1, c->s_req));
    return MR_OK;
}

int finish_timing(Ctx* c) {
    MR_HIP(hipStreamSynchronize(c->s_req));
    float ms = 0.0f;
    MR_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_kernel_ms = ms;
    c->last_tile_ms = ms;
    c->last_tile_launches = c->tile_launches;
    c->last_kern = c->kern_used;
    c->kern_used = 0;
    if (c->tile_timed) {
        MR_HIP(hipEventElapsedTime(&ms, c->evk0, c->evk1));
        c->last_tile_ms = ms;
    }
    return MR_OK;
}

int check_ctx(mr_ctx* ctx) {
    if (!ctx) return fail(MR_EVALUE, "null context");
    return MR_OK;
}

}  // namespace

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int ensure(void** p, size_t* have, size_t need) {
    if (*p && *have >= need) return MR_OK;
    if (*p) hipFree(*p);
    *p = nullptr;
    *have = 0;
    const size_t bytes = std::max<size_t>((need + (1u << 20) - 1) & ~static_cast<size_t>((1u << 20) - 1),
                                          1u << 20);
    MR_HIP(hipMalloc(p, bytes));
    *have = bytes;
    return MR_OK;
}

int ensure_host(void** p, size_t* have, size_t need) {
    if (*p && *have >= need) return MR_OK;
    if (*p) hipHostFree(*p);
    *p = nullptr;
    *have = 0;
    const size_t bytes = std::max<size_t>((need + (1u << 16) - 1) & ~static_cast<size_t>((1u << 16) - 1),
                                          1u << 16);
    MR_HIP(hipHostMalloc(p, bytes, hipHostMallocDefault));
    *have = bytes;
    return MR_OK;
}

int ramp_response_host(int n, double tau, int window, std::vector<double>& H) {
    if (n < 1) return fail(MR_EVALUE, "ramp_response: n must be >= 1");
    if (!valid_window(window)) return fail(MR_EVALUE, "unknown filter window " + std::to_string(window));
    if (!(tau > 0.0) || !std::isfinite(tau)) return fail(MR_EVALUE, "ramp_response: tau must be > 0");
    if (n > (1 << 28)) return fail(MR_EVALUE, "ramp_response: n must be <= 2^28");
    const int N = fft_length(n);
    std::vector<std::complex<double>> h(N, 0.0);
    h[0] = 1.0 / (4.0 * tau * tau);
    for (int i = 1; i < N; ++i) {
        const int idx = i < N / 2 ? i : i - N;
        if (idx % 2 != 0) {
            const double t = M_PI * idx * tau;
            h[i] = -1.0 / (t * t);
        }
    }
    fft64(h);
    H.resize(N);
    const double invN = 1.0 / N;
    for (int k = 0; k < N; ++k) {
        const int kk = k < N / 2 ? k : k - N;       // numpy.fft.fftfreq ordering
        const double f = kk * invN;
        H[k] = h[k].real() * tau * window_at(window, f);
    }
    return MR_OK;
}

}  // namespace mr

using namespace mr;

struct mr_ctx : Ctx {};
struct mr_plan : Plan {};

extern "C" {

const char* mr_last_error(void) { return g_err.c_str(); }

int mr_abi_version(void) { return kAbiVersion; }

int mr_device_count(int* n) {
    if (!n) return fail(MR_EVALUE, "null output");
    int k = 0;
    if (hipGetDeviceCount(&k) != hipSuccess) k = 0;
    *n = k;
    return MR_OK;
}

int mr_ramp_response(int n, double tau, int window, double* H, int* N) {
    std::vector<double> h;
    int rc = ramp_response_host(n, tau, window, h);
    if (rc) return rc;
    if (N) *N = static_cast<int>(h.size());
    if (H) std::memcpy(H, h.data(), sizeof(double) * h.size());
    return MR_OK;
}

int mr_ctx_create(int device, mr_ctx** out) {
    if (!out) return fail(MR_EVALUE, "null output");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        return fail(MR_ERUNTIME, "no HIP device available");
    if (device < 0 || device >= count)
        return fail(MR_EVALUE, "device " + std::to_string(device) + " out of range");
    MR_HIP(hipSetDevice(device));
    mr_ctx* c = new mr_ctx();
    c->device = device;
    if (hipDeviceGetAttribute(&c->n_cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        c->n_cus < 1)
        c->n_cus = 256;
    c->bp_explicit = std::getenv("MINIRECAST_BP_VARIANT") || std::getenv("MINIRECAST_BP_PPT") ||
                     std::getenv("MINIRECAST_BP_LAYOUT") || std::getenv("MINIRECAST_BP_DEPTH") ||
                     std::getenv("MINIRECAST_BP_SYNC");
    if (const char* v = std::getenv("MINIRECAST_BP_VARIANT")) {
        const int k = std::atoi(v);
        c->bp_variant = (k == 2 || (k >= 4 && k <= 7)) ? k : c->bp_variant;
    }
    if (const char* v = std::getenv("MINIRECAST_BP_PPT")) {
        const int k = std::atoi(v);
        c->bp_ppt = (k == 1 || k == 2) ? k : 4;
    }
    if (const char* v = std::getenv("MINIRECAST_BP_LAYOUT")) c->bp_layout = std::atoi(v) ? 1 : 0;
    if (const char* v = std::getenv("MINIRECAST_BP_DEPTH")) c->bp_depth = std::atoi(v) >= 3 ? 3 : 2;
    if (const char* v = std::getenv("MINIRECAST_BP_SYNC")) c->bp_sync = std::atoi(v) ? 1 : 0;
    if (const char* v = std::getenv("MINIRECAST_BP_VCHUNK")) {
        c->bp_vchunk = std::max(0, std::atoi(v));
        c->vchunk_explicit = true;
    }
    if (const char* v = std::getenv("MINIRECAST_BP_VCHUNK_ANY")) c->vchunk_any = std::atoi(v) != 0;
    if (const char* v = std::getenv("MINIRECAST_BP_RECORD_KB")) {
        const long kb = std::atol(v);
        if (kb >= 1) c->rec_budget = static_cast<size_t>(kb) << 10;
    }
    if (const char* v = std::getenv("MINIRECAST_COPY_RESERVE_MB")) {
        const long long mb = std::atoll(v);
        if (mb >= 0) c->copy_reserve = static_cast<size_t>(mb) << 20;
    }
    if (const char* v = std::getenv("MINIRECAST_XCD_MAP")) {
        const int k = std::atoi(v);
        c->xcd_map = (k == 1 || k == 2) ? k : 0;
    }
    // default-priority, non-blocking streams
    if (hipStreamCreateWithFlags(&c->s_ingest, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->s_req, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->evk0) != hipSuccess || hipEventCreate(&c->evk1) != hipSuccess) {
        delete c;
        return fail(MR_ERUNTIME, "stream/event creation failed");
    }
    // the address of this GPU's static scratch (mr_clip_histogram_static), once per
    // device with the device current; a failure is reported by that entry point
    static_scratch_init(device);
    *out = c;
    return MR_OK;
}

int mr_ctx_destroy(mr_ctx* c) {
    if (!c) return MR_OK;
    hipSetDevice(c->device);
    hipDeviceSynchronize();
    free_geometry(c);
    void* bufs[] = {c->d_stage, c->d_fstage, c->d_vmax, c->d_fv, c->d_req, c->d_out, c->d_pts, c->d_misc,
                    c->d_pre, c->d_recs, c->d_raw, c->d_vol, c->d_stat, c->d_corws, c->d_adj, c->d_part,
                    c->d_cor};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (c->h_req) hipHostFree(c->h_req);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->evk0) hipEventDestroy(c->evk0);
    if (c->evk1) hipEventDestroy(c->evk1);
    if (c->s_ingest) hipStreamDestroy(c->s_ingest);
    if (c->s_req) hipStreamDestroy(c->s_req);
    delete c;
    return MR_OK;
}

int mr_set_geometry_cone(mr_ctx* c, int n_total, int rows, int cols, const double* vectors,
                         int window, int storage) {
    int rc = check_ctx(c);
    if (rc) return rc;
    // geometry and view-count changes exclude ingest and requests (ingest lock first)
    std::lock_guard<std::mutex> li(c->ingest_mu);
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(n_total >= 1 && rows >= 1 && cols >= 1, MR_EVALUE, "empty geometry");
    MR_CHECK(vectors != nullptr, MR_EVALUE, "null vectors");
    MR_CHECK(valid_window(window), MR_EVALUE, "unknown filter window " + std::to_string(window));
    MR_CHECK(valid_storage(storage), MR_EVALUE, "unknown storage " + std::to_string(storage));
    rc = check_filter_cols(cols);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    MR_HIP(hipDeviceSynchronize());
    free_geometry(c);
    const int N = fft_length(cols);
    c->geom_all.assign(n_total, ViewGeom{});
    c->filt_all.assign(n_total, FilterView{});
    const double cc = 0.5 * (cols - 1), cr = 0.5 * (rows - 1);
    for (int k = 0; k < n_total; ++k) {
        // cone_constants (recon.py:70-82)
        const double* V = vectors + 12 * k;
        const V3 s = v3(V), d = v3(V + 3), u = v3(V + 6), v = v3(V + 9);
        V3 n = cross(u, v);
        n = (1.0 / norm(n)) * n;
        const double sg = dot(d - s, n);
        n = (sg > 0 ? 1.0 : (sg < 0 ? -1.0 : 0.0)) * n;
        const double sdd = dot(d - s, n);
        const double sod = dot(-1.0 * s, n);
        const V3 uu = (1.0 / dot(u, u)) * u;
        const V3 vv = (1.0 / dot(v, v)) * v;
        MR_CHECK(sod > 0 && sdd > 0, MR_EVALUE,
                 "view " + std::to_string(k) + ": source and detector must lie on opposite sides of the origin");
        // projective form (see ViewGeom): col - cc = (((S-D).UU) L + SDD (UU.rel)) / L
        ViewGeom g{};
        const V3 au = dot(s - d, uu) * n + sdd * uu;
        const V3 av = dot(s - d, vv) * n + sdd * vv;
        set_row(g.P[0], (1.0 / sod) * au, -dot(au, s) / sod);
        set_row(g.P[1], (1.0 / sod) * av, -dot(av, s) / sod);
        set_row(g.P[2], (1.0 / sod) * n, -dot(n, s) / sod);
        g.uvv = dot(u, vv);
        c->geom_all[k] = g;
        // FDK weight (recon.py:93-102)
        const V3 p0 = s + sdd * n;
        const V3 off = d - p0;
        const double un = norm(u), vn = norm(v);
        FilterView f{};
        f.a0 = static_cast<float>(dot(off, u) / un - cc * un);
        f.da = static_cast<float>(un);
        f.b0 = static_cast<float>(dot(off, v) / vn - cr * vn);
        f.db = static_cast<float>(vn);
        f.sdd = static_cast<float>(sdd);
        f.sdd2 = static_cast<float>(sdd * sdd);
        const double tau_iso = un * sod / sdd;
        f.gain = static_cast<float>(1.0 / (N * tau_iso));
        f.cone = 1;
        c->filt_all[k] = f;
    }
    rc = alloc_ring(c, n_total, rows, cols, window, storage);
    if (rc) {
        free_geometry(c);
        return rc;
    }
    c->zsep = true;
    for (const ViewGeom& g : c->geom_all) c->zsep = c->zsep && g.P[0][2] == 0.0 && g.P[2][2] == 0.0;
    c->kind = MR_KIND_CONE;
    return MR_OK;
}

int mr_set_geometry_parallel(mr_ctx* c, int n_total, int rows, int cols, const double* angles,
                             double pixel_u, double pixel_v, int window, int storage) {
    int rc = check_ctx(c);
    if (rc) return rc;
    // geometry and view-count changes exclude ingest and requests (ingest lock first)
    std::lock_guard<std::mutex> li(c->ingest_mu);
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(n_total >= 1 && rows >= 1 && cols >= 1, MR_EVALUE, "empty geometry");
    MR_CHECK(angles != nullptr, MR_EVALUE, "null angles");
    MR_CHECK(pixel_u > 0 && pixel_v > 0, MR_EVALUE, "pixel sizes must be > 0");
    MR_CHECK(valid_window(window), MR_EVALUE, "unknown filter window " + std::to_string(window));
    MR_CHECK(valid_storage(storage), MR_EVALUE, "unknown storage " + std::to_string(storage));
    rc = check_filter_cols(cols);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    MR_HIP(hipDeviceSynchronize());
    free_geometry(c);
    const int N = fft_length(cols);
    c->geom_all.assign(n_total, ViewGeom{});
    c->filt_all.assign(n_total, FilterView{});
    for (int k = 0; k < n_total; ++k) {
        // col - cc = (-x sin t + y cos t) / pu ; row - cr = z / pv  (recon.py:185-188)
        const double ct = std::cos(angles[k]), st = std::sin(angles[k]);
        ViewGeom g{};
        set_row(g.P[0], {-st / pixel_u, ct / pixel_u, 0.0}, 0.0);
        set_row(g.P[1], {0.0, 0.0, 1.0 / pixel_v}, 0.0);
        set_row(g.P[2], {0.0, 0.0, 0.0}, 1.0);
        g.uvv = 0.0;
        c->geom_all[k] = g;
        FilterView f{};
        f.gain = static_cast<float>(1.0 / (N * pixel_u));   // filter_rows(block, pixel_u) (recon.py:359)
        f.cone = 0;
        c->filt_all[k] = f;
    }
    c->pixel_u = pixel_u;
    c->pixel_v = pixel_v;
    rc = alloc_ring(c, n_total, rows, cols, window, storage);
    if (rc) {
        free_geometry(c);
        return rc;
    }
    c->zsep = false;
    c->kind = MR_KIND_PARALLEL;
    if (!c->bp_explicit) apply_bp_defaults(c, true);
    return MR_OK;
}

// Caller holds c->ingest_mu (not c->mu): the batch is staged and filtered into the free
// slots [n, n + k) of the ring and d_geom, which no request reads, and only the commit of
// the view count takes c->mu.  in_stage: the block is already in c->d_stage (device).
static int append_views_locked(mr_ctx* c, const int64_t* indices, int k, const float* block,
                               int on_device, int* n_out, bool filter, bool in_stage) {
    int rc = MR_OK;
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(k >= 0, MR_EVALUE, "negative view count");
    if (k == 0) {
        if (n_out) *n_out = c->n;
        return MR_OK;
    }
    MR_CHECK(indices && block, MR_EVALUE, "null indices or block");
    MR_CHECK(c->n + k <= c->n_total, MR_EVALUE,
             "ring full: " + std::to_string(c->n) + " + " + std::to_string(k) + " views > " +
                 std::to_string(c->n_total));
    for (int j = 0; j < k; ++j)
        MR_CHECK(indices[j] >= 0 && indices[j] < c->n_total, MR_EVALUE,
                 "view index " + std::to_string(indices[j]) + " out of range");
    MR_HIP(hipSetDevice(c->device));
    const size_t blk = sizeof(float) * static_cast<size_t>(c->R) * k * c->C;
    const float* d_block = block;
    const bool pre = filter && (c->gaussian > 0.0 || std::fabs(c->tilt) >= 1e-9);
    if (in_stage) {
        d_block = c->d_stage;
    } else if (!on_device || pre) {
        rc = ensure(reinterpret_cast<void**>(&c->d_stage), &c->stage_bytes, blk);
        if (rc) return rc;
        MR_HIP(hipMemcpyAsync(c->d_stage, block, blk,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, c->s_ingest));
        d_block = c->d_stage;
    }
    if (pre) {
        // preprocess.apply (recon.py:357): tilt, then Gaussian, on the staged block
        rc = ensure(reinterpret_cast<void**>(&c->d_pre), &c->pre_bytes, blk);
        if (rc) return rc;
        float* res = nullptr;
        rc = launch_preprocess(c, c->d_stage, c->d_pre, c->R, k, c->C, c->gaussian, c->tilt, &res,
                               c->s_ingest);
        if (rc) return rc;
        d_block = res;
    }
    std::vector<FilterView> fv(k);
    std::vector<ViewGeom> gv(k);
    for (int j = 0; j < k; ++j) {
        fv[j] = c->filt_all[indices[j]];
        fv[j].slot = c->n + j;
        gv[j] = c->geom_all[indices[j]];
    }
    rc = ensure(reinterpret_cast<void**>(&c->d_fv), &c->fv_len, sizeof(FilterView) * k);
    if (rc) return rc;
    MR_HIP(hipMemcpyAsync(c->d_fv, fv.data(), sizeof(FilterView) * k, hipMemcpyHostToDevice, c->s_ingest));
    MR_HIP(hipMemcpyAsync(c->d_geom + c->n, gv.data(), sizeof(ViewGeom) * k, hipMemcpyHostToDevice,
                          c->s_ingest));
    if (c->storage == MR_STORE_F16) {
        // per view: max |x| of (sampled) raw rows, then max |q| of the filtered view
        rc = ensure(reinterpret_cast<void**>(&c->d_vmax), &c->vmax_len, sizeof(unsigned) * kF16VmaxPerView * k);
        if (rc) return rc;
    }
    rc = filter ? launch_filter(c, d_block, k, c->n, c->s_ingest) : launch_store(c, d_block, k, c->s_ingest);
    if (rc) return rc;
    MR_HIP(hipStreamSynchronize(c->s_ingest));
    // the new views become visible to requests
    std::lock_guard<std::mutex> lk(c->mu);
    c->n += k;
    if (n_out) *n_out = c->n;
    return MR_OK;
}

static int append_views(mr_ctx* c, const int64_t* indices, int k, const float* block,
                        int on_device, int* n_out, bool filter) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> li(c->ingest_mu);
    return append_views_locked(c, indices, k, block, on_device, n_out, filter, false);
}

int mr_push_views(mr_ctx* c, const int64_t* indices, int k, const float* block,
                  int block_on_device, int* n_out) {
    return append_views(c, indices, k, block, block_on_device, n_out, true);
}

int mr_set_raw_correction(mr_ctx* c, const float* dark, const float* scale, int op, int frame_rows,
                          int frame_cols, int sample, int flip, int row0, int col0) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> li(c->ingest_mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(dark && scale, MR_EVALUE, "null dark or scale");
    MR_CHECK(op == 0 || op == 1, MR_EVALUE, "op must be 0 (multiply) or 1 (divide)");
    MR_CHECK(sample >= 1 && sample <= 16, MR_EVALUE, "sample must be in 1..16");
    MR_CHECK(frame_rows > 0 && frame_cols > 0, MR_EVALUE, "empty raw frame");
    const int hb = frame_rows / sample, wb = frame_cols / sample;
    MR_CHECK(row0 >= 0 && col0 >= 0 && row0 + c->R <= hb && col0 + c->C <= wb, MR_EVALUE,
             "crop " + std::to_string(c->R) + "x" + std::to_string(c->C) + " at (" +
                 std::to_string(row0) + ", " + std::to_string(col0) + ") exceeds the binned frame " +
                 std::to_string(hb) + "x" + std::to_string(wb));
    MR_HIP(hipSetDevice(c->device));
    const size_t bytes = sizeof(float) * static_cast<size_t>(c->R) * c->C;
    if (!c->raw.d_dark) MR_HIP(hipMalloc(reinterpret_cast<void**>(&c->raw.d_dark), bytes));
    if (!c->raw.d_scale) MR_HIP(hipMalloc(reinterpret_cast<void**>(&c->raw.d_scale), bytes));
    MR_HIP(hipMemcpy(c->raw.d_dark, dark, bytes, hipMemcpyHostToDevice));
    MR_HIP(hipMemcpy(c->raw.d_scale, scale, bytes, hipMemcpyHostToDevice));
    c->raw.H = frame_rows;
    c->raw.W = frame_cols;
    c->raw.sample = sample;
    c->raw.flip = flip ? 1 : 0;
    c->raw.row0 = row0;
    c->raw.col0 = col0;
    c->raw.op = op;
    c->raw.set = 1;
    return MR_OK;
}

int mr_push_raw(mr_ctx* c, const int64_t* indices, int k, const uint16_t* frames,
                int frames_on_device, int* n_out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> li(c->ingest_mu);
    MR_CHECK(c->raw.set, MR_ERUNTIME, "no raw correction set (mr_set_raw_correction)");
    MR_CHECK(k >= 0, MR_EVALUE, "negative view count");
    if (k == 0) {
        if (n_out) *n_out = c->n;
        return MR_OK;
    }
    MR_CHECK(indices && frames, MR_EVALUE, "null indices or frames");
    MR_CHECK(c->n + k <= c->n_total, MR_EVALUE,
             "ring full: " + std::to_string(c->n) + " + " + std::to_string(k) + " views > " +
                 std::to_string(c->n_total));
    MR_HIP(hipSetDevice(c->device));
    const size_t fbytes = sizeof(uint16_t) * static_cast<size_t>(c->raw.H) * c->raw.W * k;
    const uint16_t* d_frames = frames;
    if (!frames_on_device) {
        rc = ensure(reinterpret_cast<void**>(&c->d_raw), &c->raw_bytes, fbytes);
        if (rc) return rc;
        MR_HIP(hipMemcpyAsync(c->d_raw, frames, fbytes, hipMemcpyHostToDevice, c->s_ingest));
        d_frames = c->d_raw;
    }
    const size_t blk = sizeof(float) * static_cast<size_t>(c->R) * k * c->C;
    rc = ensure(reinterpret_cast<void**>(&c->d_stage), &c->stage_bytes, blk);
    if (rc) return rc;
    rc = launch_raw_correct(c, d_frames, k, c->d_stage, c->s_ingest);
    if (rc) return rc;
    return append_views_locked(c, indices, k, c->d_stage, 1, n_out, true, true);
}

int mr_store_views(mr_ctx* c, const int64_t* indices, int k, const float* q, int q_on_device,
                   int* n_out) {
    return append_views(c, indices, k, q, q_on_device, n_out, false);
}

int mr_set_preprocess(mr_ctx* c, double gaussian, double tilt) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(std::isfinite(gaussian) && std::isfinite(tilt), MR_EVALUE, "non-finite preprocessing");
    std::lock_guard<std::mutex> li(c->ingest_mu);
    c->gaussian = gaussian > 0.0 ? gaussian : 0.0;
    c->tilt = tilt;
    return MR_OK;
}

int mr_preprocess(mr_ctx* c, const float* in, int rows, int k, int cols, double gaussian,
                  double tilt, float* out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(rows >= 0 && k >= 0 && cols >= 0, MR_EVALUE, "negative block shape");
    MR_CHECK(std::isfinite(gaussian) && std::isfinite(tilt), MR_EVALUE, "non-finite preprocessing");
    const size_t blk = sizeof(float) * static_cast<size_t>(rows) * k * cols;
    if (blk == 0) return MR_OK;
    MR_CHECK(in && out, MR_EVALUE, "null block");
    std::lock_guard<std::mutex> li(c->ingest_mu);
    MR_HIP(hipSetDevice(c->device));
    rc = ensure(reinterpret_cast<void**>(&c->d_stage), &c->stage_bytes, blk);
    if (rc) return rc;
    rc = ensure(reinterpret_cast<void**>(&c->d_pre), &c->pre_bytes, blk);
    if (rc) return rc;
    MR_HIP(hipMemcpyAsync(c->d_stage, in, blk, hipMemcpyHostToDevice, c->s_ingest));
    float* res = c->d_stage;
    rc = launch_preprocess(c, c->d_stage, c->d_pre, rows, k, cols, gaussian > 0.0 ? gaussian : 0.0,
                           tilt, &res, c->s_ingest);
    if (rc) return rc;
    MR_HIP(hipMemcpyAsync(out, res, blk, hipMemcpyDeviceToHost, c->s_ingest));
    MR_HIP(hipStreamSynchronize(c->s_ingest));
    return MR_OK;
}

int mr_num_views(mr_ctx* c, int* n) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(n != nullptr, MR_EVALUE, "null output");
    std::lock_guard<std::mutex> lk(c->mu);
    *n = c->n;
    return MR_OK;
}

int mr_bp_points(mr_ctx* c, const double* pts, int64_t M, double cor_shift, int n_views,
                 double* out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(M >= 0, MR_EVALUE, "negative point count");
    MR_CHECK(n_views >= 0 && n_views <= c->n, MR_EVALUE,
             "n_views " + std::to_string(n_views) + " outside [0, " + std::to_string(c->n) + "]");
    if (M == 0) return MR_OK;
    MR_CHECK(pts && out, MR_EVALUE, "null points or output");
    if (n_views == 0) {
        std::fill(out, out + M, 0.0);
        return MR_OK;
    }
    MR_HIP(hipSetDevice(c->device));
    rc = ensure(reinterpret_cast<void**>(&c->d_pts), &c->pts_bytes, sizeof(double) * 3 * M);
    if (rc) return rc;
    rc = ensure(reinterpret_cast<void**>(&c->d_out), &c->out_bytes, sizeof(float) * M);
    if (rc) return rc;
    MR_HIP(hipMemcpyAsync(c->d_pts, pts, sizeof(double) * 3 * M, hipMemcpyHostToDevice, c->s_req));
    const float scale = static_cast<float>(M_PI / n_views);
    MR_HIP(hipEventRecord(c->ev0, c->s_req));
    rc = launch_bp_points(c, c->d_pts, M, cor_shift, n_views, scale, c->d_out, c->s_req);
    if (rc) return rc;
    MR_HIP(hipEventRecord(c->ev1, c->s_req));
    std::vector<float> tmp(M);
    MR_HIP(hipMemcpyAsync(tmp.data(), c->d_out, sizeof(float) * M, hipMemcpyDeviceToHost, c->s_req));
    rc = finish_timing(c);
    if (rc) return rc;
    for (int64_t i = 0; i < M; ++i) out[i] = tmp[i];
    return MR_OK;
}

int mr_bp_tasks(mr_ctx* c, const mr_task* tasks, int ntasks, int n_views, float* out,
                int64_t out_len, int out_on_device) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(ntasks >= 0 && out_len >= 0, MR_EVALUE, "negative task count or output length");
    MR_CHECK(n_views >= 0 && n_views <= c->n, MR_EVALUE,
             "n_views " + std::to_string(n_views) + " outside [0, " + std::to_string(c->n) + "]");
    if (ntasks == 0 || out_len == 0) return MR_OK;
    MR_CHECK(tasks && out, MR_EVALUE, "null tasks or output");
    rc = validate_tasks(tasks, ntasks, out_len);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    float* d_out = out;
    if (!out_on_device) {
        rc = ensure(reinterpret_cast<void**>(&c->d_out), &c->out_bytes, sizeof(float) * out_len);
        if (rc) return rc;
        d_out = c->d_out;
        MR_HIP(hipMemsetAsync(d_out, 0, sizeof(float) * out_len, c->s_req));
    }
    rc = run_tasks(c, tasks, ntasks, n_views, d_out);
    if (rc) return rc;
    if (!out_on_device)
        MR_HIP(hipMemcpyAsync(out, d_out, sizeof(float) * out_len, hipMemcpyDeviceToHost, c->s_req));
    return finish_timing(c);
}

namespace {
// Back-projection configurations the autotuner times: kernel (2: LDS records and
// device queue, 4: precomputed records, 5: LDS-staged boxes, cone only), points per
// thread, lane layout (2 x 32 or 4 x 16 waves) and views in flight.
struct BpConfig { int variant, ppt, layout, depth; };
const BpConfig kBpCandidates[] = {
    {4, 2, 0, 2}, {4, 2, 0, 3}, {4, 2, 1, 2}, {4, 2, 1, 3}, {4, 1, 0, 2}, {4, 1, 0, 3},
    {4, 1, 1, 2}, {4, 1, 1, 3}, {4, 4, 0, 2}, {4, 4, 1, 2}, {5, 2, 0, 2}, {5, 4, 0, 2},
    {2, 4, 0, 2}, {6, 2, 0, 2}, {6, 2, 0, 3}, {6, 1, 0, 3}, {6, 4, 0, 2}, {7, 2, 0, 2},
    {7, 1, 0, 2}, {7, 2, 0, 3}};

bool valid_kernel(int variant) { return variant == 2 || (variant >= 4 && variant <= 7); }

bool valid_config(int variant, int ppt, int layout, int depth) {
    return valid_kernel(variant) && (ppt == 1 || ppt == 2 || ppt == 4) &&
           (layout == 0 || layout == 1) && (depth == 2 || depth == 3);
}
}  // namespace

namespace {
// The volume on the device: a caller pointer, or a staged copy of a host array.
int stage_volume(Ctx* c, const float* vol, int64_t n, int on_device, const float** d_v) {
    if (on_device) {
        *d_v = vol;
        return MR_OK;
    }
    int rc = ensure(&c->d_vol, &c->vol_bytes, sizeof(float) * static_cast<size_t>(n));
    if (rc) return rc;
    MR_HIP(hipMemcpyAsync(c->d_vol, vol, sizeof(float) * n, hipMemcpyHostToDevice, c->s_req));
    *d_v = static_cast<const float*>(c->d_vol);
    return MR_OK;
}

float key_float(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float x;
    std::memcpy(&x, &u, 4);
    return x;
}

// np.percentile(x, 100 q), method "linear", from the order statistics a = x_(i), b =
// x_(i+1) of float32 data: virtual index (n - 1) q, gamma its fractional part, the
// difference b - a in float32 (numpy subtracts the two float32 neighbours), the
// interpolation in float64, from b when gamma >= 0.5.
double numpy_lerp(float a, float b, double gamma) {
#pragma clang fp contract(off)
    const float d = b - a;
    if (gamma >= 0.5) return static_cast<double>(b) - static_cast<double>(d) * (1.0 - gamma);
    return static_cast<double>(a) + static_cast<double>(d) * gamma;
}
}  // namespace

int mr_nonzero_percentiles(mr_ctx* c, const float* vol, int64_t n, int on_device,
                           const double* q, int nq, double* out, int64_t* nonzero) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(n >= 0 && (n == 0 || vol), MR_EVALUE, "null volume");
    MR_CHECK(nq >= 1 && nq <= 4 && q && out, MR_EVALUE, "1..4 quantiles with an output");
    for (int i = 0; i < nq; ++i)
        MR_CHECK(q[i] >= 0.0 && q[i] <= 1.0, MR_EVALUE, "quantiles must be in [0, 1]");
    MR_HIP(hipSetDevice(c->device));
    const float* d_v = nullptr;
    if (n > 0) {
        rc = stage_volume(c, vol, n, on_device, &d_v);
        if (rc) return rc;
    }
    constexpr int T = 8;
    const size_t hist_bytes = sizeof(unsigned long long) * T * 256;
    rc = ensure(&c->d_stat, &c->stat_bytes, hist_bytes + 256);
    if (rc) return rc;
    auto* d_hist = static_cast<unsigned long long*>(c->d_stat);
    auto* d_pre = reinterpret_cast<uint32_t*>(static_cast<char*>(c->d_stat) + hist_bytes);
    std::vector<unsigned long long> h(T * 256);
    // pass 0: the top byte of every non-zero key (also the count)
    int64_t m = 0;
    if (n > 0) {
        MR_HIP(hipMemsetAsync(d_hist, 0, hist_bytes, c->s_req));
        rc = launch_radix_hist(d_v, n, d_pre, 1, 24, d_hist, c->s_req);
        if (rc) return rc;
        MR_HIP(hipMemcpyAsync(h.data(), d_hist, sizeof(unsigned long long) * 256,
                              hipMemcpyDeviceToHost, c->s_req));
        MR_HIP(hipStreamSynchronize(c->s_req));
        for (int b = 0; b < 256; ++b) m += static_cast<int64_t>(h[b]);
    }
    if (nonzero) *nonzero = m;
    MR_CHECK(m > 0, MR_EVALUE, "no non-zero values");
    // target ranks (numpy _get_indexes): floor and floor + 1 of (m - 1) q, both m - 1
    // at or above the last index
    std::vector<int64_t> lo(nq), hi(nq);
    std::vector<double> gam(nq);
    std::vector<int64_t> ranks;
    for (int i = 0; i < nq; ++i) {
        const double vi = static_cast<double>(m - 1) * q[i];
        if (vi >= static_cast<double>(m - 1)) {
            lo[i] = hi[i] = m - 1;
        } else {
            lo[i] = static_cast<int64_t>(std::floor(vi));
            hi[i] = lo[i] + 1;
        }
        gam[i] = vi - std::floor(vi);
        ranks.push_back(lo[i]);
        ranks.push_back(hi[i]);
    }
    std::sort(ranks.begin(), ranks.end());
    ranks.erase(std::unique(ranks.begin(), ranks.end()), ranks.end());
    const int nt = static_cast<int>(ranks.size());
    std::vector<uint32_t> prefix(nt, 0u);
    std::vector<int64_t> rem(ranks);
    auto walk = [&](const unsigned long long* hist, int t, int shift) {
        int64_t cum = 0;
        for (int b = 0; b < 256; ++b) {
            const int64_t cnt = static_cast<int64_t>(hist[b]);
            if (rem[t] < cum + cnt) {
                prefix[t] |= static_cast<uint32_t>(b) << shift;
                rem[t] -= cum;
                return;
            }
            cum += cnt;
        }
    };
    for (int t = 0; t < nt; ++t) walk(h.data(), t, 24);
    for (int shift = 16; shift >= 0; shift -= 8) {
        MR_HIP(hipMemsetAsync(d_hist, 0, hist_bytes, c->s_req));
        MR_HIP(hipMemcpyAsync(d_pre, prefix.data(), sizeof(uint32_t) * nt, hipMemcpyHostToDevice,
                              c->s_req));
        rc = launch_radix_hist(d_v, n, d_pre, nt, shift, d_hist, c->s_req);
        if (rc) return rc;
        MR_HIP(hipMemcpyAsync(h.data(), d_hist, sizeof(unsigned long long) * 256 * nt,
                              hipMemcpyDeviceToHost, c->s_req));
        MR_HIP(hipStreamSynchronize(c->s_req));
        for (int t = 0; t < nt; ++t) walk(h.data() + 256 * t, t, shift);
    }
    auto value_of = [&](int64_t r) {
        const int t = static_cast<int>(std::lower_bound(ranks.begin(), ranks.end(), r) - ranks.begin());
        return key_float(prefix[t]);
    };
    for (int i = 0; i < nq; ++i) out[i] = numpy_lerp(value_of(lo[i]), value_of(hi[i]), gam[i]);
    return MR_OK;
}

int mr_clip_histogram(mr_ctx* c, const float* vol, int64_t n, int on_device, double lo, double hi,
                      int bins, const double* edges, int64_t* hist) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(n >= 0 && (n == 0 || vol), MR_EVALUE, "null volume");
    MR_CHECK(bins >= 1 && bins <= 8192 && edges && hist, MR_EVALUE, "1..8192 bins with edges");
    MR_CHECK(std::isfinite(lo) && std::isfinite(hi) && lo < hi, MR_EVALUE,
             "range must be finite with lo < hi");
    MR_HIP(hipSetDevice(c->device));
    const size_t hb = sizeof(unsigned long long) * bins, eb = sizeof(double) * (bins + 1);
    rc = ensure(&c->d_stat, &c->stat_bytes, hb + eb + 256);
    if (rc) return rc;
    auto* d_hist = static_cast<unsigned long long*>(c->d_stat);
    auto* d_edges = reinterpret_cast<double*>(static_cast<char*>(c->d_stat) + (hb + 255) / 256 * 256);
    MR_CHECK((hb + 255) / 256 * 256 + eb <= c->stat_bytes, MR_ERUNTIME, "statistics buffer");
    MR_HIP(hipMemsetAsync(d_hist, 0, hb, c->s_req));
    MR_HIP(hipMemcpyAsync(d_edges, edges, eb, hipMemcpyHostToDevice, c->s_req));
    if (n > 0) {
        const float* d_v = nullptr;
        rc = stage_volume(c, vol, n, on_device, &d_v);
        if (rc) return rc;
        rc = launch_clip_hist(d_v, n, lo, hi, bins, d_edges, d_hist, c->s_req);
        if (rc) return rc;
    }
    std::vector<unsigned long long> h(bins);
    MR_HIP(hipMemcpyAsync(h.data(), d_hist, hb, hipMemcpyDeviceToHost, c->s_req));
    MR_HIP(hipStreamSynchronize(c->s_req));
    for (int b = 0; b < bins; ++b) hist[b] = static_cast<int64_t>(h[b]);
    return MR_OK;
}

int mr_set_bp_config(mr_ctx* c, int variant, int ppt, int layout, int depth) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    if (variant == 0) {
        // back to the defaults of the geometry's storage form and beam
        c->bp_explicit = false;
        apply_bp_defaults(c, c->kind == MR_KIND_PARALLEL);
        return MR_OK;
    }
    MR_CHECK(valid_config(variant, ppt, layout, depth), MR_EVALUE,
             "invalid back-projection configuration");
    c->bp_variant = variant;
    c->bp_ppt = ppt;
    c->bp_layout = layout;
    c->bp_depth = depth;
    c->bp_explicit = true;
    return MR_OK;
}

int mr_set_ring_copy_budget(mr_ctx* c, int64_t bytes) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_HIP(hipSetDevice(c->device));
    c->copy_budget = bytes < 0 ? -1 : bytes;
    trim_ring_copies(c, c->s_req);
    return MR_OK;
}

int mr_device_bytes(mr_ctx* c, int64_t* out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(out != nullptr, MR_EVALUE, "null output");
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t ring = c->ring ? static_cast<size_t>(c->n_total) * c->view_stride *
                                      (c->storage == MR_STORE_F32 ? 4 : 2)
                                : 0;
    out[0] = static_cast<int64_t>(ring);
    out[1] = c->bring ? static_cast<int64_t>(c->bring_bytes) : 0;
    out[2] = c->pring ? static_cast<int64_t>(c->pring_bytes) : 0;
    out[3] = static_cast<int64_t>(c->fs_bytes);
    return MR_OK;
}

int mr_get_bp_config(mr_ctx* c, int* cfg) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(cfg, MR_EVALUE, "null configuration");
    std::lock_guard<std::mutex> lk(c->mu);
    cfg[0] = c->bp_variant;
    cfg[1] = c->bp_ppt;
    cfg[2] = c->bp_layout;
    cfg[3] = c->bp_depth;
    return MR_OK;
}

int mr_autotune(mr_ctx* c, const mr_task* tasks, int ntasks, int n_views, int64_t out_len,
                int max_tiles, int reps, int* cfg, float* best_ms) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(ntasks >= 1 && tasks && out_len >= 1, MR_EVALUE, "autotuning needs tasks and an output");
    MR_CHECK(n_views >= 1 && n_views <= c->n, MR_EVALUE,
             "n_views " + std::to_string(n_views) + " outside [1, " + std::to_string(c->n) + "]");
    MR_CHECK(max_tiles >= 1 && reps >= 1 && reps <= 100, MR_EVALUE, "max_tiles >= 1, reps in 1..100");
    rc = validate_tasks(tasks, ntasks, out_len);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    std::vector<Tile> all, tiles;
    build_tiles(tasks, ntasks, all);
    // an evenly spaced sample of at most max_tiles tiles keeps tuning cheap
    if (static_cast<int64_t>(all.size()) > max_tiles) {
        for (int i = 0; i < max_tiles; ++i)
            tiles.push_back(all[static_cast<size_t>(i) * all.size() / max_tiles]);
    } else {
        tiles = all;
    }
    if (cfg) { cfg[0] = c->bp_variant; cfg[1] = c->bp_ppt; cfg[2] = c->bp_layout; cfg[3] = c->bp_depth; }
    if (best_ms) *best_ms = 0.0f;
    if (tiles.empty()) return MR_OK;
    const size_t tb = sizeof(mr_task) * ntasks, tl = sizeof(Tile) * tiles.size();
    const size_t off_tiles = (tb + 255) / 256 * 256;
    rc = ensure(&c->d_req, &c->req_bytes, off_tiles + tl + 256);
    if (rc) return rc;
    rc = ensure_host(&c->h_req, &c->h_req_bytes, off_tiles + tl + 256);
    if (rc) return rc;
    std::memcpy(c->h_req, tasks, tb);
    std::memcpy(static_cast<char*>(c->h_req) + off_tiles, tiles.data(), tl);
    MR_HIP(hipMemcpyAsync(c->d_req, c->h_req, off_tiles + tl, hipMemcpyHostToDevice, c->s_req));
    rc = ensure(reinterpret_cast<void**>(&c->d_out), &c->out_bytes, sizeof(float) * out_len);
    if (rc) return rc;
    const mr_task* d_tasks = static_cast<const mr_task*>(c->d_req);
    const Tile* d_tiles = reinterpret_cast<const Tile*>(static_cast<char*>(c->d_req) + off_tiles);
    const float scale = static_cast<float>(M_PI / n_views);
    const int nt = static_cast<int>(tiles.size());
    c->req_slices = true;
    for (int i = 0; i < ntasks; ++i) c->req_slices = c->req_slices && tasks[i].nq <= 1;
    BpConfig best{c->bp_variant, c->bp_ppt, c->bp_layout, c->bp_depth};
    float best_t = 1e30f;
    for (const BpConfig& k : kBpCandidates) {
        if (k.variant == 5 && c->kind != MR_KIND_CONE) continue;
        c->bp_variant = k.variant;
        c->bp_ppt = k.ppt;
        c->bp_layout = k.layout;
        c->bp_depth = k.depth;
        rc = launch_bp_tiles(c, d_tasks, d_tiles, nt, n_views, scale, c->d_out, c->s_req);  // warm-up
        if (rc) break;
        MR_HIP(hipEventRecord(c->ev0, c->s_req));
        for (int r = 0; r < reps && !rc; ++r)
            rc = launch_bp_tiles(c, d_tasks, d_tiles, nt, n_views, scale, c->d_out, c->s_req);
        if (rc) break;
        MR_HIP(hipEventRecord(c->ev1, c->s_req));
        MR_HIP(hipEventSynchronize(c->ev1));
        float ms = 0.0f;
        MR_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        ms /= reps;
        if (ms < best_t) { best_t = ms; best = k; }
    }
    c->bp_variant = best.variant;
    c->bp_ppt = best.ppt;
    c->bp_layout = best.layout;
    c->bp_depth = best.depth;
    if (rc) return rc;
    // views per launch, when the whole request was timed and its launches can be split
    const int subs = (best.variant >= 4 && best.variant != 5) ? BP_TP / (8 * best.ppt) : 0;
    if (subs > 0 && tiles.size() == all.size() && c->kind == MR_KIND_CONE && c->req_slices &&
        nt * subs >= 8 * c->n_cus) {
        const int saved = c->bp_vchunk;
        int best_vc = saved;
        for (const int vc : {16, 32, 64, 128, 0}) {
            c->bp_vchunk = vc;
            rc = launch_bp_tiles(c, d_tasks, d_tiles, nt, n_views, scale, c->d_out, c->s_req);
            if (rc) break;
            MR_HIP(hipEventRecord(c->ev0, c->s_req));
            for (int r = 0; r < reps && !rc; ++r)
                rc = launch_bp_tiles(c, d_tasks, d_tiles, nt, n_views, scale, c->d_out, c->s_req);
            if (rc) break;
            MR_HIP(hipEventRecord(c->ev1, c->s_req));
            MR_HIP(hipEventSynchronize(c->ev1));
            float ms = 0.0f;
            MR_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
            ms /= reps;
            if (ms < best_t) { best_t = ms; best_vc = vc; }
        }
        c->bp_vchunk = best_vc;
        c->vchunk_explicit = true;
        if (rc) return rc;
    }
    if (cfg) { cfg[0] = best.variant; cfg[1] = best.ppt; cfg[2] = best.layout; cfg[3] = best.depth; }
    if (best_ms) *best_ms = best_t;
    return MR_OK;
}

int mr_set_bp_views_per_pass(mr_ctx* c, int views) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(views >= 0, MR_EVALUE, "views per pass must be >= 0 (0: all views in one launch)");
    std::lock_guard<std::mutex> lk(c->mu);
    c->bp_vchunk = views;
    c->vchunk_explicit = true;
    return MR_OK;
}

int mr_get_bp_views_per_pass(mr_ctx* c, int* views) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(views != nullptr, MR_EVALUE, "null output");
    std::lock_guard<std::mutex> lk(c->mu);
    *views = c->bp_vchunk;
    return MR_OK;
}

int mr_bp_tasks_dshift(mr_ctx* c, const mr_task* tasks, int ntasks, int n_views, float* out,
                       int64_t out_len, int out_on_device) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(ntasks >= 0 && out_len >= 0, MR_EVALUE, "negative task count or output length");
    MR_CHECK(n_views >= 0 && n_views <= c->n, MR_EVALUE, "n_views outside the stored range");
    if (ntasks == 0 || out_len == 0) return MR_OK;
    MR_CHECK(tasks && out, MR_EVALUE, "null tasks or output");
    rc = validate_tasks(tasks, ntasks, out_len);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    float* d_out = out;
    if (!out_on_device) {
        rc = ensure(reinterpret_cast<void**>(&c->d_out), &c->out_bytes, sizeof(float) * out_len);
        if (rc) return rc;
        d_out = c->d_out;
        MR_HIP(hipMemsetAsync(d_out, 0, sizeof(float) * out_len, c->s_req));
    }
    rc = run_tasks(c, tasks, ntasks, n_views, d_out, OP_DSHIFT);
    if (rc) return rc;
    if (!out_on_device)
        MR_HIP(hipMemcpyAsync(out, d_out, sizeof(float) * out_len, hipMemcpyDeviceToHost, c->s_req));
    return finish_timing(c);
}

int mr_bp_tasks_adjoint(mr_ctx* c, const mr_task* tasks, int ntasks, int n_views, const float* g,
                        int64_t g_len, int g_on_device, float* gq, int gq_on_device) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(ntasks >= 0 && g_len >= 0, MR_EVALUE, "negative task count or length");
    MR_CHECK(n_views >= 0 && n_views <= c->n, MR_EVALUE, "n_views outside the stored range");
    MR_CHECK(gq != nullptr, MR_EVALUE, "null gradient output");
    const size_t gq_bytes = sizeof(float) * static_cast<size_t>(n_views) * c->R * c->C;
    MR_HIP(hipSetDevice(c->device));
    float* d_gq = gq;
    if (!gq_on_device) {
        rc = ensure(reinterpret_cast<void**>(&c->d_fstage), &c->fstage_bytes, gq_bytes);
        if (rc) return rc;
        d_gq = c->d_fstage;
    }
    if (gq_bytes) MR_HIP(hipMemsetAsync(d_gq, 0, gq_bytes, c->s_req));
    if (ntasks > 0 && g_len > 0 && n_views > 0) {
        MR_CHECK(tasks && g, MR_EVALUE, "null tasks or weights");
        rc = validate_tasks(tasks, ntasks, g_len);
        if (rc) return rc;
        const float* d_g = g;
        if (!g_on_device) {
            rc = ensure(reinterpret_cast<void**>(&c->d_out), &c->out_bytes, sizeof(float) * g_len);
            if (rc) return rc;
            MR_HIP(hipMemcpyAsync(c->d_out, g, sizeof(float) * g_len, hipMemcpyHostToDevice, c->s_req));
            d_g = c->d_out;
        }
        rc = run_tasks(c, tasks, ntasks, n_views, d_gq, OP_ADJOINT, d_g);
        if (rc) return rc;
    }
    if (!gq_on_device && gq_bytes)
        MR_HIP(hipMemcpyAsync(gq, d_gq, gq_bytes, hipMemcpyDeviceToHost, c->s_req));
    MR_HIP(hipStreamSynchronize(c->s_req));
    return MR_OK;
}

int mr_filter_adjoint(mr_ctx* c, const int64_t* indices, int k, const float* gq, int gq_on_device,
                      float* graw, int graw_on_device) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> li(c->ingest_mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(k >= 0, MR_EVALUE, "negative view count");
    if (k == 0) return MR_OK;
    MR_CHECK(indices && gq && graw, MR_EVALUE, "null argument");
    for (int j = 0; j < k; ++j)
        MR_CHECK(indices[j] >= 0 && indices[j] < c->n_total, MR_EVALUE, "view index out of range");
    MR_HIP(hipSetDevice(c->device));
    const size_t bytes = sizeof(float) * static_cast<size_t>(c->R) * k * c->C;
    const float* d_in = gq;
    float* d_o = graw;
    if (!gq_on_device) {
        rc = ensure(reinterpret_cast<void**>(&c->d_stage), &c->stage_bytes, bytes);
        if (rc) return rc;
        MR_HIP(hipMemcpyAsync(c->d_stage, gq, bytes, hipMemcpyHostToDevice, c->s_ingest));
        d_in = c->d_stage;
    }
    if (!graw_on_device) {
        rc = ensure(reinterpret_cast<void**>(&c->d_pre), &c->pre_bytes, bytes);
        if (rc) return rc;
        d_o = c->d_pre;
    }
    std::vector<FilterView> fv(k);
    for (int j = 0; j < k; ++j) fv[j] = c->filt_all[indices[j]];
    rc = ensure(reinterpret_cast<void**>(&c->d_fv), &c->fv_len, sizeof(FilterView) * k);
    if (rc) return rc;
    MR_HIP(hipMemcpyAsync(c->d_fv, fv.data(), sizeof(FilterView) * k, hipMemcpyHostToDevice, c->s_ingest));
    rc = launch_filter_adjoint(c, d_in, k, d_o, c->s_ingest);
    if (rc) return rc;
    if (!graw_on_device)
        MR_HIP(hipMemcpyAsync(graw, d_o, bytes, hipMemcpyDeviceToHost, c->s_ingest));
    MR_HIP(hipStreamSynchronize(c->s_ingest));
    return MR_OK;
}

static int64_t task_candidates(const mr_task& T) {
    if (T.nq <= 0 || T.np <= 0 || T.ns <= 0) return 0;
    return static_cast<int64_t>(T.nq) * ((T.np + BP_TP - 1) / BP_TP) * ((T.ns + BP_TS - 1) / BP_TS);
}

static void plan_free(Plan* p) {
    for (auto& g : p->graphs) hipGraphExecDestroy(g.second);
    p->graphs.clear();
    if (p->cor_exec) hipGraphExecDestroy(p->cor_exec);
    p->cor_exec = nullptr;
    void* dev[] = {p->d_tasks, p->d_toff, p->d_queue, p->d_counters, p->d_out, p->d_block, p->d_meta,
                   p->d_vmax, p->d_scores, p->d_acc, p->d_corws, p->d_cor};
    for (void* d : dev)
        if (d) hipFree(d);
    if (p->h_block) hipHostFree(p->h_block);
    if (p->h_meta) hipHostFree(p->h_meta);
}

int mr_plan_create(mr_ctx* c, const mr_task* tasks, int ntasks, int64_t out_len, mr_plan** out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(out != nullptr, MR_EVALUE, "null output");
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(ntasks >= 1 && tasks, MR_EVALUE, "a plan needs at least one task");
    MR_CHECK(out_len >= 1, MR_EVALUE, "a plan needs an output");
    rc = validate_tasks(tasks, ntasks, out_len);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    mr_plan* p = new mr_plan();
    p->c = c;
    p->ntasks = ntasks;
    p->out_len = out_len;
    p->h_tasks.assign(tasks, tasks + ntasks);
    p->geom_epoch = c->geom_epoch;
    std::vector<int64_t> toff(ntasks + 1, 0);
    for (int i = 0; i < ntasks; ++i) toff[i + 1] = toff[i] + task_candidates(tasks[i]);
    p->total_candidates = toff[ntasks];
    hipError_t e = hipSuccess;
    auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
    chk(hipMalloc(reinterpret_cast<void**>(&p->d_tasks), sizeof(mr_task) * ntasks));
    chk(hipMalloc(reinterpret_cast<void**>(&p->d_toff), sizeof(int64_t) * (ntasks + 1)));
    chk(hipMalloc(reinterpret_cast<void**>(&p->d_queue), sizeof(Tile) * std::max<int64_t>(p->total_candidates, 1)));
    chk(hipMalloc(reinterpret_cast<void**>(&p->d_counters), sizeof(int) * 4));
    chk(hipMalloc(reinterpret_cast<void**>(&p->d_out), sizeof(float) * out_len));
    if (e == hipSuccess) {
        chk(hipMemcpy(p->d_tasks, tasks, sizeof(mr_task) * ntasks, hipMemcpyHostToDevice));
        chk(hipMemcpy(p->d_toff, toff.data(), sizeof(int64_t) * (ntasks + 1), hipMemcpyHostToDevice));
        chk(hipMemset(p->d_out, 0, sizeof(float) * out_len));
        const int nv[2] = {c->n, 0};
        chk(hipMemcpy(p->d_counters + 2, nv, sizeof(int) * 2, hipMemcpyHostToDevice));
    }
    if (e != hipSuccess) {
        plan_free(p);
        delete p;
        return fail(MR_ERUNTIME, std::string("plan allocation: ") + hipGetErrorString(e));
    }
    rc = launch_enumerate_tiles(p->d_tasks, p->d_toff, ntasks, p->total_candidates, p->d_queue,
                                p->d_counters, c->s_req);
    if (!rc) {
        e = hipStreamSynchronize(c->s_req);
        if (e != hipSuccess) rc = fail(MR_ERUNTIME, std::string("plan queue: ") + hipGetErrorString(e));
    }
    if (rc) {
        plan_free(p);
        delete p;
        return rc;
    }
    p->grid = persistent_grid(c);
    *out = p;
    return MR_OK;
}

int mr_plan_destroy(mr_plan* p) {
    if (!p) return MR_OK;
    {
        std::lock_guard<std::mutex> lk(p->c->mu);
        hipSetDevice(p->c->device);
        hipStreamSynchronize(p->c->s_req);
        plan_free(p);
    }
    delete p;
    return MR_OK;
}

static int plan_copy_out(mr_plan* p, float* out_host) {
    if (out_host)
        MR_HIP(hipMemcpyAsync(out_host, p->d_out, sizeof(float) * p->out_len, hipMemcpyDeviceToHost,
                              p->c->s_req));
    return MR_OK;
}

int mr_plan_run(mr_plan* p, int n_views, float* out_host) {
    MR_CHECK(p != nullptr, MR_EVALUE, "null plan");
    Ctx* c = p->c;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(p->geom_epoch == c->geom_epoch, MR_ERUNTIME, "the plan was made for an earlier geometry");
    if (n_views < 0) n_views = c->n;
    MR_CHECK(n_views <= c->n, MR_EVALUE, "n_views beyond the stored views");
    MR_HIP(hipSetDevice(c->device));
    TileQueue q;
    q.head = p->d_counters + 1;
    q.count = p->d_counters;
    q.grid = p->grid;
    const float scale = n_views > 0 ? static_cast<float>(M_PI / n_views) : 0.0f;
    int rc = prepare_persistent_bp(c, n_views, c->s_req);
    if (rc) return rc;
    MR_HIP(hipEventRecord(c->ev0, c->s_req));
    rc = launch_reset_head(p->d_counters, c->s_req);
    if (!rc) rc = launch_bp_tiles(c, p->d_tasks, p->d_queue, 0, n_views, scale, p->d_out, c->s_req, &q);
    if (rc) return rc;
    MR_HIP(hipEventRecord(c->ev1, c->s_req));
    rc = plan_copy_out(p, out_host);
    if (rc) return rc;
    rc = finish_timing(c);
    p->last_ms = c->last_kernel_ms;
    return rc;
}

int mr_plan_cor_scores(mr_plan* p, int n_views, double* scores) {
    MR_CHECK(p != nullptr, MR_EVALUE, "null plan");
    MR_CHECK(scores != nullptr, MR_EVALUE, "null scores");
    Ctx* c = p->c;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(p->geom_epoch == c->geom_epoch, MR_ERUNTIME, "the plan was made for an earlier geometry");
    if (n_views < 0) n_views = c->n;
    MR_CHECK(n_views <= c->n, MR_EVALUE, "n_views beyond the stored views");
    // each task is one image the scoring kernels read as np * ns contiguous pixels from
    // its out_offset (sizes may differ: the trial images of several slices)
    std::vector<int64_t> offs(p->ntasks), npix(p->ntasks);
    for (int i = 0; i < p->ntasks; ++i) {
        const mr_task& T = p->h_tasks[i];
        MR_CHECK(T.nq == 1 && T.stride_p == T.ns && T.stride_s == 1, MR_EVALUE,
                 "score plans need single images of contiguous rows (nq = 1, stride_p = ns, stride_s = 1)");
        offs[i] = T.out_offset;
        npix[i] = static_cast<int64_t>(T.np) * T.ns;
    }
    MR_HIP(hipSetDevice(c->device));
    if (!p->d_scores) MR_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_scores), sizeof(double) * p->ntasks));
    if (!p->d_corws) {
        // kept only once its image descriptors are in place
        void* ws = nullptr;
        MR_HIP(hipMalloc(&ws, cor_workspace_bytes(p->ntasks)));
        std::vector<CorImage> imgs;
        const int chunks = cor_images(offs.data(), npix.data(), p->ntasks, imgs);
        const int rc = write_cor_images(imgs, ws, c->s_req);
        if (rc) {
            hipFree(ws);
            return rc;
        }
        p->d_corws = ws;
        p->cor_chunks = chunks;
    }
    // a COR search (shift groups) runs the shift-group kernel on the ring itself, rows for
    // up to n_total views; other images the persistent v2 launch
    CorPlan CP;
    const bool groups = cor_request(c, p->h_tasks.data(), p->ntasks, n_views, c->n_total, CP);
    int rk = -1;
    if (!groups) {
        int rc0 = prepare_persistent_bp(c, n_views, c->s_req);
        if (rc0) return rc0;
        rk = v2_ring_kind(c);
    }
    if (p->cor_exec && (p->cor_groups != groups ||
                        (!groups && (p->cor_ring != rk || p->cor_copies != c->copies_epoch)))) {
        // the path, the configuration or a ring copy changed since the capture
        hipGraphExecDestroy(p->cor_exec);
        p->cor_exec = nullptr;
    }
    // the device view count of the launch
    MR_HIP(hipMemcpyAsync(p->d_counters + 2, &n_views, sizeof(int), hipMemcpyHostToDevice, c->s_req));
    MR_HIP(hipStreamSynchronize(c->s_req));
    if (!p->cor_exec && groups) {
        p->cor_groups = true;
        const CorLayout L = cor_layout(CP, c->n_total, c->C);
        if (p->d_cor) hipFree(p->d_cor);
        p->d_cor = nullptr;
        MR_HIP(hipMalloc(&p->d_cor, L.total));
        CorLaunch CL;
        int r = cor_prepare(c, CP, L, p->d_cor, p->d_tasks, c->n_total, CL, c->s_req);
        if (r) return r;
        CL.d_nviews = p->d_counters + 2;
        CL.d_out = p->d_out;
        hipGraph_t graph = nullptr;
        MR_HIP(hipStreamBeginCapture(c->s_req, hipStreamCaptureModeThreadLocal));
        r = launch_cor_groups(c, CL, c->s_req);
        if (!r) r = launch_cor_scores(p->d_out, p->ntasks, p->cor_chunks, p->d_scores, p->d_corws, c->s_req);
        hipError_t ec = hipStreamEndCapture(c->s_req, &graph);
        if (r || ec != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return r ? r : fail(MR_ERUNTIME, std::string("graph capture: ") + hipGetErrorString(ec));
        }
        hipError_t ei = hipGraphInstantiate(&p->cor_exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            p->cor_exec = nullptr;
            return fail(MR_ERUNTIME, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
    }
    if (!p->cor_exec) {
        p->cor_groups = false;
        p->cor_ring = rk;
        p->cor_copies = c->copies_epoch;
        TileQueue q;
        q.head = p->d_counters + 1;
        q.count = p->d_counters;
        q.n_views = p->d_counters + 2;
        q.grid = p->grid;
        hipGraph_t graph = nullptr;
        MR_HIP(hipStreamBeginCapture(c->s_req, hipStreamCaptureModeThreadLocal));
        int r = launch_reset_head(p->d_counters, c->s_req);
        if (!r) r = launch_bp_tiles(c, p->d_tasks, p->d_queue, 0, 0, 0.0f, p->d_out, c->s_req, &q);
        if (!r) r = launch_cor_scores(p->d_out, p->ntasks, p->cor_chunks, p->d_scores, p->d_corws, c->s_req);
        hipError_t ec = hipStreamEndCapture(c->s_req, &graph);
        if (r || ec != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return r ? r : fail(MR_ERUNTIME, std::string("graph capture: ") + hipGetErrorString(ec));
        }
        hipError_t ei = hipGraphInstantiate(&p->cor_exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            p->cor_exec = nullptr;
            return fail(MR_ERUNTIME, std::string("graph instantiate: ") + hipGetErrorString(ei));
        }
    }
    MR_HIP(hipEventRecord(c->ev0, c->s_req));
    MR_HIP(hipGraphLaunch(p->cor_exec, c->s_req));
    MR_HIP(hipEventRecord(c->ev1, c->s_req));
    c->kern_used = p->cor_groups ? 10 : 2;
    MR_HIP(hipMemcpyAsync(scores, p->d_scores, sizeof(double) * p->ntasks, hipMemcpyDeviceToHost, c->s_req));
    int rc = finish_timing(c);
    p->last_ms = c->last_kernel_ms;
    return rc;
}

// Running sums of a plan: allocated on first use; restarted from zero when there are
// none yet or the stored views were reset or replaced since.  Writes the device view
// count (it follows the host one: other ingest paths may have run) and the first view
// not yet summed into the plan's counters [2] and [3].
static int plan_prepare_sums(Plan* p) {
    Ctx* c = p->c;
    const int64_t acc_len = p->out_len * (c->kind == MR_KIND_PARALLEL ? 2 : 1);
    if (!p->d_acc) MR_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_acc), sizeof(float) * acc_len));
    if (p->acc_n < 0 || p->acc_epoch != c->views_epoch || p->acc_n > c->n) {
        MR_HIP(hipMemsetAsync(p->d_acc, 0, sizeof(float) * acc_len, c->s_req));
        p->acc_n = 0;
        p->acc_epoch = c->views_epoch;
    }
    const int nv_begin[2] = {c->n, p->acc_n};
    MR_HIP(hipMemcpyAsync(p->d_counters + 2, nv_begin, sizeof(int) * 2, hipMemcpyHostToDevice, c->s_req));
    MR_HIP(hipStreamSynchronize(c->s_req));
    return MR_OK;
}

int mr_plan_update(mr_plan* p, float* out_host) {
    MR_CHECK(p != nullptr, MR_EVALUE, "null plan");
    Ctx* c = p->c;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(p->geom_epoch == c->geom_epoch, MR_ERUNTIME, "the plan was made for an earlier geometry");
    MR_HIP(hipSetDevice(c->device));
    int rc = prepare_persistent_bp(c, c->n, c->s_req);
    if (!rc) rc = plan_prepare_sums(p);
    if (rc) return rc;
    TileQueue q;
    q.head = p->d_counters + 1;
    q.count = p->d_counters;
    q.n_views = p->d_counters + 2;
    q.grid = p->grid;
    q.v_begin = p->d_counters + 3;
    q.accum = p->d_acc;
    q.accum_half = p->out_len;
    p->acc_n = -1;                     // until the launch is known to have completed
    MR_HIP(hipEventRecord(c->ev0, c->s_req));
    rc = launch_reset_head(p->d_counters, c->s_req);
    if (!rc) rc = launch_bp_tiles(c, p->d_tasks, p->d_queue, 0, 0, 0.0f, p->d_out, c->s_req, &q);
    if (rc) return rc;
    MR_HIP(hipEventRecord(c->ev1, c->s_req));
    rc = plan_copy_out(p, out_host);
    if (rc) return rc;
    rc = finish_timing(c);
    if (rc) return rc;
    p->last_ms = c->last_kernel_ms;
    p->acc_n = c->n;
    return MR_OK;
}

int mr_plan_stream(mr_plan* p, const int64_t* indices, int k, const float* block, float* out_host) {
    MR_CHECK(p != nullptr, MR_EVALUE, "null plan");
    Ctx* c = p->c;
    // a streaming update filters (ingest buffers) and back-projects: both locks
    std::lock_guard<std::mutex> li(c->ingest_mu);
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(p->geom_epoch == c->geom_epoch, MR_ERUNTIME, "the plan was made for an earlier geometry");
    MR_CHECK(k >= 1 && indices && block, MR_EVALUE, "empty or null batch");
    MR_CHECK(c->n + k <= c->n_total, MR_EVALUE, "ring full");
    MR_CHECK(c->gaussian <= 0.0 && std::fabs(c->tilt) < 1e-9, MR_ERUNTIME,
             "streaming plans do not preprocess; push the batch with mr_push_views");
    for (int j = 0; j < k; ++j)
        MR_CHECK(indices[j] >= 0 && indices[j] < c->n_total, MR_EVALUE, "view index out of range");
    MR_HIP(hipSetDevice(c->device));
    const size_t blk = sizeof(float) * static_cast<size_t>(c->R) * k * c->C;
    const size_t meta = (sizeof(FilterView) + sizeof(ViewGeom)) * k;
    if (k > p->kmax) {
        // staging grows: existing graphs captured the old addresses
        for (auto& g : p->graphs) hipGraphExecDestroy(g.second);
        p->graphs.clear();
        if (p->h_block) hipHostFree(p->h_block);
        if (p->h_meta) hipHostFree(p->h_meta);
        if (p->d_block) hipFree(p->d_block);
        if (p->d_meta) hipFree(p->d_meta);
        if (p->d_vmax) hipFree(p->d_vmax);
        p->h_block = nullptr; p->h_meta = nullptr; p->d_block = nullptr; p->d_meta = nullptr;
        p->d_vmax = nullptr;
        p->kmax = 0;
        MR_HIP(hipHostMalloc(reinterpret_cast<void**>(&p->h_block), blk, hipHostMallocDefault));
        MR_HIP(hipHostMalloc(&p->h_meta, meta, hipHostMallocDefault));
        MR_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_block), blk));
        MR_HIP(hipMalloc(&p->d_meta, meta));
        if (c->storage == MR_STORE_F16) {
            // the graphs' own per-view max buffer: the context's d_vmax is regrown (freed)
            // by other entry points, which would leave a graph on freed memory
            MR_HIP(hipMalloc(reinterpret_cast<void**>(&p->d_vmax), sizeof(unsigned) * kF16VmaxPerView * k));
        }
        p->kmax = k;
    }
    int rc = 0;
    FilterView* fv = static_cast<FilterView*>(p->h_meta);
    ViewGeom* gv = reinterpret_cast<ViewGeom*>(static_cast<char*>(p->h_meta) + sizeof(FilterView) * k);
    for (int j = 0; j < k; ++j) {
        fv[j] = c->filt_all[indices[j]];
        fv[j].slot = c->n + j;
        gv[j] = c->geom_all[indices[j]];
    }
    // the ring copy the v2 loop reads here (band or row pair) holds the views stored so
    // far; the graph converts the batch it stores
    rc = prepare_persistent_bp(c, c->n, c->s_req);
    if (rc) return rc;
    const int rk = v2_ring_kind(c);
    if (!p->graphs.empty() && (p->graphs_ring != rk || p->graphs_copies != c->copies_epoch)) {
        // configuration changed or a ring copy was dropped since the capture
        for (auto& g : p->graphs) hipGraphExecDestroy(g.second);
        p->graphs.clear();
    }
    // running sums: the batch adds views [acc_n, n + k) to them
    rc = plan_prepare_sums(p);
    if (rc) return rc;
    hipGraphExec_t exec = nullptr;
    for (auto& g : p->graphs)
        if (g.first == k) exec = g.second;
    if (!exec) {
        TileQueue q;
        q.head = p->d_counters + 1;
        q.count = p->d_counters;
        q.n_views = p->d_counters + 2;
        q.grid = p->grid;
        q.v_begin = p->d_counters + 3;
        q.accum = p->d_acc;
        q.accum_half = p->out_len;
        hipGraph_t graph = nullptr;
        MR_HIP(hipStreamBeginCapture(c->s_req, hipStreamCaptureModeThreadLocal));
        hipError_t e2 = hipMemcpyAsync(p->d_meta, p->h_meta, meta, hipMemcpyHostToDevice, c->s_req);
        int r = e2 != hipSuccess ? MR_ERUNTIME : MR_OK;
        if (!r) r = launch_stream_prep(c, p->d_meta, k, p->d_counters + 2, c->s_req);
        if (!r) r = launch_filter(c, p->d_block, k, 0, c->s_req, static_cast<const FilterView*>(p->d_meta),
                                  p->d_vmax);
        if (!r && rk == 2) r = launch_band_convert_tail(c, p->d_counters + 2, k, c->s_req);
        if (!r && rk == 1) r = launch_pair_convert_tail(c, p->d_counters + 2, k, c->s_req);
        if (!r) r = launch_reset_head(p->d_counters, c->s_req);
        if (!r) r = launch_bp_tiles(c, p->d_tasks, p->d_queue, 0, 0, 0.0f, p->d_out, c->s_req, &q);
        hipError_t ec = hipStreamEndCapture(c->s_req, &graph);
        if (r || ec != hipSuccess) {
            if (graph) hipGraphDestroy(graph);
            return r ? r : fail(MR_ERUNTIME, std::string("graph capture: ") + hipGetErrorString(ec));
        }
        hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (ei != hipSuccess) return fail(MR_ERUNTIME, std::string("graph instantiate: ") + hipGetErrorString(ei));
        if (p->graphs.empty()) {
            p->graphs_ring = rk;
            p->graphs_copies = c->copies_epoch;
        }
        p->graphs.emplace_back(k, exec);
    }
    // the batch goes to the device in chunks through pinned memory: host threads copy
    // the chunks, and each is sent on the request stream as soon as it is complete, so the
    // host copy of the later chunks overlaps the transfer (the previous call has
    // synchronised the stream, so the staging buffer is free).  Everything that waits on
    // the stream (the running-sum counters, a capture) is done before, so the graph launch
    // queues behind the transfer instead of waiting for it; the batch's metadata is sent
    // by the graph
    {
        auto send = [&](size_t o, size_t len) {
            const hipError_t e = hipMemcpyAsync(reinterpret_cast<char*>(p->d_block) + o,
                                                reinterpret_cast<const char*>(p->h_block) + o, len,
                                                hipMemcpyHostToDevice, c->s_req);
            return e == hipSuccess ? MR_OK : fail(MR_ERUNTIME, std::string("batch upload: ") + hipGetErrorString(e));
        };
        if (blk >= (size_t{1} << 20)) {
            if (!p->copy_pool) p->copy_pool.reset(new HostCopyPool(3));
            rc = p->copy_pool->copy(p->h_block, block, blk, 8, send);
        } else {
            std::memcpy(p->h_block, block, blk);
            rc = send(0, blk);
        }
        if (rc) return rc;
    }
    p->acc_n = -1;                     // until the replay is known to have completed
    MR_HIP(hipEventRecord(c->ev0, c->s_req));
    MR_HIP(hipGraphLaunch(exec, c->s_req));
    MR_HIP(hipEventRecord(c->ev1, c->s_req));
    rc = plan_copy_out(p, out_host);
    if (rc) return rc;
    rc = finish_timing(c);
    if (rc) return rc;
    p->last_ms = c->last_kernel_ms;
    c->n += k;
    if (rk == 2) c->bring_n = std::max(c->bring_n, c->n);
    if (rk == 1) c->pring_n = std::max(c->pring_n, c->n);
    p->acc_n = c->n;
    return MR_OK;
}

int mr_plan_output(mr_plan* p, float** out_device, int64_t* out_len) {
    MR_CHECK(p != nullptr, MR_EVALUE, "null plan");
    if (out_device) *out_device = p->d_out;
    if (out_len) *out_len = p->out_len;
    return MR_OK;
}

int mr_plan_last_ms(mr_plan* p, double* ms) {
    MR_CHECK(p && ms, MR_EVALUE, "null argument");
    *ms = p->last_ms;
    return MR_OK;
}

int mr_reset_views(mr_ctx* c) {
    int rc = check_ctx(c);
    if (rc) return rc;
    // geometry and view-count changes exclude ingest and requests (ingest lock first)
    std::lock_guard<std::mutex> li(c->ingest_mu);
    std::lock_guard<std::mutex> lk(c->mu);
    c->n = 0;
    c->pring_n = 0;
    c->bring_n = 0;
    ++c->views_epoch;
    return MR_OK;
}

int mr_cor_scores(mr_ctx* c, const mr_task* tasks, int ntasks, int n_views, double* scores) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(n_views >= 0 && n_views <= c->n, MR_EVALUE, "n_views outside the stored range");
    if (ntasks <= 0) return MR_OK;
    MR_CHECK(tasks && scores, MR_EVALUE, "null tasks or scores");
    // the images go one after another into a device buffer, whatever their sizes
    std::vector<mr_task> T(tasks, tasks + ntasks);
    std::vector<int64_t> offs(ntasks), npix(ntasks);
    int64_t total = 0;
    for (int i = 0; i < ntasks; ++i) {
        MR_CHECK(T[i].nq == 1 && T[i].np >= 0 && T[i].ns >= 0, MR_EVALUE, "score tasks must be single images");
        npix[i] = static_cast<int64_t>(T[i].np) * T[i].ns;
        offs[i] = total;
        T[i].out_offset = total;
        T[i].stride_q = npix[i];
        T[i].stride_p = T[i].ns;
        T[i].stride_s = 1;
        total += npix[i];
    }
    MR_HIP(hipSetDevice(c->device));
    rc = ensure(reinterpret_cast<void**>(&c->d_out), &c->out_bytes, sizeof(float) * total);
    if (rc) return rc;
    rc = ensure(reinterpret_cast<void**>(&c->d_misc), &c->misc_bytes, sizeof(double) * ntasks);
    if (rc) return rc;
    // image descriptors first: their copy waits for the stream, which is idle here
    rc = ensure(&c->d_corws, &c->corws_bytes, cor_workspace_bytes(ntasks));
    if (rc) return rc;
    std::vector<CorImage> imgs;
    const int chunks = cor_images(offs.data(), npix.data(), ntasks, imgs);
    rc = write_cor_images(imgs, c->d_corws, c->s_req);
    if (rc) return rc;
    MR_HIP(hipMemsetAsync(c->d_out, 0, sizeof(float) * total, c->s_req));
    rc = run_tasks(c, T.data(), ntasks, n_views, c->d_out);
    if (rc) return rc;
    rc = launch_cor_scores(c->d_out, ntasks, chunks, c->d_misc, c->d_corws, c->s_req);
    if (rc) return rc;
    MR_HIP(hipMemcpyAsync(scores, c->d_misc, sizeof(double) * ntasks, hipMemcpyDeviceToHost, c->s_req));
    return finish_timing(c);
}

int mr_filter_rows(mr_ctx* c, const float* in, int rows, int n, double tau, int window, float* out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> li(c->ingest_mu);
    MR_CHECK(rows >= 0 && n >= 1, MR_EVALUE, "bad row shape");
    MR_CHECK(valid_window(window), MR_EVALUE, "unknown filter window " + std::to_string(window));
    MR_CHECK(tau > 0, MR_EVALUE, "tau must be > 0");
    if (rows == 0) return MR_OK;
    MR_CHECK(in && out, MR_EVALUE, "null input or output");
    rc = check_filter_cols(n);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    const int N = fft_length(n);
    float* d_H = nullptr;
    float2* d_tw = nullptr;
    rc = make_tables(N, window, &d_H, &d_tw);
    if (rc) return rc;
    const size_t bytes = sizeof(float) * static_cast<size_t>(rows) * n;
    const size_t fs_bytes = fft_scratch_bytes(N, (rows + 1) / 2);
    float *d_in = nullptr, *d_o = nullptr;
    float2* d_fs = nullptr;
    hipError_t e1 = hipMalloc(reinterpret_cast<void**>(&d_in), bytes);
    hipError_t e2 = hipMalloc(reinterpret_cast<void**>(&d_o), bytes);
    hipError_t e3 = fs_bytes ? hipMalloc(reinterpret_cast<void**>(&d_fs), fs_bytes) : hipSuccess;
    if (e1 == hipSuccess && e2 == hipSuccess && e3 == hipSuccess) {
        // an ingest-side operation (the context's filter-view buffer): the ingest stream
        hipError_t e = hipMemcpyAsync(d_in, in, bytes, hipMemcpyHostToDevice, c->s_ingest);
        if (e == hipSuccess)
            rc = launch_filter_rows_plain(c, d_in, rows, n, static_cast<float>(1.0 / (N * tau)), d_H,
                                          d_tw, N, d_o, d_fs, fs_bytes, c->s_ingest);
        if (!rc) {
            if (e == hipSuccess)
                e = hipMemcpyAsync(out, d_o, bytes, hipMemcpyDeviceToHost, c->s_ingest);
            if (e == hipSuccess) e = hipStreamSynchronize(c->s_ingest);
            if (e != hipSuccess) rc = fail(MR_ERUNTIME, std::string("filter_rows: ") + hipGetErrorString(e));
        }
    } else {
        rc = fail(MR_ERUNTIME, "filter_rows: device allocation failed");
    }
    if (d_in) hipFree(d_in);
    if (d_o) hipFree(d_o);
    if (d_fs) hipFree(d_fs);
    hipFree(d_H);
    hipFree(d_tw);
    return rc;
}

int mr_read_views(mr_ctx* c, int first, int count, float* out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(c->kind != MR_KIND_NONE, MR_ERUNTIME, "no geometry set");
    MR_CHECK(first >= 0 && count >= 0 && first + count <= c->n, MR_EVALUE, "view range outside the stored views");
    if (count == 0) return MR_OK;
    MR_CHECK(out != nullptr, MR_EVALUE, "null output");
    MR_HIP(hipSetDevice(c->device));
    const size_t bytes = sizeof(float) * static_cast<size_t>(count) * c->R * c->C;
    rc = ensure(reinterpret_cast<void**>(&c->d_out), &c->out_bytes, bytes);
    if (rc) return rc;
    rc = launch_read_views(c, first, count, c->d_out, c->s_req);
    if (rc) return rc;
    MR_HIP(hipMemcpyAsync(out, c->d_out, bytes, hipMemcpyDeviceToHost, c->s_req));
    MR_HIP(hipStreamSynchronize(c->s_req));
    return MR_OK;
}

static int phantom_common(mr_ctx* c, const double* ell, int K, int kind, const double* geo, int ngeo,
                          int A, int rows, int cols, double pu, double pv, double far, float* out) {
    int rc = check_ctx(c);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    MR_CHECK(K >= 0 && A >= 0 && rows >= 0 && cols >= 0, MR_EVALUE, "negative phantom size");
    if (A == 0 || rows == 0 || cols == 0) return MR_OK;
    MR_CHECK(out && geo && (K == 0 || ell), MR_EVALUE, "null phantom argument");
    MR_HIP(hipSetDevice(c->device));
    double *d_ell = nullptr, *d_geo = nullptr;
    MR_HIP(hipMalloc(reinterpret_cast<void**>(&d_ell), sizeof(double) * (13 * K + 1)));
    MR_HIP(hipMalloc(reinterpret_cast<void**>(&d_geo), sizeof(double) * ngeo));
    if (K) MR_HIP(hipMemcpy(d_ell, ell, sizeof(double) * 13 * K, hipMemcpyHostToDevice));
    MR_HIP(hipMemcpy(d_geo, geo, sizeof(double) * ngeo, hipMemcpyHostToDevice));
    rc = launch_phantom(d_ell, K, kind, d_geo, A, rows, cols, pu, pv, far, out, c->s_req);
    hipError_t e = hipStreamSynchronize(c->s_req);
    hipFree(d_ell);
    hipFree(d_geo);
    if (rc) return rc;
    if (e != hipSuccess) return fail(MR_ERUNTIME, std::string("phantom: ") + hipGetErrorString(e));
    return MR_OK;
}

int mr_phantom_cone(mr_ctx* c, const double* ell, int K, const double* vectors, int A, int rows,
                    int cols, float* out_device) {
    return phantom_common(c, ell, K, MR_KIND_CONE, vectors, 12 * A, A, rows, cols, 0, 0, 0, out_device);
}

int mr_phantom_parallel(mr_ctx* c, const double* ell, int K, const double* angles, int A, int rows,
                        int cols, double pixel_u, double pixel_v, double far, float* out_device) {
    return phantom_common(c, ell, K, MR_KIND_PARALLEL, angles, A, A, rows, cols, pixel_u, pixel_v, far,
                          out_device);
}

int mr_device_alloc(mr_ctx* c, int64_t bytes, void** ptr) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(ptr && bytes >= 0, MR_EVALUE, "bad allocation request");
    MR_HIP(hipSetDevice(c->device));
    MR_HIP(hipMalloc(ptr, bytes > 0 ? bytes : 1));
    return MR_OK;
}

int mr_device_free(mr_ctx* c, void* ptr) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    if (ptr) MR_HIP(hipFree(ptr));
    return MR_OK;
}

int mr_memcpy(mr_ctx* c, void* dst, const void* src, int64_t bytes, int kind) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(bytes >= 0 && kind >= 0 && kind <= 2, MR_EVALUE, "bad copy request");
    if (bytes == 0) return MR_OK;
    MR_HIP(hipSetDevice(c->device));
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                                      : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    MR_HIP(hipMemcpy(dst, src, bytes, k));
    return MR_OK;
}

int mr_synchronize(mr_ctx* c) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_HIP(hipSetDevice(c->device));
    MR_HIP(hipDeviceSynchronize());
    return MR_OK;
}

int mr_last_kernel_ms(mr_ctx* c, double* ms) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(ms != nullptr, MR_EVALUE, "null output");
    *ms = c->last_kernel_ms;
    return MR_OK;
}

int mr_last_tile_kernel_ms(mr_ctx* c, double* ms) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(ms != nullptr, MR_EVALUE, "null output");
    *ms = c->last_tile_ms;
    return MR_OK;
}

int mr_last_tile_kernel_launches(mr_ctx* c, int* n) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(n != nullptr, MR_EVALUE, "null output");
    *n = c->last_tile_launches;
    return MR_OK;
}

int mr_last_bp_kernel(mr_ctx* c, int* kernel) {
    int rc = check_ctx(c);
    if (rc) return rc;
    MR_CHECK(kernel != nullptr, MR_EVALUE, "null output");
    *kernel = c->last_kern;
    return MR_OK;
}

}  // extern "C"
