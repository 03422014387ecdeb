// Internal declarations shared by the libminirecast_hip translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/minirecast_hip.h"
#include "mr_hostpool.h"
#include "mr_tiles.h"

namespace mr {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define MR_HIP(call)                                                              \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return ::mr::fail(MR_ERUNTIME, std::string(#call) + ": " +            \
                                               hipGetErrorString(e_));            \
    } while (0)

#define MR_CHECK(cond, code, msg)                                                 \
    do {                                                                          \
        if (!(cond)) return ::mr::fail(code, msg);                                \
    } while (0)

// ---------------------------------------------------------------- geometry
// Projective form of one stored view (float64).  For a point x (mm):
//   uh = P[0].[x,1], vh = P[1].[x,1], wh = P[2].[x,1]
//   col = uh / wh + (C-1)/2,  row = vh / wh + (R-1)/2
// Cone beam: rows are divided by SOD so the FDK distance weight (SOD/L)^2 is 1/wh^2.
// Parallel beam: wh == 1 and no distance weight.
// A rotation-centre shift s adds s*wh to uh and s*uvv*wh to vh (the cone detector
// moved to D - s*u, recon.py:420; the parallel column shift, recon.py:181).
struct ViewGeom {
    double P[3][4];
    double uvv;
    double pad[3];
};
static_assert(sizeof(ViewGeom) == 128, "ViewGeom layout");

// FDK cosine weight / pitch of one arriving view (recon.py:93-102) in float32:
//   w(r, c) = sdd / sqrt(sdd^2 + (a0 + c*da)^2 + (b0 + r*db)^2)
// `gain` = 1 / (N * tau_iso): the inverse-FFT normalisation and the ramp
// response's 1/tau dependence (H(tau) = H(1) / tau).
struct FilterView {
    float a0, da, b0, db;
    float sdd, sdd2, gain;
    int32_t slot;
    int32_t cone;
    int32_t pad[3];
};
static_assert(sizeof(FilterView) == 48, "FilterView layout");
// f16 ingest: per batch view, the scale bound of sampled raw rows in kF16SampleParts parts and
// the max |q| of the filtered view gathered over kF16QmaxSlots addresses in separate cache
// lines (entries of the max buffer: kF16VmaxPerView per view)
constexpr int kF16SampleParts = 16;
constexpr int kF16QmaxSlots = 16;
constexpr int kF16QmaxStride = 32;     // entries between slots: each slot in its own 128-byte line
constexpr int kF16VmaxPerView = kF16SampleParts + kF16QmaxSlots * kF16QmaxStride;

// Device work queue of a persistent back-projection launch.
struct TileQueue {
    int* head = nullptr;             // atomic tile counter (null: one tile per workgroup)
    const int* count = nullptr;      // number of queued tiles
    const int* n_views = nullptr;    // view count read on the device (null: argument)
    int grid = 0;                    // persistent grid size
    // running sums (persistent v2 launches): views [*v_begin, n) are added to the
    // unscaled FP32 sums in accum (cone: [out_len]; parallel: two rows' sums,
    // [2][out_len] at accum_half), which are written back; out gets the scaled result
    const int* v_begin = nullptr;
    float* accum = nullptr;
    int64_t accum_half = 0;
};

// Bands of the band ring (kernel 7): padded rows 0..R+1 can hold a point's upper tap,
// band b covers padded rows 3b..3b+3.
constexpr int band_count(int R) { return (R + 1) / 3 + 1; }

// BP workgroup size (the tile shape BP_TP x BP_TS is in mr_tiles.h).
constexpr int BP_THREADS = 256;
constexpr int BP_VPASS = 64;     // views per LDS record pass
constexpr int BP_REC = 16;       // floats per (tile, view) record

// ---------------------------------------------------------------- context
// Raw-count correction set by mr_set_raw_correction.
struct RawCorr {
    int set = 0;
    int H = 0, W = 0, sample = 1, flip = 0, row0 = 0, col0 = 0, op = 0;
    float* d_dark = nullptr;         // [R][C]
    float* d_scale = nullptr;        // [R][C]
};

// Locks: ingest_mu serialises ingest (staging buffers d_stage, d_pre, d_raw, d_fv,
// d_vmax, d_fs, the raw correction, the preprocessing parameters and s_ingest); mu guards
// the stored-view count n, the ring copies, the request buffers and s_req.  An ingest
// writes only ring and d_geom slots >= n, which no request reads, and takes mu just to
// commit n; a request runs while a batch is being filtered.  Changes of the geometry or
// of n take both, ingest_mu first.
struct Ctx {
    int device = 0;
    std::mutex ingest_mu;
    std::mutex mu;
    hipStream_t s_ingest = nullptr, s_req = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_kernel_ms = 0.0;
    hipEvent_t evk0 = nullptr, evk1 = nullptr;   // around the tile kernel alone (v4 / v5)
    bool tile_timed = false;
    int tile_launches = 0;           // tile-kernel dispatches of the request in flight
    int last_tile_launches = 0;      // ... of the last finished request
    int kern_used = 0;               // back-projection kernel of the request in flight
                                     // (2, 4-7; 8, 9: z columns; 10: COR shift groups)
    int last_kern = 0;               // ... of the last finished request
    double last_tile_ms = 0.0;
    int bp_variant = 4;              // back-projection inner loop (MINIRECAST_BP_VARIANT)
    int bp_ppt = 2;                  // v4/v5 points per thread (MINIRECAST_BP_PPT)
    int xcd_map = 2;                 // 1: XCD-contiguous tiles, 2: also sorted by height
                                     // (MINIRECAST_XCD_MAP)
    int bp_layout = 0;               // v4 lane layout: 0 = 2 x 32, 1 = 4 x 16 (MINIRECAST_BP_LAYOUT)
    int bp_depth = 2;                // v4 views in flight, 2 or 3 (MINIRECAST_BP_DEPTH)
    bool bp_explicit = false;        // configuration chosen by the caller or the environment
    int bp_vchunk = 0;               // kernels 4/6 (cone): views per launch, 0 = all (MINIRECAST_BP_VCHUNK)
    bool vchunk_explicit = false;    // views per launch set by the environment
    bool vchunk_any = false;         // split grids of any size (MINIRECAST_BP_VCHUNK_ANY, tests)
    size_t rec_budget = size_t{512} << 20;   // kernel-4 records buffer (MINIRECAST_BP_RECORD_KB)
    bool req_slices = false;         // the request in flight has single-plane tasks only
    int n_cus = 256;                 // compute units of the device
    int bp_sync = 0;                 // kernels 4/6: waves of a workgroup in lockstep per view
                                     // (MINIRECAST_BP_SYNC)
    RawCorr raw;                     // raw-count correction of mr_push_raw
    void* pring = nullptr;           // row-pair copy of the ring (kernel 6)
    size_t pring_bytes = 0;
    int pring_n = 0;                 // views converted into pring
    void* bring = nullptr;           // band copy of a 16-bit ring (kernel 7)
    size_t bring_bytes = 0;
    int bring_n = 0;                 // views converted into bring
    // ring copies are caches (ensure_pair_ring / ensure_band_ring): built only with
    // copy_reserve bytes of device memory left after them and within copy_budget
    // (-1: none) for both together; a copy dropped to make room bumps copies_epoch
    size_t copy_reserve = size_t{1} << 30;   // MINIRECAST_COPY_RESERVE_MB
    int64_t copy_budget = -1;
    uint64_t copies_epoch = 0;
    bool band_no_room = false, pair_no_room = false;   // the last attempt found no room
    uint16_t* d_raw = nullptr;       // staged raw frames
    size_t raw_bytes = 0;
    void* d_vol = nullptr;           // staged volume for the statistics entry points
    size_t vol_bytes = 0;
    void* d_stat = nullptr;          // histograms, prefixes, edges
    size_t stat_bytes = 0;
    float* d_recs = nullptr;         // v4 per-(tile, view) records
    size_t recs_bytes = 0;
    void* d_adj = nullptr;           // adjoint gather: masked weights + detector boxes
    size_t adj_bytes = 0;
    void* d_part = nullptr;          // per-part sums of view-split launches (shift derivative)
    size_t part_bytes = 0;

    int kind = MR_KIND_NONE;
    // cone geometry whose every view has P[0][2] = P[2][2] = 0 exactly: the detector column
    // and depth of a point do not depend on z (volume requests can run the z-column kernel)
    bool zsep = false;
    int n_total = 0, R = 0, C = 0;
    int window = MR_WINDOW_SHEPP_LOGAN, storage = MR_STORE_F32;
    int N = 0;                       // FFT length
    int pitch = 0, Rp = 0;           // padded row pitch (elements) and padded rows
    int64_t view_stride = 0;         // elements per stored view
    int n = 0;                       // views stored
    // bumped when the geometry is set (geom_epoch, views_epoch) or the stored views are
    // forgotten (views_epoch): plans check them before reusing graphs or running sums
    uint64_t geom_epoch = 0, views_epoch = 0;
    std::vector<TileCache> tile_cache;   // most recent first, at most kTileCacheEntries
    int v2_ring = 0;                 // ring copy of the v2 launch in progress (v2_ring_kind)

    // host copies
    std::vector<ViewGeom> geom_all;   // per geometry index
    std::vector<FilterView> filt_all; // per geometry index (slot filled at push)
    std::vector<uint8_t> received;
    double pixel_u = 1.0, pixel_v = 1.0;
    double gaussian = 0.0, tilt = 0.0;  // preprocessing applied at ingest

    // device buffers
    void* ring = nullptr;            // [n_total][Rp][pitch] storage elements
    ViewGeom* d_geom = nullptr;      // [n_total] per slot
    float* d_vscale = nullptr;       // [n_total] per slot
    float* d_H = nullptr;            // [N] ramp response for tau = 1 (windowed)
    double h_l1 = 0.0;               // sum_m |h(m)| of that response's unnormalised inverse
                                     // DFT (the f16 scale bound of the filter)
    float2* d_tw = nullptr;          // [N] twiddles exp(-2 pi i k / N), then, when N >
                                     // 2^kFilterLdsLogN, [N / 16] for N / 16
    float2* d_fs = nullptr;          // four-step filter scratch (N > 2^kFilterLdsLogN)
    size_t fs_bytes = 0;
    float* d_stage = nullptr;        // raw block staging
    size_t stage_bytes = 0;
    float* d_pre = nullptr;          // preprocessing ping-pong buffer
    size_t pre_bytes = 0;
    float* d_fstage = nullptr;       // f32 staging of a gradient (adjoint with a host output)
    size_t fstage_bytes = 0;
    unsigned* d_vmax = nullptr;      // per-batch per-view max |x| or |q| (float bits) [kF16VmaxPerView k]
    size_t vmax_len = 0;
    FilterView* d_fv = nullptr;
    size_t fv_len = 0;

    // request scratch
    void* d_req = nullptr;           // tasks + tiles
    size_t req_bytes = 0;
    void* h_req = nullptr;           // pinned host staging
    size_t h_req_bytes = 0;
    float* d_out = nullptr;          // device output staging
    size_t out_bytes = 0;
    double* d_pts = nullptr;
    size_t pts_bytes = 0;
    double* d_misc = nullptr;
    size_t misc_bytes = 0;
    void* d_corws = nullptr;         // COR scoring workspace
    size_t corws_bytes = 0;
    void* d_cor = nullptr;           // COR shift groups: detector rows, coefficients, tables
    size_t cor_bytes = 0;
};

// A compiled request: tasks and their device work queue, re-run on demand or after
// every ingested batch as one captured HIP graph per batch size.
struct Plan {
    Ctx* c = nullptr;
    int ntasks = 0;
    int64_t out_len = 0;
    int64_t total_candidates = 0;
    mr_task* d_tasks = nullptr;
    int64_t* d_toff = nullptr;
    Tile* d_queue = nullptr;
    int* d_counters = nullptr;       // [0] queue length, [1] head, [2] stored views
    float* d_out = nullptr;          // plan output
    int grid = 0;                    // persistent grid
    // streaming update staging (pinned host, fixed device addresses for graph replay)
    int kmax = 0;
    float* h_block = nullptr;        // [R][kmax][C]
    float* d_block = nullptr;
    void* h_meta = nullptr;          // FilterView[kmax] + ViewGeom[kmax] + int slots[kmax]
    void* d_meta = nullptr;
    size_t meta_bytes = 0;
    unsigned* d_vmax = nullptr;      // f16 storage: per-view max |x|, max |q| [kF16VmaxPerView kmax]
    std::vector<std::pair<int, hipGraphExec_t>> graphs;   // per batch size
    double last_ms = 0.0;
    uint64_t geom_epoch = 0;         // context geometry the plan was made for
    // streaming: unscaled FP32 sums over views [0, acc_n) of the context's ring epoch
    // acc_epoch (acc_n < 0: none yet), so that a batch adds only its new views
    float* d_acc = nullptr;
    int acc_n = -1;
    uint64_t acc_epoch = 0;
    // ring copy the captured graphs read (v2_ring_kind) and the context's copies_epoch
    // at their capture: recaptured when either changes
    int graphs_ring = 0, cor_ring = 0;
    uint64_t graphs_copies = 0, cor_copies = 0;
    int cor_chunks = 0;           // grid width of the scoring kernels (largest chunk count)
    // rotation-centre scoring of the plan's images (one graph: persistent BP + scores)
    std::vector<mr_task> h_tasks;    // host copy of the tasks
    double* d_scores = nullptr;
    void* d_corws = nullptr;         // scoring workspace of the captured graph
    hipGraphExec_t cor_exec = nullptr;
    void* d_cor = nullptr;           // COR shift groups of the captured graph (cor_groups)
    bool cor_groups = false;         // the captured graph runs the shift-group kernel
    std::unique_ptr<HostCopyPool> copy_pool;   // staging copies of streamed batches
};

int ensure(void** p, size_t* have, size_t need);   // device buffer growth
int ensure_host(void** p, size_t* have, size_t need);

// ---------------------------------------------------------------- kernels (launchers)
// fv, vmax: the batch's filter views and per-view max buffer [k] (f16 storage); null
// selects the context's own (d_fv, d_vmax).  A captured graph passes buffers it owns,
// since the context's may be reallocated.
int launch_filter(Ctx* c, const float* d_block, int k, int n0, hipStream_t s,
                  const FilterView* fv = nullptr, unsigned* vmax = nullptr);
int launch_store(Ctx* c, const float* d_q, int k, hipStream_t s);
int launch_preprocess(Ctx* c, float* d_a, float* d_b, int R, int k, int C, double gaussian,
                      double tilt, float** d_res, hipStream_t s);
// d_fs, fs_bytes: the four-step scratch (fft_scratch_bytes) when N > 2^14, else unused
int launch_filter_rows_plain(Ctx* c, const float* d_in, int rows, int n, float gain_tau,
                             const float* d_H, const float2* d_tw, int N, float* d_out,
                             float2* d_fs, size_t fs_bytes, hipStream_t s);
// FFT lengths: one workgroup's LDS holds 2^kFilterLdsLogN points; longer rows, up to
// 2^kFilterMaxLogN, take the four-step path through a device scratch of
// fft_scratch_bytes(N, pairs) bytes (0 when not needed)
constexpr int kFilterLdsLogN = 14;
constexpr int kFilterMaxLogN = 18;
size_t fft_scratch_bytes(int N, int pairs);
int launch_bp_tiles(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles,
                    int n_views, float scale, float* d_out, hipStream_t s,
                    const TileQueue* q = nullptr);
// raw frames [k][H][W] uint16 -> corrected [R][k][C] float32 (c->raw)
int launch_raw_correct(Ctx* c, const uint16_t* d_frames, int k, float* d_out, hipStream_t s);
// volume statistics (mr_volume.hip): per-pass radix histograms of up to 8 target ranks
// and the clipped uniform-bin histogram
int launch_radix_hist(const float* d_v, int64_t n, const uint32_t* d_prefix, int ntarget,
                      int shift, unsigned long long* d_hist, hipStream_t s);
int launch_clip_hist(const float* d_v, int64_t n, double lo, double hi, int bins,
                     const double* d_edges, unsigned long long* d_hist, hipStream_t s);
// zeros every point of the given tiles (tiles culled by the box test)
int launch_zero_tiles(const mr_task* d_tasks, const Tile* d_tiles, int ntiles, float* d_out,
                      hipStream_t s);
int launch_enumerate_tiles(const mr_task* d_tasks, const int64_t* d_toff, int ntasks,
                           int64_t total, Tile* d_queue, int* d_counters, hipStream_t s);
int launch_reset_head(int* d_counters, hipStream_t s);
// persistent (v2) launches of a kernel-7 context read the band copy: it must hold the
// views a launch reads (prepare before launching or capturing; in a streaming graph the
// tail conversion brings in the batch just stored)
int64_t band_view_stride(const Ctx* c);
int64_t pair_view_stride(const Ctx* c);
int v2_ring_kind(const Ctx* c);      // 0 the ring, 1 the row-pair copy, 2 the band copy
int prepare_persistent_bp(Ctx* c, int n_views, hipStream_t s);
int launch_band_convert_tail(Ctx* c, const int* d_n, int k, hipStream_t s);
int launch_pair_convert_tail(Ctx* c, const int* d_n, int k, hipStream_t s);
int ensure_pair_ring(Ctx* c, int n_views, hipStream_t s);
size_t copies_held(const Ctx* c);    // device bytes of the resident ring copies
void trim_ring_copies(Ctx* c, hipStream_t s);   // drops copies beyond copy_budget
int launch_stream_prep(Ctx* c, const void* d_meta, int k, int* d_nviews, hipStream_t s);
int persistent_grid(Ctx* c);
int launch_bp_points(Ctx* c, const double* d_pts, int64_t M, double cor_shift, int n_views,
                     float scale, float* d_out, hipStream_t s);
int launch_read_views(Ctx* c, int first, int count, float* d_out, hipStream_t s);
int launch_bp_adjoint(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles,
                      int n_views, float scale, const float* d_g, float* d_gq, hipStream_t s);
int launch_bp_dshift(Ctx* c, const mr_task* d_tasks, const Tile* d_tiles, int ntiles, int n_views,
                     float scale, float* d_out, hipStream_t s);
// the z-column kernel of a volume request: zcol_form picks its form (and builds the ring
// copy it reads), build_zcol_tiles(tasks, n, form == ZCOL_STAGED ? ZQL : ZQ) its tiles;
// n_views > 0
enum { ZCOL_RING = 0, ZCOL_STAGED = 1 };
// Shift groups of a rotation-centre search (parallel beam, cor_shift_groups): detector
// rows per set and view, then the group kernel.  With d_nviews (graph replays) the view
// count is read on the device and rows are prepared for up to nvt views.
struct CorLaunch {
    int q = 0;                       // shifts s0 + k / q
    int nsets = 0, ntiles = 0;
    const mr_task* d_tasks = nullptr;
    const CorRowSet* d_sets = nullptr;
    const CorGroup* d_groups = nullptr;
    const CorTile* d_tiles = nullptr;
    float* d_rows = nullptr;         // [nsets][nvt][lw]
    double* d_coef = nullptr;        // [nsets][nvt][4]
    int lw = 0, nvt = 0;             // floats per row (>= C + 2 kCorPad), views per set
    const int* d_nviews = nullptr;
    int n_views = 0;
    float* d_out = nullptr;
};
int launch_cor_groups(Ctx* c, const CorLaunch& L, hipStream_t s);
// floats per detector row of a COR shift set
inline int cor_row_floats(int C) { return (C + 2 * kCorPad + 3) / 4 * 4; }

int zcol_form(Ctx* c, int n_views, hipStream_t s, int* form);
int launch_bp_zcol(Ctx* c, int form, const mr_task* d_tasks, int ntasks, const Tile* d_tiles,
                   int ntiles, int n_views, float scale, float* d_out, hipStream_t s);
int launch_filter_adjoint(Ctx* c, const float* d_gq, int k, float* d_graw, hipStream_t s);
// COR entropy scores of nimg images, image i = npix[i] floats from d_img + offset[i].
// d_ws: a device workspace of cor_workspace_bytes(nimg) bytes (owned by the caller: a
// captured graph keeps it) that also holds the image descriptors: cor_images forms them
// (and returns the largest chunk count, the launch's grid width), write_cor_images copies
// them in (outside any capture; they stay valid until the workspace is rewritten).
struct CorImage {
    int64_t offset, npix;
    int nchunk, pad;
};
size_t cor_workspace_bytes(int nimg);
int cor_images(const int64_t* offsets, const int64_t* npix, int nimg, std::vector<CorImage>& out);
int write_cor_images(const std::vector<CorImage>& imgs, void* d_ws, hipStream_t s);
int launch_cor_scores(const float* d_img, int nimg, int max_chunks, double* d_scores, void* d_ws,
                      hipStream_t s);
int launch_phantom(const double* d_ell, int K, int kind, const double* d_geo, int A, int R,
                   int C, double pu, double pv, double far, float* out, hipStream_t s);

// the static scratch of the entry points that allocate nothing (mr_noalloc.hip): takes
// the device address of this GPU's copy, once per device, with that device current
int static_scratch_init(int device);

// host float64 helpers
int ramp_response_host(int n, double tau, int window, std::vector<double>& H);

}  // namespace mr
