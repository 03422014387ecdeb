/*
 * minirecast_hip.h -- C ABI of libminirecast_hip.so, the MI355X (gfx950) streaming
 * FDK/FBP engine behind the minirecast.recon / minirecast.preprocess Python API.
 *
 * Plain C types only: pointers, sizes, doubles.  Every function returns MR_OK (0)
 * or a negative MR_E* code; mr_last_error() then holds a thread-local message.
 * Host arrays are caller-owned, C-contiguous and only read during the call.
 * Pointers documented as "device" must come from the same GPU as the context.
 *
 * Each entry point cites the reference function it replaces (paths relative to
 * the minirecast reference package).
 */
#ifndef MINIRECAST_HIP_H
#define MINIRECAST_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes: the Python layer maps them to the reference's exception types */
#define MR_OK            0
#define MR_EVALUE       -1   /* ValueError: bad shape, count, window, index        */
#define MR_ERUNTIME     -2   /* RuntimeError: HIP/device failure, no data          */
#define MR_ETYPE        -3   /* TypeError: wrong geometry kind for the call        */

/* ---- filter windows (recon.py:45-54) */
#define MR_WINDOW_RAMLAK        0   /* "ram-lak" / "none" / None */
#define MR_WINDOW_SHEPP_LOGAN   1
#define MR_WINDOW_HANN          2
#define MR_WINDOW_COSINE        3

/* ---- storage of the filtered-projection ring (recon.py:320 stores float32) */
#define MR_STORE_F32   0
#define MR_STORE_F16   1   /* IEEE half, one power-of-two scale per view */
#define MR_STORE_BF16  2   /* bfloat16, round-to-nearest-even           */

/* ---- geometry kinds */
#define MR_KIND_NONE      0
#define MR_KIND_CONE      1
#define MR_KIND_PARALLEL  2

typedef struct mr_ctx mr_ctx;

/* Thread-local message of the last failing call on this thread. */
const char* mr_last_error(void);

/* ABI version (bumped on any signature change). */
int mr_abi_version(void);

/* Number of visible HIP devices (0 when none); never fails on a machine without GPUs. */
int mr_device_count(int* n);

/* Windowed discrete Ram-Lak response: replaces recon.py:34-55 ramp_response.
 * Writes N = 2^ceil(log2(2n)) to *N and, when H is non-NULL, H[0..N-1] (float64),
 * computed in float64 on the host.  MR_EVALUE for an unknown window or n < 1. */
int mr_ramp_response(int n, double tau, int window, double* H, int* N);

/* A reconstruction context owns the device ring, per-view geometry and streams of
 * one SliceReconstructor (recon.py:298-343).  device = HIP ordinal. */
int mr_ctx_create(int device, mr_ctx** out);
int mr_ctx_destroy(mr_ctx* ctx);

/* Cone-beam geometry (recon.py:327-330 via cone_constants, recon.py:70-82).
 * vectors: [n_total][12] float64 ASTRA cone_vec rows (src, det centre, u, v in mm).
 * Allocates and zeroes the ring for n_total views of rows x cols. */
int mr_set_geometry_cone(mr_ctx* ctx, int n_total, int rows, int cols,
                         const double* vectors, int window, int storage);

/* Parallel-beam geometry (recon.py:322-326): angles [n_total] in radians,
 * pixel sizes in mm. */
int mr_set_geometry_parallel(mr_ctx* ctx, int n_total, int rows, int cols,
                             const double* angles, double pixel_u, double pixel_v,
                             int window, int storage);

/* Ingest: replaces the filter + store steps of recon.py:346-376 add_projections
 * (fdk_filter_block recon.py:85-104 / filter_rows recon.py:58-64).
 * indices[k]: geometry indices of the views (the caller has already dropped the
 * ones received before, recon.py:353-356).  block: [rows][k][cols] float32, a host
 * pointer (block_on_device = 0) or a device pointer (1).  The filtered views are
 * appended in arrival order at slots n..n+k-1; *n_out receives the new count. */
int mr_push_views(mr_ctx* ctx, const int64_t* indices, int k, const float* block,
                  int block_on_device, int* n_out);

/* Appends already-filtered views q [k][rows][cols] float32 (host or device pointer)
 * at slots n..n+k-1 with the geometry of `indices`, quantising for F16/BF16 storage:
 * the store step of recon.py:363-375 without filtering.  Used by the reference's
 * module-level bp_cone_* / bp_parallel_* signatures (recon.py:129-219), which take
 * filtered views, and by the adjoint. */
int mr_store_views(mr_ctx* ctx, const int64_t* indices, int k, const float* q, int q_on_device,
                   int* n_out);

/* Raw-count correction for mr_push_raw, replacing the host-side correction of the
 * loaders (flexray.py:174-184 corrected(), with _read_binned flexray.py:111-116;
 * hdf5.py:77-78, with _bin hdf5.py:24-31).  Raw frames are [frame_rows][frame_cols]
 * uint16 detector images.  Each is binned by sample x sample block means, optionally
 * flipped vertically (flip = 1: flexray's np.flipud), and cropped to the geometry's
 * rows x cols at binned (row0, col0).  dark and scale are [rows][cols] float32 in that
 * cropped frame; the line integral is
 *     -log(max(y, 1e-3)),  y = (x - dark) * scale (op 0, flexray: scale = 1/max(flat-dark, 1))
 *                          y = (x - dark) / scale (op 1, hdf5:    scale = max(flat-dark, 1e-6)).
 * Valid until the geometry is set again. */
int mr_set_raw_correction(mr_ctx* ctx, const float* dark, const float* scale, int op,
                          int frame_rows, int frame_cols, int sample, int flip, int row0,
                          int col0);

/* mr_push_views for raw frames [k][frame_rows][frame_cols] uint16 (host or device
 * pointer): the correction above runs on the device, then preprocessing, filtering
 * and storage as in mr_push_views.  Uploads 2 bytes per raw pixel instead of 4 per
 * corrected pixel (no counterpart in the reference: its loaders correct on the host
 * before recon.py:346 add_projections). */
int mr_push_raw(mr_ctx* ctx, const int64_t* indices, int k, const uint16_t* frames,
                int frames_on_device, int* n_out);

/* Preprocessing applied by mr_push_views to every later block (preprocess.py:59-64
 * apply, called at recon.py:357): rotation by -tilt degrees about the detector centre,
 * then a Gaussian low-pass of sigma = gaussian detector pixels; 0 disables each. */
int mr_set_preprocess(mr_ctx* ctx, double gaussian, double tilt);

/* preprocess.apply on a host block [rows][k][cols] float32 (preprocess.py:59-64;
 * gaussian_blur :8-28 with tilt = 0, rotate_projections :31-56 with gaussian = 0 and
 * tilt = -degrees).  out: host [rows][k][cols] float32. */
int mr_preprocess(mr_ctx* ctx, const float* in, int rows, int k, int cols, double gaussian,
                  double tilt, float* out);

/* Views stored so far (recon.py:319 self.n). */
int mr_num_views(mr_ctx* ctx, int* n);

/* Back-projection of an arbitrary point list: replaces recon.py:405-424
 * reconstruct_points -> bp_cone_numba (recon.py:193-219) / bp_parallel_numba
 * (recon.py:175-191).  pts: host [M][3] float64 (mm); uses the first n_views
 * stored views (a snapshot, recon.py:401-403) with scale pi / n_views.
 * out: host [M] float64.  n_views = 0 gives zeros. */
int mr_bp_points(mr_ctx* ctx, const double* pts, int64_t M, double cor_shift,
                 int n_views, double* out);

/* A slice / grid task: points x(q, p, s) = o + q*c + p*a + s*b (mm, float64),
 * q < nq, p < np, s < ns, written to
 * out[out_offset + q*stride_q + p*stride_p + s*stride_s].
 * mask_box = 1 evaluates the in-box mask of recon.py:253-265 (|x| <= box per axis)
 * with float64 points built in numpy's operation order pts = (mo + t_s*me1) + t_p*me2,
 * t_x = (x - mhalf) * mstep, and writes 0 outside (recon.py:434-440).
 * cor_shift: rotation-centre shift in detector columns (recon.py:412-424). */
typedef struct mr_task {
    double o[3], a[3], b[3], c[3];
    double mo[3], me1[3], me2[3];
    double box[3];
    double mhalf, mstep;
    double cor_shift;
    int64_t out_offset, stride_q, stride_p, stride_s;
    int32_t nq, np, ns, mask_box;
} mr_task;

/* Back-projection of a batch of tasks in one launch: replaces recon.py:426-440
 * reconstruct_plane (one task per plane), recon.py:442-475 reconstruct_volume /
 * preview_volume (one task for the whole grid) and the trial slices of
 * recon.py:498-512 estimate_cor.  out: float32, host (out_on_device = 0) or device
 * (1), sized by the caller to hold every task's output range.  Every output element
 * of every task is written (points outside the volume box with 0); elements that no
 * task addresses are left as they are.  A slice batch whose per-(tile, view) records
 * would exceed the 512 MB records buffer runs on the work-queue loop of kernel 2 (records
 * built per tile in LDS) under the default configuration; an explicit kernel choice
 * processes it in chunks of tiles instead. */
int mr_bp_tasks(mr_ctx* ctx, const mr_task* tasks, int ntasks, int n_views,
                float* out, int64_t out_len, int out_on_device);

/* ---- volume statistics for the threshold after a volume reconstruction (model.py:18-29
 * otsu_threshold, which the reference computes with numpy on the host)
 *
 * Percentiles of the non-zero voxels: out[i] = np.percentile(vol[vol != 0], 100 q[i])
 * (method "linear") for nq <= 4 fractions q[i] in [0, 1] (pass them as numpy computes
 * them: np.true_divide(percent, np.float32(100))).  Exact order statistics by radix
 * select on the device; *nonzero receives the count (MR_EVALUE when it is 0).  vol is
 * float32 [n], a host (on_device = 0) or device (1) pointer. */
int mr_nonzero_percentiles(mr_ctx* ctx, const float* vol, int64_t n, int on_device,
                           const double* q, int nq, double* out, int64_t* nonzero);

/* hist[bins] = np.histogram(np.clip(vol[vol != 0], lo, hi), bins, range=(lo, hi))[0],
 * with edges[bins + 1] = np.linspace(lo, hi, bins + 1) supplied by the caller. */
int mr_clip_histogram(mr_ctx* ctx, const float* vol, int64_t n, int on_device, double lo,
                      double hi, int bins, const double* edges, int64_t* hist);

/* ---- the entry point that allocates no device memory */
#define MR_STATIC_SCRATCH_BYTES  (256 * 1024)   /* per GPU, held until the process ends */
#define MR_STATIC_HIST_MAX_BINS  8192

/* mr_clip_histogram of a volume already in device memory (vol_device, float32 [n]),
 * the same counts, without allocating: no hipMalloc, hipFree or allocating library
 * routine.  No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid it).  The
 * counters and edges live in the library's static per-GPU scratch of
 * MR_STATIC_SCRATCH_BYTES, cleared by a kernel at every call: bins above
 * MR_STATIC_HIST_MAX_BINS are refused (MR_EVALUE).  Calls on one GPU share that scratch:
 * the library serialises its own calls, but work that overlaps them on that GPU (for
 * example a captured graph replayed during a call) races with them. */
int mr_clip_histogram_static(mr_ctx* ctx, const float* vol_device, int64_t n, double lo,
                             double hi, int bins, const double* edges, int64_t* hist);

/* ---- back-projection tuning (no counterpart in the reference: its Numba loop has
 * no shape parameters)
 *
 * cfg[4] = {kernel, points per thread, lane layout, views in flight}: kernel 2 (device
 * work queue, records staged in LDS), 4 (records precomputed in global memory; default
 * for F32 cone beam, 2 points, 2 views in flight), 5 (cone only: per-view detector boxes
 * staged in LDS), 6 (kernel 4 reading a row-pair copy of the ring, one wide load per tap
 * quad; default for parallel beam, 4 points, 2 views in flight) or 7 (kernel 4 reading
 * a copy of a 16-bit ring in bands of 4 rows stepping 3, one 16-byte load per tap quad;
 * F32 storage runs it as 6; default for F16 and BF16 storage
 * in cone beam, 2 points, 3 views in flight, with volume requests run as kernel 6, 4
 * points, 2 views);
 * points per thread 1, 2 or 4; lane layout 0 (waves of 2 x 32 points) or 1 (4 x 16);
 * 2 or 3 views in flight (kernels 4, 6 and 7).
 * Until this is called the configuration follows the storage form of the geometry;
 * kernel 0 (other arguments ignored) returns to that default.  Volume requests of a cone
 * geometry whose detector column and depth do not depend on z run the z-column kernel
 * (8 or 9, see mr_last_bp_kernel), and rotation-centre searches of a parallel geometry
 * the shift-group kernel (10), in the default configuration only. */
int mr_set_bp_config(mr_ctx* ctx, int kernel, int ppt, int layout, int depth);
int mr_get_bp_config(mr_ctx* ctx, int* cfg);

/* ---- device memory of a context (no counterpart in the reference, whose projections
 * live in host memory)
 *
 * Kernels 6 and 7 read copies of the ring built on first use: the row-pair copy (twice
 * the ring) and, for 16-bit storage, the band copy (4/3 of the ring).  They are caches:
 * a copy is built only if the device keeps MINIRECAST_COPY_RESERVE_MB (default 1024) MiB
 * free after it and both copies stay within the budget set here (bytes < 0: no budget,
 * the default).  When freeing the copy a request does not read makes room, that copy is
 * freed first; when nothing does, the request runs without the copy (kernel 6, then 7,
 * then 4 on the ring itself; the work-queue loop on the ring), with the same results to
 * float32 rounding.  Setting a budget frees the copies beyond it at once.
 * mr_device_bytes: out[4] = {ring, band copy, row-pair copy, filter scratch} bytes held
 * on the device now. */
int mr_set_ring_copy_budget(mr_ctx* ctx, int64_t bytes);
int mr_device_bytes(mr_ctx* ctx, int64_t* out);

/* Autotuner: times every configuration on an evenly spaced sample of at most
 * max_tiles 32 x 32 tiles of `tasks` (reps timed runs each, after one warm-up, against
 * the first n_views stored views), keeps the fastest for this context and returns it
 * in cfg[4] with its time per run in *best_ms.  out_len sizes a scratch output.  When
 * the tiles timed are all of the request's, the request is cone beam of single-plane
 * tasks and its grid fills the device, the views per launch of the chosen configuration
 * (mr_set_bp_views_per_pass) are then tuned too, among 16, 32, 64, 128 and all. */
int mr_autotune(mr_ctx* ctx, const mr_task* tasks, int ntasks, int n_views, int64_t out_len,
                int max_tiles, int reps, int* cfg, float* best_ms);

/* Views per part of kernels 4, 6 and 7 (cone and parallel beam): a request runs as one
 * launch over view parts of this many views (each part summed in view order, then the
 * parts added in part order), or, when the part sums would exceed 1 GiB, as one launch per
 * block of this many views for single-plane tasks whose grid fills the device (>= 8
 * workgroups per CU).  0 = one sequential sum in one launch.  Until set (here, by
 * MINIRECAST_BP_VCHUNK or by mr_autotune) it follows the storage form: 32 for F32, 64 for
 * F16 and BF16.  No counterpart in the reference. */
int mr_set_bp_views_per_pass(mr_ctx* ctx, int views);
int mr_get_bp_views_per_pass(mr_ctx* ctx, int* views);

/* Rotation-centre search scores: recon.py:498-513.  Each task is one trial image
 * (nq = 1; out_offset/strides are ignored; sizes may differ, so the trial images of
 * several slices can be scored in one call): it is back-projected into a device buffer
 * and scored on the device as sum p ln p of a 128-bin histogram of its non-zero values
 * (np.histogram range min..max).  scores: host [ntasks] float64.  With a parallel
 * geometry, trial images of axial slices whose shifts lie on a grid of 1, 1/2 or 1/4 px
 * are back-projected by the shift-group kernel (kernel 10, mr_last_bp_kernel). */
int mr_cor_scores(mr_ctx* ctx, const mr_task* tasks, int ntasks, int n_views, double* scores);

/* ---- differentiation (no counterpart in the reference: the transposes of the maps above)
 *
 * Adjoint of mr_bp_tasks with respect to the stored views: scatters the task-output
 * weights g (laid out as mr_bp_tasks' output, g_len elements) with the forward's
 * bilinear and FDK distance weights and scale pi / n_views into gq [n_views][rows][cols]
 * float32 (zeroed by the call).  Storage scales are not applied: the gradient is with
 * respect to the dequantised views. */
int mr_bp_tasks_adjoint(mr_ctx* ctx, const mr_task* tasks, int ntasks, int n_views,
                        const float* g, int64_t g_len, int g_on_device, float* gq,
                        int gq_on_device);

/* Derivative of every mr_bp_tasks output element with respect to its task's
 * cor_shift (recon.py:412-424: the detector moved to D - s*u), same layout and
 * write rule as mr_bp_tasks. */
int mr_bp_tasks_dshift(mr_ctx* ctx, const mr_task* tasks, int ntasks, int n_views, float* out,
                       int64_t out_len, int out_on_device);

/* Adjoint of the ingest filter (fdk_filter_block, recon.py:85-104) for views with
 * geometry `indices`: graw [rows][k][cols] = w * F(gq) for gq [k][rows][cols]; the
 * zero-padded ramp filter F is a real even circular convolution and its own transpose,
 * w is the FDK cosine weight (1 for parallel beam). */
int mr_filter_adjoint(mr_ctx* ctx, const int64_t* indices, int k, const float* gq,
                      int gq_on_device, float* graw, int graw_on_device);

/* Forgets the stored views (n = 0) so that the ring can be refilled. */
int mr_reset_views(mr_ctx* ctx);

/* ---- plans: batched requests on a device work queue (recon.py:426-475 callers that
 * re-request the same slices after every arriving batch; no counterpart in the reference)
 *
 * A plan compiles a fixed batch of tasks once: its tiles are enumerated on the device
 * (in-box test included) into a work queue that a persistent back-projection launch
 * drains (reading the band copy of a 16-bit ring in a context configured for kernel 7,
 * or for kernel 6 with parallel beam; for F32 parallel beam with kernel 4 or 6 and later,
 * the row-pair copy).
 * out_len floats of output live in device memory owned by the plan. */
typedef struct mr_plan mr_plan;
int mr_plan_create(mr_ctx* ctx, const mr_task* tasks, int ntasks, int64_t out_len, mr_plan** out);
int mr_plan_destroy(mr_plan* plan);

/* Runs the plan over the first n_views stored views (n_views < 0: all stored views);
 * copies the output to out_host when it is not NULL. */
int mr_plan_run(mr_plan* plan, int n_views, float* out_host);

/* Streaming update (recon.py:346-376 followed by the plan's requests): ingests k views
 * (indices, host block [rows][k][cols]) and brings the plan's output up to all stored
 * views, as one HIP graph captured per batch size and replayed for every later batch of
 * that size.  The block is copied to pinned memory in chunks by three host threads the
 * plan keeps (and the calling thread) and each chunk uploaded as soon as it is complete.  The plan keeps the unscaled FP32 sums of its points and adds only the
 * views not yet summed (the same bits as a full pass); they restart after
 * mr_reset_views.  MR_ERUNTIME once the context's geometry has been set again (as for
 * every plan call).  Copies the output to out_host when it is not NULL. */
int mr_plan_stream(mr_plan* plan, const int64_t* indices, int k, const float* block,
                   float* out_host);

/* Brings the plan's output up to all stored views, however they were ingested
 * (mr_push_views, mr_push_raw, mr_store_views, mr_plan_stream): the views not yet in the
 * plan's running FP32 sums are added by one persistent launch (the same bits as
 * mr_plan_run over all views; they restart after mr_reset_views).  For callers that
 * re-request the same slices as views arrive (the viewer and server pattern of
 * recon.py:426-475).  Copies the output to out_host when it is not NULL. */
int mr_plan_update(mr_plan* plan, float* out_host);

/* Rotation-centre scores (recon.py:498-513) of a plan whose tasks are trial images:
 * single planes of contiguous rows (nq = 1, stride_p = ns, stride_s = 1), any sizes (one
 * plan can hold the searches of several slices).  Back-projects them from the first
 * n_views stored views (< 0: all) in the plan's persistent launch, or with the
 * shift-group kernel when mr_cor_scores would use it (the plan keeps detector rows for
 * every view the ring holds), and scores each image (histogram entropy of its non-zero
 * values, as mr_cor_scores) on the device, both in one HIP graph captured on the first
 * call and replayed after (captured again when the kernel choice changes);
 * scores[ntasks]. */
int mr_plan_cor_scores(mr_plan* plan, int n_views, double* scores);

/* Device pointer and length of the plan's output. */
int mr_plan_output(mr_plan* plan, float** out_device, int64_t* out_len);

/* Duration of the last mr_plan_run / mr_plan_stream on the device (milliseconds). */
int mr_plan_last_ms(mr_plan* plan, double* ms);

/* Parity hook for filter_rows (recon.py:58-64): ramp-filters `rows` host rows of
 * length n (float32) with pitch tau on the context's device. */
int mr_filter_rows(mr_ctx* ctx, const float* in, int rows, int n, double tau,
                   int window, float* out);

/* Reads back stored filtered views [first, first+count) as float32 [count][rows][cols]
 * (dequantised for F16/BF16 storage).  Parity hook for recon.py:365 (self._q). */
int mr_read_views(mr_ctx* ctx, int first, int count, float* out);

/* Synthetic acquisition (phantom.py:56-98 line integrals of constant ellipsoids,
 * float64 per ray, stored float32).  ellipsoids: [K][13] = value, centre[3], M[9]
 * (row-major; M maps x - centre to the unit sphere).  Cone: vectors [A][12];
 * parallel: angles [A], pixel sizes, start distance `far`.  out: device pointer
 * to [rows][A][cols] float32. */
int mr_phantom_cone(mr_ctx* ctx, const double* ellipsoids, int K, const double* vectors,
                    int A, int rows, int cols, float* out_device);
int mr_phantom_parallel(mr_ctx* ctx, const double* ellipsoids, int K, const double* angles,
                        int A, int rows, int cols, double pixel_u, double pixel_v,
                        double far, float* out_device);

/* Device allocation helpers for callers without another allocator (bench, tests). */
int mr_device_alloc(mr_ctx* ctx, int64_t bytes, void** ptr);
int mr_device_free(mr_ctx* ctx, void* ptr);
int mr_memcpy(mr_ctx* ctx, void* dst, const void* src, int64_t bytes, int kind /* 0 H2D, 1 D2H, 2 D2D */);
int mr_synchronize(mr_ctx* ctx);

/* Timing of the last mr_bp_tasks / mr_bp_points kernel on the request stream,
 * measured with HIP events around the back-projection kernel (milliseconds). */
int mr_last_kernel_ms(mr_ctx* ctx, double* ms);

/* The tile kernel alone within the last mr_bp_tasks (kernels 4 and 5 first build their
 * per-view records with a separate kernel; for the others this equals
 * mr_last_kernel_ms). */
int mr_last_tile_kernel_ms(mr_ctx* ctx, double* ms);

/* Dispatches of kernels 4/6/7 within the last mr_bp_tasks: 1 for a request run over view
 * parts in one launch (mr_set_bp_views_per_pass), or one per block of views
 * when its part sums exceed their budget and it is a request of single-plane tasks whose
 * grid fills the device.  0 for the other kernels.  No counterpart in the reference
 * (measurement only). */
int mr_last_tile_kernel_launches(mr_ctx* ctx, int* n);

/* The back-projection kernel of the last mr_bp_tasks, mr_cor_scores or
 * mr_plan_cor_scores: 2 (records in LDS, the work-queue loop), 4, 5, 6 or 7 (the
 * plane-tile kernels), 8 or 9 (the z-column kernel of volume requests: a cone geometry
 * whose detector column and depth do not depend on z, tasks stacking planes along z; 8
 * gathers from the ring, 9 stages each view's detector box of the 16-bit row-pair copy
 * through LDS), 10 (the shift-group kernel of rotation-centre searches: a parallel
 * geometry, axial planes that differ only in cor_shift, the shifts on a grid of 1, 1/2 or
 * 1/4 px, at least 4 per plane on average), 0 when nothing was launched.  Kernels 8-10
 * run in the default configuration only.  No counterpart in the reference
 * (introspection only). */
int mr_last_bp_kernel(mr_ctx* ctx, int* kernel);

/* L2 gather microbenchmark (no counterpart in the reference; measurement only): the rate
 * *gbps (1e9 bytes/s of loaded data) at which the device serves scattered 16-byte loads
 * from an L2-resident buffer of buffer_bytes (a power of two, 8 KiB..1 GiB), when the 64
 * lanes of each load instruction touch lines_per_wave (1..64) distinct 128-byte lines;
 * lane_group 1: consecutive lanes on different lines; 2, 4 or 8: runs of that many
 * consecutive lanes on one line (lines_per_wave x lane_group <= 64); *ms (optional) the
 * best of three timed runs after a warm-up.  It is the L2 side of the back-projection
 * roofline: the band-ring kernel issues one such 16-byte load per point-view. */
int mr_l2_gather_bandwidth(mr_ctx* ctx, int lines_per_wave, int lane_group,
                           int64_t buffer_bytes, double* gbps, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* MINIRECAST_HIP_H */
