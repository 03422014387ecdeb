"""Benchmark: streaming cone-beam FDK on one MI355X, full-resolution workload.

Workload (full_cone): 30-ellipsoid phantom, 1200 views over 360 deg, 1536 x 1944
detector at 0.075 mm, SOD 400 / SDD 600, ram-lak.  Projections are generated on the
GPU (exact line integrals), filtered into HBM-resident rings stored as float32,
float16 (per-view scale) and bfloat16.  One step = one 3-slice update: three oblique
1024^2 slices through the centre back-projected from all 1200 views, the three
slices in one launch, output left in HBM.

Reported (one JSON line on stdout):
  value      GUPS of the FP16-stored batched update = in-box slice points x views / s
  ms_per_step, plus per-form ms / GUPS / rel-L2 (f32 batched, f32 as three separate
  launches, f16, bf16); rel-L2 and max-abs against the float64 oracle at 3 x 2048
  sampled in-box pixels (float64 filtering and back-projection of the same float32
  projections, oracle/liboracle_exact.so).
  roofline   for the back-projection tile kernel: compulsory HBM bytes (union of the 2x2
             tap footprints over all requested points, per view, x bytes per sample,
             plus the output) / the duration of one update's tile-kernel launch and the
             part-sum pass that adds its view parts (HIP events around them on their
             stream; one launch per update unless the part sums exceed their budget,
             then one launch per block of 64 views), against 8 TB/s; traffic = measured
             HBM bytes per launch (tile kernel and part-sum pass; traffic_per_update = x
             the launches of one update) from two
             rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) of a short run of the same
             f16 update, each scaled by the counter's reading for a 1 GiB device copy
             (null when rocprofv3 is unavailable or --no-traffic).
  cpu_baseline  the Numba-equivalent CPU restatement (oracle/liboracle_fast.so:
             fast-math, OpenMP over points, views serial, float64 accumulation),
             timed on this host on a bounded sample of the same update.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-check] [--no-cpu]
"""
import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tomography-local-demo_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def footprint_pixels(geom, pts, views):
    """Mean number of detector pixels touched by the 2x2 taps of `pts` over `views`."""
    from minirecast.recon import cone_constants
    K = cone_constants(geom.vectors[views])
    R, C = geom.rows, geom.cols
    counts = []
    mark = np.zeros(R * C, dtype=bool)
    for k in range(len(views)):
        rel = pts - K['s'][k]
        L = rel @ K['n'][k]
        X = K['s'][k] + (K['sdd'][k] / L)[:, None] * rel - K['d'][k]
        col = X @ K['uu'][k] + (C - 1) / 2.0
        row = X @ K['vv'][k] + (R - 1) / 2.0
        c0 = np.floor(col).astype(np.int64)
        r0 = np.floor(row).astype(np.int64)
        mark[:] = False
        for dr in (0, 1):
            for dc in (0, 1):
                rr, cc = r0 + dr, c0 + dc
                ok = (rr >= 0) & (rr < R) & (cc >= 0) & (cc < C)
                mark[rr[ok] * C + cc[ok]] = True
        counts.append(int(mark.sum()))
    return float(np.mean(counts))


def band_lines_per_load(geom, box, planes, views, max_tiles=64, seed=0):
    """(lines, quad lines): the mean number of distinct 128-byte lines that one 16-byte
    band-ring load instruction of the default 16-bit slice kernel touches, and the sum
    over its 16 lane quads (lanes 4q..4q+3) of the distinct lines each quad touches (64
    lanes: 2 point rows x 32 point columns of a 32 x 32 tile, the kernel's lane layout),
    on a sample of the request's tiles and views: each point's band element (b, c0 + 1)
    and the next, b = floor((floor(row) + 1) / 3), in the band copy's byte layout
    (pitch = roundup(C + 3, 64) elements of 8 bytes)."""
    from minirecast import recon
    from minirecast.recon import cone_constants
    K = cone_constants(geom.vectors[views])
    R, C = geom.rows, geom.cols
    pitch = (C + 3 + 63) // 64 * 64
    rng = np.random.default_rng(seed)
    groups = []
    for p in planes:
        T = recon.plane_task(box, p['origin'], p['normal'], p['n'])
        n = p['n']
        nt = (n + 31) // 32
        for tp in range(nt):
            for ts in range(nt):
                for w in range(16):                 # 2 workgroups x 4 waves x 2 point slots
                    sub, wave, m = w >> 3, (w >> 1) & 3, w & 1
                    pr = tp * 32 + sub * 16 + 2 * wave + 8 * m + np.repeat([0, 1], 32)
                    sc = ts * 32 + np.tile(np.arange(32), 2)
                    if pr.max() >= n or sc.max() >= n:
                        continue
                    groups.append(T['o'] + pr[:, None] * T['a'] + sc[:, None] * T['b'])
    groups = np.array(groups)
    if len(groups) > max_tiles * 16:
        groups = groups[np.sort(rng.choice(len(groups), max_tiles * 16, replace=False))]
    counts, quads = [], []
    for k in range(len(views)):
        rel = groups - K['s'][k]
        L = rel @ K['n'][k]
        X = K['s'][k] + (K['sdd'][k] / L)[..., None] * rel - K['d'][k]
        col = np.clip(X @ K['uu'][k] + (C - 1) / 2.0, -1.0, C)
        row = np.clip(X @ K['vv'][k] + (R - 1) / 2.0, -1.0, R)
        b = (np.floor(row).astype(np.int64) + 1) // 3
        off = b * pitch * 8 + (np.floor(col).astype(np.int64) + 1) * 8
        l0, l1 = off // 128, (off + 15) // 128
        lines = np.sort(np.concatenate([l0, l1], axis=1), axis=1)
        counts.append(1 + (np.diff(lines, axis=1) != 0).sum(axis=1))
        ql = np.sort(np.concatenate([l0.reshape(-1, 16, 4), l1.reshape(-1, 16, 4)], axis=2), axis=2)
        quads.append((1 + (np.diff(ql, axis=2) != 0).sum(axis=2)).sum(axis=1))
    return float(np.mean(np.concatenate(counts))), float(np.mean(np.concatenate(quads)))


def _pmc_pass(counter):
    """One rocprofv3 --pmc pass of a short f16 run of this benchmark (child process):
    the counter per tile launch of the f16 cone update (its tile-kernel dispatches and the
    part-sum pass that adds the view parts of a split launch, over the updates of the
    run, divided by the launches of one update), the launches of one update (which the
    child reports), and the value for the 1 GiB copy run first by the child."""
    out = tempfile.mkdtemp(prefix="mr_pmc_")
    cmd = ["rocprofv3", "--pmc", counter, "-d", out, "-o", "run", "--output-format", "csv",
           "--", sys.executable, "-u", os.path.abspath(__file__), "--pmc-child", "--steps", "2",
           "--warmup", "1", "--no-check", "--no-cpu", "--forms", "f16", "--calibrate-copy"]
    try:
        with open(os.path.join(out, "child.log"), "w") as fh:
            subprocess.run(cmd, stdout=fh, stderr=subprocess.STDOUT, timeout=240, check=False)
        files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            return None
        bp, ps, cp = [], [], []
        with open(files[0]) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                name, val = row.get("Kernel_Name", ""), float(row["Counter_Value"])
                if ("bp_tiles_kernel4<true, __half" in name or
                        "bp_tiles_kernel4ILb1E6__half" in name):
                    bp.append(val)
                elif "part_sum_kernel" in name:
                    ps.append(val)
                elif "copyBuffer" in name or "copy" in name.lower():
                    cp.append(val)
        if not bp or not cp:
            return None
        # one update = `launches` dispatches of the tile kernel (one split launch, or one
        # per block of views) and, for a split launch, one part-sum pass
        launches = 1
        with open(os.path.join(out, "child.log")) as fh:
            for ln in fh:
                if "tile launches per update:" in ln:
                    launches = int(ln.rsplit(":", 1)[1])
        updates = len(bp) / launches
        per_update = (sum(bp) + sum(ps)) / updates
        return per_update / launches, launches, float(max(cp))
    except (OSError, subprocess.SubprocessError, ValueError, KeyError):
        return None
    finally:
        shutil.rmtree(out, ignore_errors=True)


def pmc_traffic():
    """HBM bytes per launch (mean per dispatch) of the f16 tile kernel, and the launches
    of one update: FETCH_SIZE and WRITE_SIZE (two passes, one counter group each), each
    scaled by its reading for the 1 GiB copy."""
    if shutil.which("rocprofv3") is None:
        log("traffic: rocprofv3 not found")
        return None
    res = {}
    for counter, key in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write")):
        r = _pmc_pass(counter)
        if r is None:
            log(f"traffic: {counter} pass gave no reading")
            return None
        bp, launches, copy = r
        scale = float(1 << 30) / copy            # bytes per counter unit
        res[key] = bp * scale
        res[key + "_bytes_per_unit"] = round(scale, 3)
        res["launches_per_update"] = launches
    log(f"traffic: {res}")
    return res


def cpu_threads():
    """(threads, detail) for the CPU baseline: every CPU of the affinity mask, unless the
    environment sets this process's share (OMP_NUM_THREADS, which the GPU hosts set to the
    CPUs a job may use), and the cgroup quota when one is set."""
    mask = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(mask, int(env)) if env.isdigit() and int(env) > 0 else mask
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = round(int(q) / int(period), 2)
    except (OSError, ValueError):
        pass
    return threads, {"affinity_cpus": mask, "omp_num_threads": env or None,
                     "cgroup_cpu_quota": quota}


def cpu_model():
    """The host CPU's model string (/proc/cpuinfo), or "unknown CPU"."""
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


# when a list, _line collects the workload records instead of printing them (the other
# workloads measured by the default run)
_SINK = None


def _line(metric, value, unit, ms, steps, warmup, cfg, extra, higher=True, dtype="f32"):
    d = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps,
         "warmup": warmup, "ms_per_step": ms, "higher_is_better": higher, "scaling": "weak",
         "vs_baseline": None, "dtype": dtype,
         "data": "synthetic (seeded ellipsoid phantom, exact line integrals)", "config": cfg}
    d.update(extra)
    if _SINK is not None:
        _SINK.append(d)
        return
    print(json.dumps(d), flush=True)


def other_workloads(args):
    """The other configurations of the benchmark set, measured with a few steps each so
    that the default run records them too: {name: record} (each record as its own
    --workload line would print it, plus the wall time of the measurement)."""
    global _SINK
    out = {}
    sub = argparse.Namespace(**vars(args))
    sub.steps = min(args.steps, 3)
    for name, fn in (("stream_cone", bench_stream_cone), ("parallel_cor", bench_parallel_cor),
                     ("full_volume", bench_full_volume), ("diff_cone", bench_diff_cone)):
        _SINK = []
        t0 = time.perf_counter()
        try:
            fn(sub)
            rec = _SINK[-1]
            for k in ("n_gpus", "scaling", "vs_baseline", "data"):
                rec.pop(k, None)
            rec["wall_s"] = round(time.perf_counter() - t0, 2)
            out[name] = rec
        finally:
            _SINK = None
        log(f"{name}: {out[name]['value']} {out[name]['unit']} ({out[name]['wall_s']} s)")
    return out


def bench_full_volume(args):
    """The 768 x 768 x 607 volume over the full_cone scan's box (4.3e11 updates), reconstructed into HBM
    from the f16 ring in one request (the z-column kernel: 16 x 16 voxel columns x 16
    planes per workgroup, each view's detector box staged through LDS), then the Otsu
    threshold of model.py computed on the device-resident volume."""
    from minirecast import _hip, model, phantom, recon, workloads
    w = workloads.full_cone()
    g = w.geom
    A, R, C = g.n_angles, g.rows, g.cols
    ctx = _hip.Context(0)
    ctx.set_cone(g.vectors, R, C, w.window, "f16")
    batch = 40
    d_raw = ctx.alloc(R * batch * C * 4)
    for b0 in range(0, A, batch):
        idx = np.arange(b0, min(A, b0 + batch))
        phantom.project_cone_device(w.phantom, type(g)(g.vectors[idx], R, C), ctx, d_raw)
        ctx.push_device(idx, d_raw)
    ctx.free(d_raw)
    box = g.volume_box()
    dims, origin, h = recon.volume_grid(box, w.extra["volume"])
    nvox = int(np.prod(dims))
    T = recon.volume_task(origin, h, dims, 0, int(dims[2]))[None]
    dvol = ctx.alloc(nvox * 4)
    ctx.bp_tasks_device(T, A, dvol, nvox)                  # warm-up (band ring, records)
    ctx.synchronize()
    times, tile_ms = [], []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        ctx.bp_tasks_device(T, A, dvol, nvox)
        ctx.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
        tile_ms.append(ctx.last_tile_ms())
    ms = float(np.median(times))
    t0 = time.perf_counter()
    q = np.true_divide(np.asarray([0.1, 99.95]), np.float32(100))
    (lo, hi), m = ctx.nonzero_percentiles(None, q, dptr=dvol, n=nvox)
    edges = np.linspace(lo, hi, 513)
    hist = ctx.clip_histogram(None, lo, hi, edges, dptr=dvol, n=nvox)
    centers = 0.5 * (edges[1:] + edges[:-1])
    w0 = np.cumsum(hist)
    w1 = w0[-1] - w0
    m0 = np.cumsum(hist * centers) / np.maximum(w0, 1)
    m1 = (np.sum(hist * centers) - np.cumsum(hist * centers)) / np.maximum(w1, 1)
    thr = float(centers[int(np.argmax(w0 * w1 * (m0 - m1) ** 2))])
    otsu_ms = (time.perf_counter() - t0) * 1e3
    ctx.free(dvol)
    upd = nvox * A
    res = dict(dims=[int(d) for d in dims], voxel_mm=float(h), ms=ms,
               tile_kernel_ms=float(np.median(tile_ms)), gups=upd / (ms * 1e-3) / 1e9,
               otsu_ms=otsu_ms, otsu_threshold=thr, nonzero_voxels=m, config=ctx.bp_config(),
               kernel=ctx.last_bp_kernel())
    log(f"full_volume: {res}")
    _line("ms per 768^3 volume from 1200 views (f16 ring) + GUPS", round(res["gups"], 2),
          "GUPS", round(ms, 3), len(times), 1,
          {"workload": "full_cone volume: 768^3 grid over the box from 1200 views, 1536x1944 "
                       "detector", "storage": "f16"}, {"volume": res})


def bench_stream_cone(args):
    """Streaming ingest: 600 views in interleaved batches of 10 (Poisson noise, 384 x 486
    detector); after every batch the three oblique 512^2 slices and the 64^3 preview
    are brought up to date.  Graph path: one mr_plan whose per-batch update (upload,
    filter, persistent back-projection of the batch's new views into the running FP32
    sums of all tasks, from the device work queue) is a captured hipGraph replayed per
    batch; eager path: SliceReconstructor calls (every request back-projects all views
    received so far, as the reference does).  The line's value is the median ms per
    batch of the graph path."""
    from minirecast import _hip, phantom, recon, workloads
    from minirecast.stream import arrival_order
    w = workloads.stream_cone()
    g = w.geom
    A, R, C = g.n_angles, g.rows, g.cols
    proj = phantom.project_cone(w.phantom, g)
    proj = workloads.add_poisson_noise(proj, w.extra["photons"], w.seed + 100)
    box = g.volume_box()
    dims, origin, h = recon.volume_grid(box, w.extra["preview"])
    tasks, off, n_in = [], 0, 0
    for p in w.planes:
        tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], out_offset=off))
        _, m = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        n_in += int(m.sum())
        off += p['n'] ** 2
    vt = recon.volume_task(origin, h, dims, 0, int(dims[2]))
    vt['out_offset'] = off
    tasks.append(vt)
    nvox = int(np.prod(dims))
    out_len = off + nvox
    order = arrival_order(A, w.extra["order"])
    batch = w.extra["batch"]
    blocks = [(order[i:i + batch], np.ascontiguousarray(proj[:, order[i:i + batch], :]))
              for i in range(0, A, batch)]
    # a full pass per batch (the reference's requests) and each view once (running sums)
    updates = sum(min((b + 1) * batch, A) * (n_in + nvox) for b in range(len(blocks)))
    updates_once = A * (n_in + nvox)

    def run(storage):
        # graph path
        ctx = _hip.Context(0)
        ctx.set_cone(g.vectors, R, C, w.window, storage)
        plan = _hip.Plan(ctx, np.array(tasks, dtype=_hip.TASK_DTYPE), out_len)
        gpu_ms, wall_ms = [], []
        for idx, blk in blocks:
            t0 = time.perf_counter()
            plan.stream(idx, blk, copy=False)
            wall_ms.append((time.perf_counter() - t0) * 1e3)
            gpu_ms.append(plan.last_ms())
        out_graph = plan.run(-1, copy=True)
        plan.close()
        # eager path
        rec = recon.SliceReconstructor(None, g, window=w.window, verbose=False, storage=storage)
        eager_ms = []
        for idx, blk in blocks:
            t0 = time.perf_counter()
            rec.add_projections(idx, blk)
            imgs = rec.reconstruct_planes(w.planes)
            vol, _, _ = rec.preview_volume(w.extra["preview"])
            eager_ms.append((time.perf_counter() - t0) * 1e3)
        ref = np.concatenate([im.ravel() for im in imgs] + [vol.ravel()])
        err = float(np.linalg.norm(out_graph - ref) / np.linalg.norm(ref))
        res = dict(graph_ms_per_batch_median=float(np.median(wall_ms)),
                   graph_gpu_ms_per_batch_median=float(np.median(gpu_ms)),
                   graph_ms_last_batch=float(wall_ms[-1]),
                   eager_ms_per_batch_median=float(np.median(eager_ms)),
                   eager_ms_last_batch=float(eager_ms[-1]), graph_vs_eager_rel_l2=err,
                   batches=len(blocks), first_batch_ms=float(wall_ms[0]),
                   gups_full_pass_equivalent=updates / (sum(wall_ms) * 1e-3) / 1e9,
                   gups_computed=updates_once / (sum(wall_ms) * 1e-3) / 1e9,
                   eager_gups=updates / (sum(eager_ms) * 1e-3) / 1e9)
        return res, ref

    res, ref32 = run("f32")
    res16, ref16 = run("f16")
    res16["f16_vs_f32_rel_l2"] = float(np.linalg.norm(ref16 - ref32) / np.linalg.norm(ref32))
    res["f16"] = res16
    log(f"stream_cone: {res}")
    _line("ms per streamed batch (10 views ingested + 3 oblique 512^2 slices + 64^3 preview)",
          round(res["graph_ms_per_batch_median"], 4), "ms", round(res["graph_ms_per_batch_median"], 4),
          len(blocks), 0,
          {"workload": "stream_cone: 600 views in 60 interleaved batches of 10, 384x486 "
                       "detector, 3 oblique 512^2 slices + 64^3 preview per batch",
           "path": "mr_plan hipGraph per batch (upload + filter + persistent BP of the new "
                   "views into running sums), f32 ring; f16 ring under stream.f16"},
          {"stream": res}, higher=False)


def bench_parallel_cor(args):
    """Rotation-centre search, parallel beam: 1800 views, 64 x 2048 detector, 25 shifts
    x axial slices 1024^2 (z = 0), 768^2 (z = +16), 512^2 (z = -16); every trial image of
    a slice size back-projected in one launch and scored (entropy) on the device.  Timed
    for the float32 and float16 rings (per-view scale) as one request per slice size, as
    one replayed graph per slice size (persistent launch + scoring), as one request of all
    75 trial images, and as one graph of all 75 (one persistent launch over the tiles of
    the three sizes + scoring; float32 is the line's value).  The trial images of a slice
    differ only by their shift on a 0.25 px grid, so every form runs the shift-group
    kernel (kernel 10: one window of detector values per point and view serves all 25
    shifts of a slice)."""
    from minirecast import _hip, phantom, recon, workloads
    w = workloads.parallel_cor()
    g = w.geom
    proj = phantom.project_parallel(w.phantom, g)
    box = g.volume_box()
    shifts = w.extra["shifts"]
    groups, upd = [], 0
    for p in w.planes:
        T = np.array([recon.plane_task(box, p['origin'], p['normal'], p['n'], cor_shift=float(sh))
                      for sh in shifts], dtype=_hip.TASK_DTYPE)
        _, m = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        upd += int(m.sum()) * len(shifts) * g.n_angles
        groups.append(T)
    forms = {}
    for storage in ("f32", "f16"):
        rec = recon.SliceReconstructor(proj, g, window=w.window, verbose=False, storage=storage)
        ctx = rec._ctx
        for T in groups:
            ctx.cor_scores(T, rec.n)                     # warm-up
        times = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            sc = [ctx.cor_scores(T, rec.n) for T in groups]
            times.append((time.perf_counter() - t0) * 1e3)
        ms = float(np.median(times))
        # the same search as one replayed graph per slice size (persistent launch over a
        # device queue of every trial image's tiles + device scoring)
        plans = []
        for T in groups:
            P = T.copy()
            npix = int(P[0]['np']) * int(P[0]['ns'])
            P['out_offset'] = np.arange(len(P)) * npix
            plans.append(_hip.Plan(ctx, P, len(P) * npix))
            plans[-1].cor_scores(rec.n)                  # capture + warm-up
        ptimes = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            psc = [pl.cor_scores(rec.n) for pl in plans]
            ptimes.append((time.perf_counter() - t0) * 1e3)
        pms = float(np.median(ptimes))
        for pl in plans:
            pl.close()
        # every trial image of the three slice sizes in one request, and in one plan:
        # one persistent launch over a queue of all their tiles + scoring, one graph
        allT = np.concatenate(groups)
        ctx.cor_scores(allT, rec.n)                      # warm-up
        otimes = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            osc = ctx.cor_scores(allT, rec.n)
            otimes.append((time.perf_counter() - t0) * 1e3)
        P = allT.copy()
        npx = P['np'].astype(np.int64) * P['ns'].astype(np.int64)
        P['out_offset'] = np.concatenate([[0], np.cumsum(npx)[:-1]])
        one = _hip.Plan(ctx, P, int(npx.sum()))
        one.cor_scores(rec.n)                            # capture + warm-up
        gtimes = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            gsc = one.cor_scores(rec.n)
            gtimes.append((time.perf_counter() - t0) * 1e3)
        kernel = ctx.last_bp_kernel()
        one.close()
        cuts = np.cumsum([len(T) for T in groups])[:-1]
        best = lambda x: [float(shifts[int(np.argmax(y))]) for y in np.split(x, cuts)]
        forms[storage] = dict(ms_per_search=ms, gups=upd / (ms * 1e-3) / 1e9,
                              best_shift_per_slice=[float(shifts[int(np.argmax(x))]) for x in sc],
                              config=ctx.bp_config(),
                              plan_graph_ms_per_search=pms,
                              plan_graph_best_shift_per_slice=[float(shifts[int(np.argmax(x))])
                                                               for x in psc],
                              one_request_ms_per_search=float(np.median(otimes)),
                              one_request_best_shift_per_slice=best(osc),
                              one_graph_ms_per_search=float(np.median(gtimes)),
                              one_graph_gups=upd / (float(np.median(gtimes)) * 1e-3) / 1e9,
                              one_graph_best_shift_per_slice=best(gsc), kernel=kernel)
        del rec, ctx, plans, one                         # frees the context's rings
    res = dict(forms, trial_images=25 * 3, updates=upd)
    log(f"parallel_cor: {res}")
    # the line: float32 ring, the search as one captured graph of all 75 trial images
    ms = forms["f32"]["one_graph_ms_per_search"]
    _line("ms per COR search (25 shifts x 3 axial slices, 1800 views) + GUPS",
          round(upd / (ms * 1e-3) / 1e9, 2), "GUPS", round(ms, 3), args.steps, 1,
          {"workload": "parallel_cor: 1800 views / 180 deg, 64x2048 detector, hann, "
                       "25 shifts x axial 1024^2, 768^2, 512^2; one graph of all 75 "
                       "trial images, float32 ring"}, {"cor": res})


def bench_filter(args):
    """Ingest filtering of the full-resolution scan's views: FDK cosine weight + ramp filter
    (N = 4096 FFTs of 1536 rows x 1944 columns per view) of device-resident raw batches of
    40 views into the ring, for each storage form.  Per view the algorithmic traffic is
    the raw view read (R C 4 B) and the stored view written (R C s B, s = 4 or 2); the FFT
    work is 2 complex N-point transforms per row pair at 5 N log2 N flops each."""
    from minirecast import _hip, phantom, workloads
    w = workloads.full_cone()
    g = w.geom
    R, C = g.rows, g.cols
    batch = 40
    N = 1 << int(np.ceil(np.log2(2 * C)))
    raw_ctx = _hip.Context(0)
    d_raw = raw_ctx.alloc(4 * R * batch * C)
    sub = type(g)(g.vectors[:batch], R, C)
    phantom.project_cone_device(w.phantom, sub, raw_ctx, d_raw)
    raw_ctx.synchronize()
    flops = 2 * ((R + 1) // 2) * 5 * N * np.log2(N)
    res = {}
    for storage in ("f32", "f16", "bf16"):
        esz = 4 if storage == "f32" else 2
        ctx = _hip.Context(0)
        ctx.set_cone(g.vectors, R, C, w.window, storage)
        idx = np.arange(batch)
        for _ in range(max(1, args.warmup)):
            ctx.reset_views()
            ctx.push_device(idx, d_raw)
        times = []
        for _ in range(max(1, args.steps)):
            ctx.reset_views()
            t0 = time.perf_counter()
            ctx.push_device(idx, d_raw)
            times.append((time.perf_counter() - t0) * 1e3)
        ms = float(np.median(times))
        byts = R * C * (4 + esz) * batch
        res[storage] = dict(ms_per_batch=ms, ms_per_view=ms / batch,
                            algorithmic_gbs=byts / (ms * 1e-3) / 1e9,
                            hbm_frac=byts / (ms * 1e-3) / 8e12,
                            fft_tflops=flops * batch / (ms * 1e-3) / 1e12)
        log(f"filter {storage}: {res[storage]}")
        ctx.close()
    raw_ctx.free(d_raw)
    h = res["f32"]
    _line("ms per filtered view (FDK weight + ramp filter, 1536x1944 detector, N = 4096)",
          round(h["ms_per_view"], 5), "ms", round(h["ms_per_batch"], 4), args.steps, args.warmup,
          {"workload": "filter: full_cone ingest, device-resident raw batches of 40 views, "
                       "f32 / f16 / bf16 ring", "batch": batch, "fft_length": N},
          {"filter": res}, higher=False)


def bench_l2_gather(args):
    """The gather microbenchmark behind the back-projection roofline's L2 side: the rate
    of scattered 16-byte loads (the band-ring element load) when the 64 lanes of each
    load instruction touch L distinct 128-byte lines, for buffers from L2-resident (1 and
    4 MiB) through the Infinity Cache (16, 128 MiB) to HBM (1 GiB)."""
    from minirecast import _hip
    ctx = _hip.Context(0)
    sizes = [1 << 20, 1 << 22, 1 << 24, 1 << 27, 1 << 30]
    lines = [1, 2, 4, 6, 7, 8, 16, 32, 64]
    table = {}
    for nb in sizes:
        table[f"{nb >> 20} MiB"] = {str(L): round(ctx.l2_gather_bandwidth(L, nb)[0], 1)
                                    for L in lines}
        log(f"gather GB/s, buffer {nb >> 20} MiB: {table[f'{nb >> 20} MiB']}")
    # runs of g consecutive lanes on one line (L lines per instruction either way): the
    # cost of a load instruction as a function of the lines each group of lanes touches
    groups = {}
    for g in (1, 2, 4, 8):
        groups[str(g)] = {str(L): round(ctx.l2_gather_bandwidth(L, 1 << 20, g)[0], 1)
                          for L in (1, 2, 4, 8, 16, 32) if L * g <= 64}
        log(f"gather GB/s, 1 MiB, runs of {g} lanes per line: {groups[str(g)]}")
    _line("GB/s of scattered 16-byte loads from an L2-resident buffer, 7 lines per load",
          table["1 MiB"]["7"], "GB/s", None, 4, 1,
          {"workload": "l2_gather: 16-byte buffer loads, L distinct 128-byte lines per "
                       "64-lane load instruction, 8 loads in flight per thread, "
                       "8 workgroups of 256 threads per CU"},
          {"gather_gbps": table, "gather_gbps_lane_runs": groups})


def bench_diff_cone(args):
    """Differentiable back-projection (256^3-box phantom, 360 views, 256 x 256 detector,
    3 oblique 256^2 slices): forward + backward through the hand-written adjoint,
    gradient-based COR refinement (true shift 0.75 px) and 20 SIRT-style steps
    (gradient steps on the projections toward target slice values)."""
    import torch
    from minirecast import autograd, phantom, workloads
    w = workloads.diff_cone()
    g = w.geom
    op = autograd.FDKSlices(g, w.planes, w.window)
    p_true = torch.tensor(phantom.project_cone(w.phantom, g), device="cuda")
    wts = torch.tensor(np.random.default_rng(3).normal(size=op.out_len), dtype=torch.float32,
                       device="cuda")
    # forward + backward timing
    def fb():
        proj = p_true.clone().requires_grad_(True)
        s = torch.tensor(0.25, device="cuda", requires_grad=True)
        y = autograd.fdk_slices(proj, s, op)
        (y * wts).sum().backward()
        torch.cuda.synchronize()
    for _ in range(2):
        fb()
    t = []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        fb()
        t.append((time.perf_counter() - t0) * 1e3)
    fb_ms = float(np.median(t))
    # COR refinement: match slices reconstructed at the true shift
    with torch.no_grad():
        y_star = autograd.fdk_slices(p_true, torch.tensor(0.75, device="cuda"), op)
    def gn_step(s):
        y = autograd.fdk_slices(p_true, s, op)
        loss = 0.5 * ((y - y_star) ** 2).sum()
        g_s, = torch.autograd.grad(loss, s)
        # Gauss-Newton on the shift: dy/ds from the library
        ds = op.dshift(p_true, float(s.detach()))
        return float(g_s.detach()) / max(float((ds * ds).sum()), 1e-30)

    # one untimed step: PyTorch's first use of the loss and gradient kernels
    gn_step(torch.tensor(0.0, device="cuda", requires_grad=True))
    s = torch.tensor(0.0, device="cuda", requires_grad=True)
    hist, it_ms = [], []
    t0 = time.perf_counter()
    for _ in range(15):
        t1 = time.perf_counter()
        step = gn_step(s)
        with torch.no_grad():
            s -= step
        hist.append(float(s.detach()))
        it_ms.append((time.perf_counter() - t1) * 1e3)
    cor_ms = (time.perf_counter() - t0) * 1e3
    # 20 SIRT-style (Landweber) steps on the projections toward target slice values:
    # p <- p - g / L, g the autograd gradient of 0.5 |A p - y|^2 (A: projections ->
    # slices, linear), L the largest eigenvalue of A^T A from 8 power iterations
    zero = torch.tensor(0.0, device="cuda")
    with torch.no_grad():
        y_t = autograd.fdk_slices(p_true, zero, op)
    gen = torch.Generator(device=p_true.device)
    gen.manual_seed(11)                  # a fixed start: the estimate is reproducible
    v = torch.randn(p_true.shape, generator=gen, device=p_true.device, dtype=p_true.dtype)
    v /= v.norm()
    for _ in range(8):
        vv = v.clone().requires_grad_(True)
        av = autograd.fdk_slices(vv, zero, op)
        atav, = torch.autograd.grad(0.5 * (av * av).sum(), vv)
        L = float(atav.norm())
        v = atav / L
    p = torch.zeros_like(p_true).requires_grad_(True)
    losses = []
    t0 = time.perf_counter()
    for k in range(20):
        y = autograd.fdk_slices(p, zero, op)
        loss = 0.5 * ((y - y_t) ** 2).sum()
        gp, = torch.autograd.grad(loss, p)
        with torch.no_grad():
            p -= gp / L
        losses.append(float(loss))
    torch.cuda.synchronize()
    sirt_ms = (time.perf_counter() - t0) * 1e3 / 20
    res = dict(forward_backward_ms=fb_ms, cor_refinement=dict(true=0.75, final=hist[-1],
                                                                 iterations=len(hist),
                                                                 ms_total=cor_ms,
                                                                 ms_per_iteration_median=float(np.median(it_ms))),
               sirt=dict(steps=20, ms_per_step=sirt_ms, lipschitz=L, loss_first=losses[0],
                         loss_last=losses[-1], monotone=bool(all(b <= a for a, b in
                                                                  zip(losses, losses[1:])))))
    log(f"diff_cone: {res}")
    n_upd = op.out_len * g.n_angles
    _line("ms per forward+backward (slices -> grad wrt projections and COR shift)",
          round(fb_ms, 4), "ms", round(fb_ms, 4), len(t), 2,
          {"workload": "diff_cone: 360 views, 256x256 detector, 3 oblique 256^2 slices",
           "updates_per_forward": n_upd}, {"diff": res}, higher=False)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="full_cone",
                    choices=["full_cone", "full_volume", "stream_cone", "parallel_cor",
                             "diff_cone", "filter", "l2_gather"])
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-check", action="store_true", help="skip the float64 oracle error")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--views", type=int, default=None, help="debug: fewer views")
    ap.add_argument("--forms", default="f32,f16,bf16", help="storage forms to time")
    ap.add_argument("--no-workloads", action="store_true",
                    help="full_cone: skip the short runs of the other workloads")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE / WRITE_SIZE passes")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run a 1 GiB device copy (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported "
              f"(the path runs on one GPU; there is no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    if args.workload != "full_cone":
        {"stream_cone": bench_stream_cone, "parallel_cor": bench_parallel_cor,
         "diff_cone": bench_diff_cone, "full_volume": bench_full_volume,
         "filter": bench_filter, "l2_gather": bench_l2_gather}[args.workload](args)
        return
    from minirecast import _hip, phantom, recon, workloads
    if args.calibrate_copy:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 1 GiB read + 1 GiB written")
        del a, b
    w = workloads.full_cone()
    geom = w.geom
    if args.views:
        from minirecast.geometry import ConeVecGeometry
        geom = ConeVecGeometry(geom.vectors[:: max(1, 1200 // args.views)][: args.views],
                               geom.rows, geom.cols)
    A, R, C = geom.n_angles, geom.rows, geom.cols
    box = w.geom.volume_box()
    log(f"workload full_cone: {A} views, {R}x{C} detector, box {box}")

    # ---------------------------------------------------------------- rings
    t0 = time.perf_counter()
    forms = tuple(f for f in args.forms.split(",") if f)
    if "f32" not in forms:
        forms = ("f32",) + forms          # the oracle baseline and CPU q come from f32
    ctx = {s: _hip.Context(0) for s in forms}
    for s in forms:
        ctx[s].set_cone(geom.vectors, R, C, w.window, s)
    gen = ctx["f32"]
    batch = 40
    d_raw = gen.alloc(R * batch * C * 4)
    check = not args.no_check
    # float64 oracle at sampled in-box pixels of the three planes
    rng = np.random.default_rng(12345)
    samp_idx, samp_pts, masks = [], [], []
    for p in w.planes:
        pts, mask = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        masks.append(mask)
        flat = np.flatnonzero(mask.ravel())
        sel = np.sort(rng.choice(flat, size=2048, replace=False))
        samp_idx.append(sel)
        samp_pts.append(pts.reshape(-1, 3)[sel])
    samp_all = np.concatenate(samp_pts)
    acc64 = np.zeros(len(samp_all))
    n_in = int(sum(m.sum() for m in masks))
    if check:
        from oracle import oracle_c as OC
        OC.set_threads(cpu_threads()[0])
    host_raw = np.empty((R, batch, C), np.float32)
    oracle_s = 0.0
    filt_ms = {s: 0.0 for s in forms}
    for b0 in range(0, A, batch):
        k = min(batch, A - b0)
        idx = np.arange(b0, b0 + k)
        ph_geom = type(geom)(geom.vectors[idx], R, C)
        phantom.project_cone_device(w.phantom, ph_geom, gen, d_raw)
        for s in forms:
            ts = time.perf_counter()
            ctx[s].push_device(idx, d_raw)
            filt_ms[s] += (time.perf_counter() - ts) * 1e3
        if check:
            to = time.perf_counter()
            raw = host_raw[:, :k, :] if k == batch else np.empty((R, k, C), np.float32)
            gen.d2h(raw, d_raw)
            for j in range(k):
                q64 = OC.fdk_filter_view(raw, j, geom.vectors[b0 + j], w.window)
                OC.bp_cone_acc_view(q64, geom.vectors[b0 + j], 0.0, samp_all, acc64)
            oracle_s += time.perf_counter() - to
    gen.free(d_raw)
    ref = acc64 * (np.pi / A)
    log(f"ingest done in {time.perf_counter() - t0:.1f}s (oracle {oracle_s:.1f}s); "
        f"filter+store ms/view: " + ", ".join(f"{s} {filt_ms[s] / A:.3f}" for s in forms))

    # ---------------------------------------------------------------- timed updates
    tasks = []
    off = 0
    for p in w.planes:
        tasks.append(recon.plane_task(box, p['origin'], p['normal'], p['n'], out_offset=off))
        off += p['n'] ** 2
    tasks = np.array(tasks, dtype=_hip.TASK_DTYPE)
    out_len = off
    results = {}
    for s in forms:
        c = ctx[s]
        dout = c.alloc(out_len * 4)
        for _ in range(args.warmup):
            c.bp_tasks_device(tasks, A, dout, out_len)
        c.synchronize()
        kms = []
        t1 = time.perf_counter()
        tms = []
        for _ in range(args.steps):
            c.bp_tasks_device(tasks, A, dout, out_len)
            kms.append(c.last_kernel_ms())
            tms.append(c.last_tile_ms())
        c.synchronize()
        wall = (time.perf_counter() - t1) * 1e3 / args.steps
        img = np.empty(out_len, np.float32)
        c.d2h(img, dout)
        res = dict(ms=wall, kernel_ms=float(np.mean(kms)), tile_kernel_ms=float(np.mean(tms)),
                   tile_launches=c.last_tile_launches(),
                   gups=n_in * A / (wall * 1e-3) / 1e9, config=c.bp_config())
        if args.pmc_child and s == "f16":
            log(f"tile launches per update: {res['tile_launches']}")
        if check:
            got = np.concatenate([img[o:o + p['n'] ** 2][si] for o, p, si in
                                  zip(np.cumsum([0] + [p['n'] ** 2 for p in w.planes[:-1]]),
                                      w.planes, samp_idx)]).astype(np.float64)
            res["rel_l2"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
            res["max_abs"] = float(np.max(np.abs(got - ref)))
        results[s] = res
        if s == "f32":
            # plain path: the three slices as three separate launches
            for _ in range(args.warmup):
                for t in tasks:
                    c.bp_tasks_device(t[None], A, dout, out_len)
            c.synchronize()
            t1 = time.perf_counter()
            kms = []
            for _ in range(args.steps):
                km = 0.0
                for t in tasks:
                    c.bp_tasks_device(t[None], A, dout, out_len)
                    km += c.last_kernel_ms()
                kms.append(km)
            c.synchronize()
            wall = (time.perf_counter() - t1) * 1e3 / args.steps
            results["f32_separate"] = dict(ms=wall, kernel_ms=float(np.mean(kms)),
                                           gups=n_in * A / (wall * 1e-3) / 1e9)
        if s == "f16":
            # PCIe-inclusive: the same update with the slices copied back to host memory
            # (what reconstruct_planes returns through the Python API)
            c.bp_tasks(tasks, A, out_len)
            t1 = time.perf_counter()
            for _ in range(args.steps):
                c.bp_tasks(tasks, A, out_len)
            res["host_output_ms"] = (time.perf_counter() - t1) * 1e3 / args.steps
        c.free(dout)
        log(f"{s}: {results[s]}")
    log(f"f32 three launches: {results.get('f32_separate')}")

    # ---------------------------------------------------------------- roofline
    all_pts = []
    for p, m in zip(w.planes, masks):
        pts, _ = recon.slice_points(box, p['origin'], p['normal'], p['n'])
        all_pts.append(pts.reshape(-1, 3)[m.ravel()])
    all_pts = np.concatenate(all_pts)
    fviews = np.linspace(0, A - 1, 8).astype(int)
    F = footprint_pixels(geom, all_pts, fviews)
    head = "f16" if "f16" in forms else "f32"
    esz = 2 if head in ("f16", "bf16") else 4
    hbm_bytes = F * A * esz + out_len * 4
    kern_s = results[head]["tile_kernel_ms"] * 1e-3
    launches = max(1, int(results[head].get("tile_launches", 1)))
    # per launch: the update's algorithmic bytes and tile-kernel time over its launches
    # (blocks of views), i.e. the same rate as for the whole update
    achieved = (hbm_bytes / launches) / (kern_s / launches) / 1e9
    traffic = None
    if not args.no_traffic and not args.pmc_child and head == "f16":
        traffic = pmc_traffic()
    tl = int(traffic["fetch"] + traffic["write"]) if traffic else None
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=tl,
                    kernel=f"bp_tiles_kernel4 + part_sum_kernel (cone, {head})",
                    launches_per_update=launches,
                    algorithmic_bytes_per_launch=int(hbm_bytes / launches),
                    algorithmic_bytes=int(hbm_bytes),
                    traffic_per_update=(int(tl * launches) if tl else None),
                    footprint_px_per_view=round(F),
                    gather_bytes_l2=int(4 * esz * n_in * A))
    if traffic:
        roofline["traffic_detail"] = traffic
    # the L2 side: every point-view is one 16-byte band-element load (16-bit forms).  A
    # load instruction's cost follows the distinct lines of each lane quad (the gather
    # microbenchmark: runs of 4 lanes per line serve 32-36 TB/s at 1-8 lines per
    # instruction, runs of 2 ~29 TB/s, every quad on 4 lines ~9.6 TB/s), so the rate is
    # interpolated in the kernel's quad lines (CPU emulation on a sample) between those
    # patterns at its number of lines per instruction
    if head in ("f16", "bf16"):
        lpl, lq = band_lines_per_load(geom, box, w.planes, np.arange(0, A, max(1, A // 6)))
        L = int(min(16, max(1, round(lpl))))
        pts_lq = [(16.0, ctx[head].l2_gather_bandwidth(L, 1 << 20, 4)[0])]
        if L <= 32:
            pts_lq.append((32.0, ctx[head].l2_gather_bandwidth(L, 1 << 20, 2)[0]))
        if L >= 4:
            pts_lq.append((64.0, ctx[head].l2_gather_bandwidth(L, 1 << 20, 1)[0]))
        xs = np.array([x for x, _ in pts_lq])
        inv = np.array([1.0 / g for _, g in pts_lq])     # time per byte, linear in quad lines
        gbps = 1.0 / float(np.interp(lq, xs, inv))
        load_bytes = 16 * n_in * A
        t_l2 = load_bytes / (gbps * 1e9)
        roofline["l2"] = dict(lines_per_load=round(lpl, 2), quad_lines_per_load=round(lq, 2),
                              gather_gbps=round(gbps, 1),
                              gather_gbps_at={f"{int(x)} quad lines, {L} lines": round(g, 1)
                                              for x, g in pts_lq},
                              load_bytes=int(load_bytes), t_l2_ms=round(t_l2 * 1e3, 4),
                              t_hbm_ms=round(hbm_bytes / (HBM_PEAK_GBS * 1e9) * 1e3, 4),
                              t_kernel_ms=round(kern_s * 1e3, 4),
                              frac_l2=round(t_l2 / kern_s, 4))
    log(f"roofline: {roofline}")

    # ---------------------------------------------------------------- CPU baseline
    cpu = None
    if not args.no_cpu:
        from oracle import oracle_c as OC
        threads, cpu_detail = cpu_threads()
        OC.set_threads(threads, "fast")
        # the whole update (all in-box points x all views), median of 3 runs
        sub = np.arange(A)
        q = np.empty((len(sub), R, C), np.float32)
        for i, v in enumerate(sub):
            q[i] = ctx["f32"].read_views(int(v), 1, R, C)[0]
        K = recon.cone_constants(geom.vectors[sub])     # batch ingest: slot v holds view v
        times = []
        for _ in range(3):
            tc = time.perf_counter()
            OC.bp_cone(q, K['s'], K['d'], K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], all_pts,
                       np.pi / A, kind="fast")
            times.append(time.perf_counter() - tc)
        del q
        dt = float(np.median(times))
        gups = len(all_pts) * len(sub) / dt / 1e9
        cpu = dict(value=round(gups, 4), unit="GUPS", cores=threads, kind="port",
                   sample=f"the full 3-slice update: all {len(all_pts)} in-box points x all {A} "
                          f"views, median of 3 runs ({dt * 1e3:.0f} ms each; "
                          f"{sum(times):.1f} s of CPU work); Numba-equivalent C/OpenMP path "
                          f"(numba unavailable) on {cpu_model()}, {threads} OpenMP threads",
                   **cpu_detail)
        log(f"cpu baseline: {cpu}")

    others = None
    if not args.no_workloads and not args.pmc_child and args.views is None:
        ctx.clear()                      # the full_cone rings are freed before the others
        gen = c = None
        import gc
        gc.collect()
        others = other_workloads(args)

    h = results[head]
    line = {
        "metric": "ms per 3-slice update + GUPS at 1200 views; fp16-store vs fp32 @ rel-L2 err",
        "value": round(h["gups"], 2), "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(h["ms"], 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded 30-ellipsoid phantom, exact line integrals)",
        "config": {"workload": "full_cone: 3 oblique 1024^2 slices from 1200 views, "
                               "1536x1944 detector, ram-lak",
                   "storage": head, "batched": True, "points_in_box": n_in, "views": A},
        "forms": {k: {kk: (round(vv, 6) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                  for k, v in results.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if others is not None:
        line["workloads"] = others
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
