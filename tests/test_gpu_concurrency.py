"""Requests while a batch is being ingested (recon.py:357-375 filters outside the lock and
locks only the store): a slice request on one thread must not wait for another thread's
push (host-to-device copy and filter) and must see exactly the views committed before it."""
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_requests_run_during_ingest_and_see_committed_views():
    from minirecast import _hip, recon
    from minirecast.geometry import circular_cone_vectors
    A, R, C, k = 240, 768, 972, 40
    g = circular_cone_vectors(A, 400.0, 600.0, R, C, 0.3)
    ctx = _hip.Context(0)
    ctx.set_cone(g.vectors, R, C, "shepp-logan")
    rng = np.random.default_rng(3)
    block = rng.uniform(0.0, 1.0, size=(R, k, C)).astype(np.float32)   # 119 MB per push
    T = np.array([recon.plane_task(g.volume_box(), (0, 0, 0), (1.0, 0.3, 0.6), 128)],
                 dtype=_hip.TASK_DTYPE)
    pushes, replies = [], []
    done = threading.Event()
    errors = []

    def ingest():
        try:
            for b in range(A // k):
                t0 = time.perf_counter()
                ctx.push(np.arange(b * k, (b + 1) * k), block)
                pushes.append((t0, time.perf_counter()))
        except Exception as e:          # reported by the main thread
            errors.append(e)
        finally:
            done.set()

    def requests():
        try:
            while not done.is_set():
                t0 = time.perf_counter()
                n = ctx.num_views()
                if n == 0:
                    continue
                img = ctx.bp_tasks(T, n, 128 * 128)
                replies.append((t0, time.perf_counter(), n, img))
        except Exception as e:
            errors.append(e)

    print("[concurrency] 6 pushes of 40 views 768 x 972 with slice requests", flush=True)
    th = [threading.Thread(target=ingest), threading.Thread(target=requests)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in th), "threads did not finish"
    assert not errors, errors
    assert ctx.num_views() == A and len(pushes) == A // k
    # every reply is over whole committed batches and equals the same request made later
    assert replies
    for _, _, n, img in replies:
        assert n % k == 0 and 0 < n <= A
    for n in sorted({r[2] for r in replies}):
        ref = ctx.bp_tasks(T, n, 128 * 128)
        for _, _, m, img in replies:
            if m == n:
                assert np.abs(img - ref).max() <= 1e-6 * np.abs(ref).max(), n
    # some requests started and finished while a push was in flight
    inside = sum(any(p0 < r0 and r1 < p1 for p0, p1 in pushes) for r0, r1, _, _ in replies)
    push_ms = np.mean([p1 - p0 for p0, p1 in pushes]) * 1e3
    req_ms = np.median([r1 - r0 for r0, r1, _, _ in replies]) * 1e3
    print(f"[concurrency] {len(replies)} replies, {inside} inside a push; push {push_ms:.1f} ms, "
          f"request {req_ms:.2f} ms", flush=True)
    assert inside >= 1
