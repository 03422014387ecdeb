"""Simulated live acquisition: projections are played into any object with
``add_projections(indices, block)`` (the local reconstructor, a server connection),
in the reference's arrival orders (stream.py:12-29) and at a set acquisition rate."""
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def arrival_order(n, mode="sequential"):
    """Acquisition order (stream.py:12-29): 'sequential' 0..n-1, or 'interleaved' =
    bit-reversed, so every group spreads over the whole angular range."""
    if mode == "sequential":
        return np.arange(n)
    if mode == "interleaved":
        bits = max(1, int(np.ceil(np.log2(n))))
        idx = np.arange(1 << bits)
        rev = np.zeros_like(idx)
        for b in range(bits):
            rev |= ((idx >> b) & 1) << (bits - 1 - b)
        return rev[rev < n]
    raise ValueError(f"unknown order {mode!r}")


class Feeder(threading.Thread):
    """Acquisition simulator: plays ``proj`` ([rows, angles, cols]) into ``sink`` as a
    fixed schedule of shots of ``group`` views in ``order`` (the reference's Feeder,
    stream.py:32-64, keeps the same arguments).

    The detector delivers ``rate`` views per second (rate <= 0: as fast as the sink
    takes them): a shot is handed over no earlier than the moment its last view has
    been acquired, measured from the start on a monotonic clock.  A helper thread
    gathers the next shot's views into a contiguous block while the sink is busy with
    the current one, so the copy out of the [rows, angles, cols] array is not on the
    sink's critical path.  ``stop()`` ends the run before the next shot; ``sent``,
    ``shots_sent`` and ``elapsed`` report progress."""

    def __init__(self, sink, proj, order="sequential", group=20, rate=100.0, verbose=True):
        super().__init__(daemon=True, name="minirecast-feeder")
        self.sink = sink
        self.proj = proj
        self.order = arrival_order(proj.shape[1], order)
        size = max(1, int(group))
        self.shots = [self.order[s:s + size] for s in range(0, len(self.order), size)]
        self.rate = float(rate)
        self.verbose = verbose
        self.stop_event = threading.Event()
        self.sent = 0
        self.shots_sent = 0
        self.elapsed = 0.0

    def _acquired_at(self, views):
        """Seconds after the start at which the detector has delivered `views` views."""
        return views / self.rate if self.rate > 0 else 0.0

    def _gather(self, shot):
        return np.ascontiguousarray(self.proj[:, shot, :])

    def run(self):
        start = time.monotonic()
        with ThreadPoolExecutor(max_workers=1) as gatherer:
            pending = gatherer.submit(self._gather, self.shots[0]) if self.shots else None
            for k, shot in enumerate(self.shots):
                # hold the shot until the detector has acquired its last view
                ahead = start + self._acquired_at(self.sent + len(shot)) - time.monotonic()
                if self.stop_event.wait(ahead) if ahead > 0 else self.stop_event.is_set():
                    break
                block = pending.result()
                if k + 1 < len(self.shots):
                    pending = gatherer.submit(self._gather, self.shots[k + 1])
                self.sink.add_projections(shot, block)
                self.sent += len(shot)
                self.shots_sent += 1
        self.elapsed = time.monotonic() - start
        if self.verbose:
            print(f"[stream] {self.shots_sent} shots, {self.sent} views in "
                  f"{self.elapsed:.2f} s = {self.sent / max(self.elapsed, 1e-9):.1f} views/s")

    def stop(self):
        self.stop_event.set()
