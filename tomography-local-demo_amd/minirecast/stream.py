"""Simulated live acquisition (stream.py of the reference): projections are pushed in
groups into any object with ``add_projections(indices, block)``."""
import threading
import time

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
    """Push ``proj`` ([rows, angles, cols]) into ``sink`` at ``rate`` projections/s
    (stream.py:32-64); rate <= 0 means as fast as possible."""

    def __init__(self, sink, proj, order="sequential", group=20, rate=100.0, verbose=True):
        super().__init__(daemon=True)
        self.sink = sink
        self.proj = proj
        self.order = arrival_order(proj.shape[1], order)
        self.group = max(1, int(group))
        self.rate = float(rate)
        self.verbose = verbose
        self.stop_event = threading.Event()
        self.sent = 0

    def run(self):
        t0 = time.perf_counter()
        for start in range(0, len(self.order), self.group):
            if self.stop_event.is_set():
                return
            idx = self.order[start:start + self.group]
            self.sink.add_projections(idx, np.ascontiguousarray(self.proj[:, idx, :]))
            self.sent = start + len(idx)
            if self.rate > 0:
                wait = t0 + self.sent / self.rate - time.perf_counter()
                if wait > 0:
                    self.stop_event.wait(wait)
        if self.verbose:
            dt = time.perf_counter() - t0
            print(f"[stream] {self.sent} projections streamed in {dt:.1f}s "
                  f"({self.sent / max(dt, 1e-9):.0f}/s)")

    def stop(self):
        self.stop_event.set()
