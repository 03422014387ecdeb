"""The slice server's messages (net.py:40-166) answered from the GPU path.

Wire format (same as net.py:40-59): a message is a JSON header sent as one bytes frame,
optionally followed by one frame of raw array bytes described by
``header["_array"] = {"dtype", "shape"}``; only numeric dtypes are accepted.  Any
connection object with ``send_bytes`` / ``recv_bytes`` carries it (multiprocessing
``Connection``, as the reference's Listener/Client use).

``SliceService.handle(header, array)`` answers the reference server's commands —
hello, geometry, projections, end, info, slice, set_cor, auto_cor, preview — with the
same reply headers and float32 arrays, plus:

* ``raw``: uint16 frames with a flat/dark correction (``dark``/``flat`` arrays sent in
  a preceding ``raw_setup`` message), corrected on the device;
* ``slices``: several slices in one request, one back-projection launch.

Concurrent ``slice`` requests (several viewers) are coalesced: the first request of a
burst waits ``coalesce_ms`` for others and all pending slices are back-projected in one
launch of the tile kernel over one pass of the views, each caller receiving its own
image.  A set of slices requested again (a viewer refreshing its planes as projections
arrive) is answered from a device plan with running sums
(``SliceReconstructor.live_planes``): only the views received since its previous request
are back-projected.  ``serve(conn)`` runs the request loop of one connection.
"""
import json
import threading
import time
from collections import OrderedDict

import numpy as np

from .geometry import ConeVecGeometry, ParallelGeometry
from .rawcorr import RawCorrection
from .recon import LivePlanesClosed, SliceReconstructor


# ------------------------------------------------------------------ framing
def send(conn, header, array=None):
    header = dict(header)
    if array is not None:
        array = np.ascontiguousarray(array)
        header["_array"] = {"dtype": array.dtype.str, "shape": list(array.shape)}
    conn.send_bytes(json.dumps(header).encode())
    if array is not None:
        conn.send_bytes(array.tobytes())


def recv(conn):
    header = json.loads(conn.recv_bytes().decode())
    array = None
    spec = header.pop("_array", None)
    if spec is not None:
        dtype = np.dtype(spec["dtype"])
        if dtype.kind not in "fiu":                     # numbers only
            raise ValueError(f"refusing array dtype {dtype}")
        array = np.frombuffer(conn.recv_bytes(), dtype=dtype).reshape(spec["shape"])
    return header, array


def geometry_message(geom):
    """(header, array) describing a geometry (net.py:62-73)."""
    if isinstance(geom, ParallelGeometry):
        return ({"cmd": "geometry", "kind": "parallel", "rows": geom.rows, "cols": geom.cols,
                 "pixel_u": geom.pixel_u, "pixel_v": geom.pixel_v,
                 "box": [float(b) for b in geom.volume_box()]},
                np.asarray(geom.angles, dtype=np.float64))
    return ({"cmd": "geometry", "kind": "cone", "rows": geom.rows, "cols": geom.cols,
             "angular_range": float(geom.angular_range),
             "box": [float(b) for b in geom.volume_box()]},
            np.asarray(geom.vectors, dtype=np.float64))


def geometry_from_message(h, arr):
    if h["kind"] == "parallel":
        return ParallelGeometry(np.asarray(arr, float), int(h["rows"]), int(h["cols"]),
                                float(h["pixel_u"]), float(h["pixel_v"]),
                                box=np.asarray(h["box"], float))
    return ConeVecGeometry(np.asarray(arr, float).reshape(-1, 12), int(h["rows"]),
                           int(h["cols"]), float(h["angular_range"]),
                           box=np.asarray(h["box"], float))


# ------------------------------------------------------------------ service
class _Pending:
    __slots__ = ("plane", "event", "image", "version", "error")

    def __init__(self, plane):
        self.plane = plane
        self.event = threading.Event()
        self.image = None
        self.version = None
        self.error = None


class SliceService:
    """Answers the server's requests with one SliceReconstructor per scan."""

    def __init__(self, window="shepp-logan", backend="auto", gaussian=0.0, tilt=0.0,
                 coalesce_ms=2.0, verbose=False):
        self.opts = dict(window=window, backend=backend, gaussian=gaussian, tilt=tilt)
        self.coalesce_s = max(0.0, float(coalesce_ms)) * 1e-3
        self.verbose = verbose
        self.recon = None
        self.lock = threading.Lock()          # scan replacement
        self._plock = threading.Lock()        # pending slice requests
        self._pending = []
        self._leader = False
        self.raw_correction = None
        self.launches = 0                     # slice launches (coalescing statistics)
        self.slices_served = 0
        # slice sets seen before: None after the first request, then a LivePlanes
        self._live = OrderedDict()
        self._live_lock = threading.Lock()
        self.live_hits = 0

    LIVE_SETS = 8                             # slice sets remembered

    def _planes(self, r, planes):
        """Images of `planes` over all views received so far: a set requested before
        is refreshed from its running sums, a new one back-projected directly."""
        key = (id(r), float(r.cor_shift), json.dumps(planes, sort_keys=True))
        with self._live_lock:
            seen = key in self._live
            live = self._live.get(key)
            if seen:
                self._live.move_to_end(key)
        if not seen:
            imgs = r.reconstruct_planes(planes)
            self._remember(key, None)
            return imgs
        if live is None:
            live = r.live_planes(planes)
            self._remember(key, live)
        try:
            imgs = live.refresh()
        except LivePlanesClosed:
            # evicted or forgotten by another request thread since the lookup above
            return r.reconstruct_planes(planes)
        self.live_hits += 1
        return imgs

    def _remember(self, key, live):
        drop = []
        with self._live_lock:
            old = self._live.get(key)
            if old is not None and old is not live:
                drop.append(old)
            self._live[key] = live
            self._live.move_to_end(key)
            while len(self._live) > self.LIVE_SETS:
                _, v = self._live.popitem(last=False)
                if v is not None:
                    drop.append(v)
        for v in drop:
            v.close()

    def _forget_live(self):
        with self._live_lock:
            items, self._live = list(self._live.values()), OrderedDict()
        for v in items:
            if v is not None:
                v.close()

    def _recon(self):
        if self.recon is None:
            raise RuntimeError("no scan yet - start an adapter (minirecast push ...)")
        return self.recon

    # -------------------------------------------------------------- slices
    def _slice(self, plane):
        """One slice request, coalesced with concurrent ones into one launch."""
        p = _Pending(plane)
        with self._plock:
            self._pending.append(p)
            lead = not self._leader
            if lead:
                self._leader = True
        if lead:
            if self.coalesce_s > 0:
                time.sleep(self.coalesce_s)
            with self._plock:
                batch, self._pending = self._pending, []
                self._leader = False
            self._run(batch)
        p.event.wait()
        if p.error is not None:
            raise p.error
        return p.image, p.version

    def _run(self, batch):
        try:
            r = self._recon()
            images = self._planes(r, [q.plane for q in batch])
            version = r.version
            self.launches += 1
            self.slices_served += len(batch)
            for q, img in zip(batch, images):
                q.image, q.version = img, version
        except Exception as e:                # every waiter gets the error
            for q in batch:
                q.error = e
        finally:
            for q in batch:
                q.event.set()

    # -------------------------------------------------------------- requests
    def handle(self, h, arr=None):
        """(reply header, reply array or None) for one request (net.py:111-160)."""
        cmd = h.get("cmd")
        if cmd == "hello":
            return {"ok": True}, None
        if cmd == "geometry":
            geom = geometry_from_message(h, arr)
            with self.lock:
                self._forget_live()
                self.recon = SliceReconstructor(None, geom, verbose=False, **self.opts)
                self.raw_correction = None
            return {"ok": True}, None
        if cmd == "projections":
            n = self._recon().add_projections(h["indices"], arr)
            return {"ok": True, "n": n}, None
        if cmd == "raw_setup":
            # arr = stacked (dark, flat) binned frames [2, Hb, Wb]
            self.raw_correction = RawCorrection(arr[0], arr[1], tuple(h["frame_shape"]),
                                                style=h.get("style", "flexray"),
                                                sample=int(h.get("sample", 1)),
                                                crop=h.get("crop"))
            return {"ok": True}, None
        if cmd == "raw":
            if self.raw_correction is None:
                raise RuntimeError("no raw correction set (send raw_setup first)")
            n = self._recon().add_raw_projections(h["indices"], arr, self.raw_correction)
            return {"ok": True, "n": n}, None
        if cmd == "end":
            return {"ok": True}, None
        if cmd == "info":
            r = self.recon
            return {"ready": r is not None, **(r.info() if r else {})}, None
        if cmd == "slice":
            plane = dict(origin=h["origin"], normal=h["normal"], n=int(h["n"]),
                         zoom=float(h.get("zoom", 1.0)), center=h.get("center"))
            img, version = self._slice(plane)
            return {"version": version}, img.astype(np.float32)
        if cmd == "slices":
            planes = [dict(origin=p["origin"], normal=p["normal"], n=int(p["n"]),
                           zoom=float(p.get("zoom", 1.0)), center=p.get("center"))
                      for p in h["planes"]]
            r = self._recon()
            imgs = self._planes(r, planes)
            self.launches += 1
            self.slices_served += len(planes)
            flat = np.concatenate([im.astype(np.float32).ravel() for im in imgs])
            return {"version": r.version, "sizes": [int(p["n"]) for p in planes]}, flat
        if cmd == "set_cor":
            self._recon().set_cor_shift(float(h["px"]))
            return {"ok": True}, None
        if cmd == "auto_cor":
            best, _ = self._recon().estimate_cor(verbose=False)
            return {"best": best}, None
        if cmd == "preview":
            vol, origin, spacing = self._recon().preview_volume(int(h.get("maxdim", 64)))
            return {"origin": [float(o) for o in origin], "spacing": float(spacing)}, vol
        return {"error": f"unknown command {cmd!r}"}, None

    def serve(self, conn):
        """Request loop of one connection until the peer closes it."""
        try:
            while True:
                h, arr = recv(conn)
                try:
                    reply, out = self.handle(h, arr)
                except Exception as e:       # keep serving: report the error
                    reply, out = {"error": repr(e)}, None
                send(conn, reply, out)
        except (EOFError, ConnectionError, OSError):
            pass
        finally:
            conn.close()
