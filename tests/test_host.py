"""CPU tests of the host side: the C ABI library loads and exports every declared
symbol, host-only library routines agree with the oracle, the Python mirror of the
reference interface keeps its names, argument meaning and error behaviour."""
import os
import re
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from conftest import ROOT, gpu_available
from oracle import oracle as O

HEADER = os.path.join(ROOT, "include", "minirecast_hip.h")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from minirecast import _hip
    lib = _hip.load()
    names = header_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    bound = {s[0] for s in _hip.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound
    assert lib.mr_abi_version() == 2


def test_task_layout_matches_header():
    """numpy's mr_task dtype has the C struct's size and field offsets."""
    from minirecast import _hip
    fields = [f for f in _hip.TASK_DTYPE.names]
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"minirecast_hip.h\"\nint main(void){\n"
    src += 'printf("%zu\\n", sizeof(mr_task));\n'
    for f in fields:
        src += f'printf("%zu\\n", offsetof(mr_task, {f}));\n'
    src += "return 0;}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-I", os.path.dirname(HEADER), c, "-o", exe])
        vals = [int(x) for x in subprocess.check_output([exe]).split()]
    assert vals[0] == _hip.TASK_DTYPE.itemsize
    assert vals[1:] == [_hip.TASK_DTYPE.fields[f][1] for f in fields]


@pytest.mark.parametrize("n", [1, 5, 96, 486, 1944])
@pytest.mark.parametrize("window", ["ram-lak", "shepp-logan", "hann", "cosine"])
def test_library_ramp_response_matches_oracle(n, window):
    """mr_ramp_response is host float64 code: it runs without a GPU."""
    from minirecast import recon
    H, N = recon.ramp_response(n, 0.075, window)
    Ho, No = O.ramp_response(n, 0.075, window)
    assert N == No
    np.testing.assert_allclose(H, Ho, rtol=1e-12, atol=1e-12 * np.abs(Ho).max())


def test_window_errors_map_to_value_error():
    from minirecast import recon
    with pytest.raises(ValueError):
        recon.ramp_response(10, 1.0, "blackman")
    from minirecast import _hip
    lib = _hip.load()
    N = _hip.ctypes.c_int(0)
    assert lib.mr_ramp_response(10, 1.0, 7, None, _hip.ctypes.byref(N)) == _hip.MR_EVALUE
    assert b"window" in lib.mr_last_error()
    assert lib.mr_ramp_response(0, 1.0, 0, None, _hip.ctypes.byref(N)) == _hip.MR_EVALUE


def test_plane_helpers_bit_identical_to_oracle():
    from minirecast import recon
    box = np.array([9.6, 9.6, 7.3])
    for normal in [(1, 0.5, 0.8), (0, 0, 1), (0, 1, 0.05), (-0.3, 1, 0.4)]:
        for zoom, center in [(1.0, None), (2.5, (1.0, -2.0, 0.5))]:
            a = recon.slice_points(box, (0.5, 1, -1), normal, 37, None, zoom, center)
            b = O.slice_points(box, (0.5, 1, -1), normal, 37, None, zoom, center)
            np.testing.assert_array_equal(a[0], b[0])
            np.testing.assert_array_equal(a[1], b[1])
        assert recon.plane_size_for_box(box, normal) == O.plane_size_for_box(box, normal)
    d1, o1, h1 = recon.volume_grid(box, 64)
    d2, o2, h2 = O.volume_grid(box, 64)
    assert list(d1) == list(d2) and h1 == h2
    np.testing.assert_array_equal(o1, o2)
    np.testing.assert_allclose(recon.planes_intersection(np.eye(3), np.eye(3)), [1, 1, 1])
    np.testing.assert_allclose(recon.planes_intersection([[0, 0, 1], [0, 0, 1]], [[0, 0, 1]] * 2,
                                                         (7, 7, 7)), [7, 7, 7])


def test_plane_task_frame_reproduces_slice_points():
    """The task's affine frame and its float64 mask frame give the same points as
    slice_points (the mask frame bit-exactly)."""
    from minirecast import recon
    box = np.array([9.6, 9.6, 7.3])
    n = 21
    T = recon.plane_task(box, (0.5, 1, -1), (1, 0.5, 0.8), n, zoom=1.3, center=(1, 0, 0))
    pts, mask = recon.slice_points(box, (0.5, 1, -1), (1, 0.5, 0.8), n, zoom=1.3, center=(1, 0, 0))
    p, s = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    aff = T['o'] + p[..., None] * T['a'] + s[..., None] * T['b']
    np.testing.assert_allclose(aff, pts, atol=1e-13)
    ts = (s - T['mhalf']) * T['mstep']
    tp = (p - T['mhalf']) * T['mstep']
    exact = (T['mo'] + ts[..., None] * T['me1']) + tp[..., None] * T['me2']
    np.testing.assert_array_equal(exact, pts)
    np.testing.assert_array_equal(np.all(np.abs(exact) <= T['box'], axis=-1), mask)


def test_plane_task_records_are_kept_and_copied():
    """plane_task keeps the records of recent arguments: a repeated request gives the same
    bytes as a fresh build (with its own shift and output offset), changing a returned
    record does not reach later ones, and more distinct planes than are kept still give
    the right records."""
    from minirecast import recon
    box = np.array([9.6, 9.6, 7.3])
    args = ((0.5, 1, -1), (1, 0.5, 0.8), 21)
    kw = dict(zoom=1.3, center=(1, 0, 0))
    T1 = recon.plane_task(box, *args, cor_shift=0.5, out_offset=7, **kw)
    T1['mask_box'] = 0
    T1['o'][0] = 99.0
    T2 = recon.plane_task(box, *args, cor_shift=-1.25, out_offset=11, **kw)
    fresh = recon._plane_task_frame(box, args[0], args[1], args[2], None, kw['zoom'], kw['center'])
    fresh['cor_shift'], fresh['out_offset'] = -1.25, 11
    assert T2.tobytes() == fresh.tobytes()
    assert T2['mask_box'] == 1 and T2['o'][0] != 99.0
    for k in range(recon._PLANE_TASKS_MAX + 5):
        normal = (1.0, 0.01 * k, 0.8)
        got = recon.plane_task(box, (0, 0, 0), normal, 8 + k % 3)
        want = recon._plane_task_frame(box, (0, 0, 0), normal, 8 + k % 3, None, 1.0, None)
        want['cor_shift'], want['out_offset'] = 0.0, 0
        assert got.tobytes() == want.tobytes(), k
    assert len(recon._PLANE_TASKS) <= recon._PLANE_TASKS_MAX


def test_geometry_contract_matches_oracle():
    from minirecast.geometry import ConeVecGeometry, ParallelGeometry, circular_cone_vectors
    g = circular_cone_vectors(17, 300.0, 600.0, 32, 48, 1.5, det_shift_u=2.0)
    np.testing.assert_array_equal(g.vectors, O.circular_cone_vectors(17, 300.0, 600.0, 1.5,
                                                                     det_shift_u=2.0))
    assert isinstance(g, ConeVecGeometry) and g.n_angles == 17
    # magnification |d - s| / |s| includes the detector's u shift (geometry.py:53-57)
    mag = np.hypot(600.0, 2.0 * 1.5) / 300.0
    np.testing.assert_allclose(g.volume_box(), [48 * 1.5 / mag / 2] * 2 + [32 * 1.5 / mag / 2],
                               rtol=1e-14)
    p = ParallelGeometry(np.arange(10) * 0.1, 4, 6, 0.5, 0.25)
    np.testing.assert_allclose(p.volume_box(), [1.5, 1.5, 0.5])


def test_cone_constants_python_matches_oracle():
    from minirecast import recon
    V = O.circular_cone_vectors(9, 400.0, 600.0, 0.3, det_shift_u=-1.0)
    a, b = recon.cone_constants(V), O.cone_constants(V)
    for k in b:
        np.testing.assert_array_equal(a[k], b[k])


def test_arrival_order_and_feeder():
    from minirecast.stream import Feeder, arrival_order
    assert list(arrival_order(10, "interleaved")) == [0, 8, 4, 2, 6, 1, 9, 5, 3, 7]
    with pytest.raises(ValueError):
        arrival_order(5, "random")

    class Sink:
        def __init__(self):
            self.got = []

        def add_projections(self, idx, block):
            assert block.shape[1] == len(idx)
            self.got.extend(int(i) for i in idx)

    sink = Sink()
    proj = np.zeros((2, 23, 3), np.float32)
    Feeder(sink, proj, order="interleaved", group=5, rate=0, verbose=False).run()
    assert sink.got == list(arrival_order(23, "interleaved"))


def test_workloads_are_deterministic():
    from minirecast import workloads as W
    a, b = W.small_cone(), W.small_cone()
    np.testing.assert_array_equal(a.phantom.packed(), b.phantom.packed())
    assert len(a.phantom.vals) == 10
    half = a.geom.volume_box()
    assert np.all(np.abs(a.phantom.centers) <= 0.3 * 2 * half + 1e-12)
    full = W.full_cone()
    assert (full.geom.rows, full.geom.cols, full.geom.n_angles) == (1536, 1944, 1200)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_reconstructor_fails_loudly_without_gpu():
    from minirecast import recon
    from minirecast.geometry import circular_cone_vectors
    g = circular_cone_vectors(8, 300.0, 600.0, 8, 8, 1.0)
    with pytest.raises(RuntimeError):
        recon.SliceReconstructor(None, g)


def test_reconstructor_argument_errors():
    from minirecast import recon
    from minirecast.geometry import circular_cone_vectors
    g = circular_cone_vectors(8, 300.0, 600.0, 8, 8, 1.0)
    with pytest.raises(ValueError):
        recon.SliceReconstructor(None, None)
    with pytest.raises(ValueError):
        recon.SliceReconstructor(None, g, window="blackman")
    with pytest.raises(ValueError):
        recon.SliceReconstructor(None, g, storage="int8")
    assert recon.HAVE_NUMBA is False


def test_no_cpu_fallback_in_product_code():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "tomography-local-demo_amd", "minirecast")
    for f in os.listdir(pkg):
        if f.endswith(".py"):
            text = open(os.path.join(pkg, f)).read()
            assert "oracle" not in text.replace("the oracle", ""), f


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: --gpus above 1 exits with a message, before any GPU
    work (runs on the CPU)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_reference_backend_names_map_to_hip(monkeypatch):
    """backend="numba" / "numpy" (the reference's names, recon.py:302-306) warn and
    select the GPU path instead of failing; with the HIP back-end forced off the
    constructor fails loudly (no CPU fallback)."""
    from minirecast import _hip, recon
    from minirecast.geometry import circular_cone_vectors
    g = circular_cone_vectors(8, 300.0, 600.0, 8, 8, 1.0)
    monkeypatch.setenv(_hip.FORCE_FAILURE_ENV, "1")
    for name in ("numba", "numpy"):
        with pytest.warns(UserWarning, match="hip"):
            with pytest.raises(RuntimeError, match=_hip.FORCE_FAILURE_ENV):
                recon.SliceReconstructor(None, g, backend=name)
    with pytest.raises(RuntimeError, match="no CPU"):
        recon.SliceReconstructor(None, g)
    with pytest.raises(_hip.HipUnavailable):
        _hip.load()


def test_live_planes_close_waits_for_refresh():
    """LivePlanes.close() from another thread waits for a refresh in flight (both hold
    the kernel lock), and a refresh after close raises LivePlanesClosed."""
    import threading
    import time
    from minirecast import recon

    class FakePlan:
        def __init__(self):
            self.closed = self.busy = self.bad = False

        def update(self):
            self.busy = True
            time.sleep(0.05)
            self.bad |= self.closed
            self.busy = False
            return np.arange(4, dtype=np.float32)

        def close(self):
            self.bad |= self.busy
            self.closed = True

    fake = FakePlan()
    lp = object.__new__(recon.LivePlanes)
    lp._plan, lp._sizes = fake, [2]
    out = []
    t = threading.Thread(target=lambda: out.append(lp.refresh()))
    t.start()
    time.sleep(0.01)
    lp.close()
    t.join()
    assert fake.closed and not fake.bad
    np.testing.assert_array_equal(out[0][0], [[0.0, 1.0], [2.0, 3.0]])
    with pytest.raises(recon.LivePlanesClosed):
        lp.refresh()
    lp.close()                                   # closing twice is harmless


def test_slice_service_eviction_during_refresh():
    """Slice sets evicted (LRU) or forgotten by one request thread while another thread
    refreshes them: no request fails and no closed plan is refreshed; a set whose plan
    was closed under it is answered by a direct back-projection."""
    import threading
    import time
    from minirecast import recon, serve

    class FakeLive:
        def __init__(self):
            self.plan_open = True
            self.lock = threading.Lock()

        def refresh(self):
            with recon._KERNEL_LOCK:
                if not self.plan_open:
                    raise recon.LivePlanesClosed("closed")
                time.sleep(0.002)
                return ["live"]

        def close(self):
            with recon._KERNEL_LOCK:
                self.plan_open = False

    class FakeRec:
        cor_shift = 0.0
        version = 1

        def reconstruct_planes(self, planes):
            return ["direct"]

        def live_planes(self, planes):
            return FakeLive()

    svc = serve.SliceService()
    svc.LIVE_SETS = 2
    rec = FakeRec()
    errors, results = [], []

    def worker(k):
        try:
            for i in range(60):
                plane = [dict(n=(k + i // 4) % 3)]
                results.append(svc._planes(rec, plane)[0])
                if i % 17 == 0:
                    svc._forget_live()
        except Exception as e:          # noqa: BLE001 - the test records every failure
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(6)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors
    assert set(results) <= {"live", "direct"} and "live" in results


def test_feeder_pacing_and_stop():
    """The feeder holds each shot until the detector rate has produced it, and stop()
    ends the run before the next shot."""
    import time
    from minirecast.stream import Feeder, arrival_order

    class Sink:
        def __init__(self):
            self.got, self.t = [], []

        def add_projections(self, idx, block):
            np.testing.assert_array_equal(block[0, :, 0], idx)
            self.got.extend(int(i) for i in idx)
            self.t.append(time.monotonic())

    proj = np.zeros((1, 23, 2), np.float32)
    proj[0, :, 0] = np.arange(23)
    sink = Sink()
    f = Feeder(sink, proj, order="interleaved", group=5, rate=200.0, verbose=False)
    t0 = time.monotonic()
    f.run()
    assert sink.got == list(arrival_order(23, "interleaved"))
    assert f.sent == 23 and f.shots_sent == 5
    # shot k (views up to 5(k+1), the last 23) is not handed over before it is acquired
    for k, t in enumerate(sink.t):
        assert t - t0 >= min(5 * (k + 1), 23) / 200.0 - 1e-3
    slow = Sink()
    f = Feeder(slow, proj, group=4, rate=20.0, verbose=False)
    f.start()
    time.sleep(0.3)
    f.stop()
    f.join(timeout=5)
    assert not f.is_alive() and 0 < f.shots_sent < 6 and f.sent == len(slow.got)
