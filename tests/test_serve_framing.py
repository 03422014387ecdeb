"""The server's wire format (net.py:40-100): JSON header frame + optional raw array
frame, numeric dtypes only, geometry messages that round-trip.  CPU only."""
from multiprocessing import Pipe

import numpy as np
import pytest


def test_framing_round_trip_and_refusals():
    from minirecast import serve
    a, b = Pipe()
    for arr in (None, np.arange(12, dtype=np.float32).reshape(3, 4),
                np.arange(6, dtype=np.uint16), np.zeros((0, 5), np.int64)):
        serve.send(a, {"cmd": "x", "k": [1, 2]}, arr)
        h, got = serve.recv(b)
        assert h == {"cmd": "x", "k": [1, 2]}
        if arr is None:
            assert got is None
        else:
            assert got.dtype == arr.dtype and got.shape == arr.shape
            np.testing.assert_array_equal(got, arr)
    # a header announcing a non-numeric array is refused before the payload is read
    a.send_bytes(b'{"cmd": "x", "_array": {"dtype": "|O", "shape": [1]}}')
    with pytest.raises(ValueError):
        serve.recv(b)


def test_geometry_messages_round_trip():
    from minirecast import serve
    from minirecast.geometry import ParallelGeometry, circular_cone_vectors
    g = circular_cone_vectors(30, 400.0, 600.0, 20, 24, 0.5, det_shift_u=0.25)
    h, arr = serve.geometry_message(g)
    g2 = serve.geometry_from_message(h, arr)
    np.testing.assert_array_equal(g2.vectors, g.vectors)
    assert (g2.rows, g2.cols, g2.angular_range) == (g.rows, g.cols, g.angular_range)
    np.testing.assert_array_equal(g2.volume_box(), g.volume_box())
    p = ParallelGeometry(np.linspace(0, np.pi, 50, endpoint=False), 8, 40, 0.7, 0.9)
    h, arr = serve.geometry_message(p)
    p2 = serve.geometry_from_message(h, arr)
    np.testing.assert_array_equal(p2.angles, p.angles)
    assert (p2.pixel_u, p2.pixel_v) == (0.7, 0.9)
    np.testing.assert_array_equal(p2.volume_box(), p.volume_box())
