"""Pins the float64 oracle (oracle/oracle.py) before it is trusted as the checker.

Each check is chosen so that a plausible mistake (dropped term, wrong sign or index,
transposed operand, off-by-one centring) fails it: brute-force DFTs and convolutions,
closed forms, known answers, and the reference's own composition fixtures.
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O


# ---------------------------------------------------------------- ramp response
@pytest.mark.parametrize("n", [1, 3, 8, 20, 32])
@pytest.mark.parametrize("window", O.WINDOWS)
def test_ramp_response_matches_bruteforce_dft(n, window):
    tau = 0.37
    H, N = O.ramp_response(n, tau, window)
    assert N == 2 ** int(np.ceil(np.log2(2 * n)))
    # explicit spatial kernel and O(N^2) DFT
    h = np.zeros(N)
    for i in range(N):
        k = i if i < N // 2 else i - N
        if k == 0:
            h[i] = 1.0 / (4 * tau * tau)
        elif k % 2:
            h[i] = -1.0 / (np.pi * k * tau) ** 2
    F = np.array([sum(h[j] * np.exp(-2j * np.pi * j * m / N) for j in range(N)) for m in range(N)])
    f = np.array([(m if m < N // 2 else m - N) / N for m in range(N)])
    win = {"ram-lak": np.ones(N), "shepp-logan": np.sinc(f),
           "hann": 0.5 * (1 + np.cos(2 * np.pi * f)), "cosine": np.cos(np.pi * f)}[window]
    np.testing.assert_allclose(H, F.real * tau * win, rtol=1e-12, atol=1e-12 / tau)
    np.testing.assert_allclose(F.imag, 0, atol=1e-12 / tau ** 2)


def test_ramp_response_is_even_and_ramp_like():
    H, N = O.ramp_response(500, 0.5, "ram-lak")
    np.testing.assert_allclose(H[1:], H[1:][::-1], rtol=0, atol=1e-12)
    f = np.fft.fftfreq(N)
    mid = (np.abs(f) > 0.05) & (np.abs(f) < 0.45)
    # discrete Ram-Lak: H(f) ~ |f| / tau away from DC (tau = sample pitch)
    np.testing.assert_allclose(H[mid], np.abs(f[mid]) / 0.5, rtol=0.02)
    assert abs(H[0]) < 1e-2 * H.max()


def test_unknown_window_raises():
    with pytest.raises(ValueError):
        O.ramp_response(10, 1.0, "blackman")


# ---------------------------------------------------------------- filter_rows
def test_filter_rows_matches_circular_convolution():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(3, 11))
    tau = 0.8
    for window in O.WINDOWS:
        H, N = O.ramp_response(11, tau, window)
        heff = np.real(np.fft.ifft(H))              # effective spatial kernel
        pad = np.zeros((3, N))
        pad[:, :11] = x
        want = np.array([[sum(pad[r, j] * heff[(i - j) % N] for j in range(N)) for i in range(11)]
                         for r in range(3)])
        np.testing.assert_allclose(O.filter_rows(x, tau, window), want, rtol=1e-10, atol=1e-12)


def test_filter_rows_delta_returns_tau_h():
    n, tau = 16, 0.3
    d = np.zeros(n)
    d[0] = 1.0
    N = O.fft_length(n)
    np.testing.assert_allclose(O.filter_rows(d, tau, "ram-lak"), tau * O.ramp_kernel(N, tau)[:n],
                               rtol=1e-12, atol=1e-12)


def test_filter_rows_linearity():
    rng = np.random.default_rng(1)
    a, b = rng.normal(size=(2, 4, 30))
    lhs = O.filter_rows(2.0 * a - 3.0 * b, 0.7, "hann")
    rhs = 2.0 * O.filter_rows(a, 0.7, "hann") - 3.0 * O.filter_rows(b, 0.7, "hann")
    np.testing.assert_allclose(lhs, rhs, atol=1e-12)


# ---------------------------------------------------------------- cone constants / FDK weight
def test_cone_constants_circular_closed_form():
    V = O.circular_cone_vectors(7, 400.0, 600.0, 0.3, det_shift_u=2.0)
    K = O.cone_constants(V)
    th = np.arange(7) * 2 * np.pi / 7
    np.testing.assert_allclose(K['sdd'], 600.0, rtol=1e-14)
    np.testing.assert_allclose(K['sod'], 400.0, rtol=1e-14)
    np.testing.assert_allclose(K['n'], np.stack([np.cos(th), np.sin(th), 0 * th], 1), atol=1e-15)
    np.testing.assert_allclose(K['uu'] * 0.3 ** 2, V[:, 6:9], atol=1e-15)


def test_fdk_weights_closed_form():
    R, C, pix, shift = 5, 7, 0.3, 2.0
    V = O.circular_cone_vectors(3, 400.0, 600.0, pix, det_shift_u=shift)
    K = O.cone_constants(V)
    w, tau = O.fdk_weights(K, 1, R, C)
    a = (np.arange(C) - (C - 1) / 2 + shift) * pix        # detector shifted along u
    b = (np.arange(R) - (R - 1) / 2) * pix
    np.testing.assert_allclose(w, 600 / np.sqrt(600 ** 2 + a[None, :] ** 2 + b[:, None] ** 2),
                               rtol=1e-14)
    assert tau == pytest.approx(pix * 400 / 600, rel=1e-14)


# ---------------------------------------------------------------- bilinear
def test_bilinear_integer_and_edges():
    img = np.arange(12.0).reshape(3, 4) + 1.0
    r, c = np.meshgrid(np.arange(3.0), np.arange(4.0), indexing="ij")
    np.testing.assert_array_equal(O.bilinear(img, r, c), img)
    # half a pixel outside: the outside taps contribute zero
    assert O.bilinear(img, np.array([-0.5]), np.array([0.0]))[0] == pytest.approx(0.5 * img[0, 0])
    assert O.bilinear(img, np.array([0.0]), np.array([3.5]))[0] == pytest.approx(0.5 * img[0, 3])
    assert O.bilinear(img, np.array([-1.0]), np.array([1.0]))[0] == 0.0
    assert O.bilinear(img, np.array([1.25]), np.array([2.5]))[0] == pytest.approx(
        0.75 * (0.5 * img[1, 2] + 0.5 * img[1, 3]) + 0.25 * (0.5 * img[2, 2] + 0.5 * img[2, 3]))


# ---------------------------------------------------------------- back-projection
def _tiny_cone(seed=0, A=5, R=6, C=9):
    rng = np.random.default_rng(seed)
    V = O.circular_cone_vectors(A, 40.0, 70.0, 1.1, det_shift_u=0.7)
    K = O.cone_constants(V)
    q = rng.normal(size=(A, R, C))
    pts = rng.uniform(-4, 4, size=(23, 3))
    return K, q, pts


def test_bp_cone_matches_scalar_loop():
    K, q, pts = _tiny_cone()
    args = (K['s'], K['d'], K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts, 0.3)
    np.testing.assert_allclose(O.bp_cone(q, *args), O.bp_cone_scalar(q, *args), rtol=1e-12,
                               atol=1e-12)


def test_bp_cone_linear_in_q():
    K, q, pts = _tiny_cone(1)
    q2 = np.random.default_rng(5).normal(size=q.shape)
    args = (K['s'], K['d'], K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts, 1.0)
    np.testing.assert_allclose(O.bp_cone(q + 2 * q2, *args),
                               O.bp_cone(q, *args) + 2 * O.bp_cone(q2, *args), atol=1e-10)


def test_cone_cor_shift_is_a_column_shift():
    # u orthogonal to v: D - s*u moves every projected point by +s columns
    K, q, pts = _tiny_cone(2, C=40)
    s = 0.75
    row, col, w = O.cone_coords(K['s'][1], K['d'][1] - s * K['u'][1], K['uu'][1], K['vv'][1],
                                K['n'][1], K['sdd'][1], K['sod'][1], pts, 6, 40)
    row0, col0, w0 = O.cone_coords(K['s'][1], K['d'][1], K['uu'][1], K['vv'][1], K['n'][1],
                                   K['sdd'][1], K['sod'][1], pts, 6, 40)
    np.testing.assert_allclose(col, col0 + s, atol=1e-12)
    np.testing.assert_allclose(row, row0, atol=1e-12)
    np.testing.assert_allclose(w, w0)


def test_bp_parallel_scalar_definition():
    rng = np.random.default_rng(3)
    A, R, C = 4, 5, 12
    q = rng.normal(size=(A, R, C))
    th = rng.uniform(0, np.pi, A)
    pts = rng.uniform(-3, 3, size=(9, 3))
    got = O.bp_parallel(q, np.cos(th), np.sin(th), 0.9, 1.2, pts, 0.5, 0.3)
    want = np.zeros(9)
    for i, (x, y, z) in enumerate(pts):
        for k in range(A):
            col = (-x * np.sin(th[k]) + y * np.cos(th[k])) / 0.9 + (C - 1) / 2 + 0.3
            row = z / 1.2 + (R - 1) / 2
            want[i] += O.bilinear(q[k], np.array([row]), np.array([col]))[0]
    np.testing.assert_allclose(got, 0.5 * want, rtol=1e-12)


def _selftest_phantom():
    table = np.array([
        [1.0, .6900, .920, .810, 0.00, 0.0000, 0.00, 0],
        [-0.8, .6624, .874, .780, 0.00, -0.0184, 0.00, 0],
        [-0.2, .1100, .310, .220, 0.22, 0.0000, 0.00, -18],
        [-0.2, .1600, .410, .280, -0.22, 0.0000, 0.00, 18],
        [0.1, .2100, .250, .410, 0.00, 0.3500, -0.15, 0],
        [0.1, .0460, .046, .050, 0.00, 0.1000, 0.25, 0],
        [0.1, .0460, .046, .050, 0.00, -0.1000, 0.25, 0],
        [0.1, .0460, .023, .050, -0.08, -0.6050, 0.00, 0],
        [0.1, .0230, .023, .020, 0.00, -0.6060, 0.00, 0],
        [0.1, .0230, .046, .020, 0.06, -0.6050, 0.00, 0]])
    vals, cents, mats = [], [], []
    for v, a, b, c, x0, y0, z0, phi in table:
        p = np.deg2rad(phi)
        rot = np.array([[np.cos(p), -np.sin(p), 0], [np.sin(p), np.cos(p), 0], [0, 0, 1]])
        vals.append(v)
        cents.append(np.array([x0, y0, z0]) * 50.0)
        mats.append(np.diag([1 / (a * 50), 1 / (b * 50), 1 / (c * 50)]) @ rot.T)
    return np.array(cents), np.array(mats), np.array(vals)


def test_analytic_phantom_reconstruction_meets_selftest_tolerance():
    """FDK of exact line integrals reproduces the phantom (selftest.py:37-57 criteria)."""
    cents, mats, vals = _selftest_phantom()
    V = O.circular_cone_vectors(120, 300.0, 600.0, 1.5 * 128 / 96, det_shift_u=2.0)
    R, C = 72, 96
    proj = O.project_cone(cents, mats, vals, V, R, C)
    s, d, u, v = np.split(V[0], 4)
    mag = np.linalg.norm(d - s) / np.linalg.norm(s)
    box = np.array([C * np.linalg.norm(u) / mag / 2] * 2 + [R * np.linalg.norm(v) / mag / 2])
    rec = O.OracleReconstructor("cone", R, C, box, vectors=V)
    rec.add_projections(np.arange(120), proj)
    img, pts = rec.reconstruct_plane((3.0, -2.0, 5.0), (1.0, 0.5, 0.8), n=48, size=110)
    truth = O.evaluate(cents, mats, vals, pts)
    m = np.linalg.norm(pts, axis=-1) < 45.0
    med = np.median(np.abs(img[m] - truth[m]))
    corr = np.corrcoef(img[m], truth[m])[0, 1]
    assert med < 0.01 and corr > 0.85, (med, corr)


# ---------------------------------------------------------------- plane helpers
def test_plane_points_geometry():
    n, size = 9, 4.5
    normal = np.array([1.0, 0.5, 0.8])
    pts = O.plane_points((1.0, 2.0, 3.0), normal, size, n)
    nh, e1, e2 = O.plane_basis(normal)
    assert abs(e1 @ nh) < 1e-15 and abs(e2 @ nh) < 1e-15 and abs(e1 @ e2) < 1e-15
    np.testing.assert_allclose((pts - (1, 2, 3)) @ nh, 0, atol=1e-13)
    np.testing.assert_allclose(np.linalg.norm(pts[0, 1] - pts[0, 0]), size / n)
    np.testing.assert_allclose(pts[0, 1] - pts[0, 0], size / n * e1)     # columns follow e1
    np.testing.assert_allclose(pts[1, 0] - pts[0, 0], size / n * e2)     # rows follow e2
    np.testing.assert_allclose(pts[n // 2, n // 2], (1, 2, 3), atol=1e-14)
    # near-axial normal switches to e1 = n x y
    _, e1z, _ = O.plane_basis((0, 0.1, 1.0))
    np.testing.assert_allclose(e1z, np.cross([0, 0.1, 1.0], [0, 1, 0]) / np.linalg.norm(
        np.cross([0, 0.1, 1.0], [0, 1, 0])))


def test_slice_points_centre_projection_and_mask():
    box = np.array([5.0, 6.0, 7.0])
    pts, mask = O.slice_points(box, (0, 0, 2.0), (0, 0, 1.0), n=11, center=(1.0, 1.0, 0.0))
    np.testing.assert_allclose(pts[5, 5], (1.0, 1.0, 2.0))
    np.testing.assert_array_equal(mask, np.all(np.abs(pts) <= box, axis=-1))
    assert O.plane_size_for_box(box, (0, 0, 1)) == pytest.approx(12.0)


def test_volume_grid():
    dims, origin, h = O.volume_grid(np.array([10.0, 5.0, 2.0]), 10)
    assert h == 2.0 and list(dims) == [10, 5, 2]
    np.testing.assert_allclose(origin, [-9.0, -4.0, -1.0])


# ---------------------------------------------------------------- COR scoring, order, phantom
def test_histogram_score_closed_form():
    img = np.zeros(300)
    img[:100] = 1.0
    img[100:300] = 2.0      # two populated bins: p = 1/3, 2/3
    want = (1 / 3) * np.log(1 / 3) + (2 / 3) * np.log(2 / 3)
    assert O.histogram_score(img) == pytest.approx(want, rel=1e-14)
    assert O.histogram_score(np.zeros(10)) == 0.0


def test_arrival_order_known_answer():
    assert list(O.arrival_order(10, "interleaved")) == [0, 8, 4, 2, 6, 1, 9, 5, 3, 7]
    assert sorted(O.arrival_order(37, "interleaved")) == list(range(37))


def test_line_integral_sphere_chord():
    cents = np.zeros((1, 3))
    mats = np.eye(3)[None] / 2.0           # sphere of radius 2
    starts = np.array([[-10.0, 0.0, 0.0], [-10.0, 1.0, 0.0], [-10.0, 3.0, 0.0]])
    d = np.array([1.0, 0.0, 0.0])
    got = O.line_integrals(cents, mats, np.array([0.5]), starts, d)
    np.testing.assert_allclose(got, 0.5 * np.array([4.0, 2 * np.sqrt(3.0), 0.0]), atol=1e-14)


def test_preprocess_selftest_invariants():
    rng = np.random.default_rng(1)
    block = rng.normal(1.0, 0.3, (40, 3, 50))
    blur = O.gaussian_blur(block, 2.0)
    assert abs(blur.mean() - block.mean()) < 0.01 and blur.std() < 0.5 * block.std()
    img = np.zeros((41, 1, 41))
    img[20, 0, 30] = 1.0
    rot = O.rotate_projections(img, 90.0)
    r, c = np.unravel_index(np.argmax(rot[:, 0, :]), (41, 41))
    assert (r, c) in ((10, 20), (30, 20))


def test_rotation_direction_closed_forms():
    """Bilinear interpolation reproduces linear images exactly, so rotating f(r, c) = r
    and g(r, c) = c about the centre has a closed form at every pixel whose source lies
    inside the detector.  The reference's inverse map (preprocess.py:42-46) takes output
    pixel (rr, cc) (centred) from source (cos t rr + sin t cc, -sin t rr + cos t cc): a
    +90 degree rotation moves a point right of the centre to above it (rows grow
    downwards), i.e. the selftest's delta at (20, 30) lands at (10, 20), not (30, 20).
    preprocess.apply corrects a tilt by rotating by -tilt (preprocess.py:59-64)."""
    R, C = 33, 48
    rows = np.broadcast_to(np.arange(R, dtype=float)[:, None, None], (R, 2, C)).copy()
    cols = np.broadcast_to(np.arange(C, dtype=float)[None, None, :], (R, 2, C)).copy()
    rr, cc = np.meshgrid(np.arange(R) - (R - 1) / 2.0, np.arange(C) - (C - 1) / 2.0,
                         indexing="ij")
    for deg in (7.5, -7.5, 31.0):
        t = np.deg2rad(deg)
        src_r = np.cos(t) * rr + np.sin(t) * cc + (R - 1) / 2.0
        src_c = -np.sin(t) * rr + np.cos(t) * cc + (C - 1) / 2.0
        inside = (src_r >= 0) & (src_r <= R - 1) & (src_c >= 0) & (src_c <= C - 1)
        assert inside.mean() > 0.5
        got_r = O.rotate_projections(rows, deg)[:, 0, :]
        got_c = O.rotate_projections(cols, deg)[:, 0, :]
        np.testing.assert_allclose(got_r[inside], src_r[inside], atol=1e-12)
        np.testing.assert_allclose(got_c[inside], src_c[inside], atol=1e-12)
        # far outside the detector every tap is zero
        far = (src_r < -1) | (src_r > R) | (src_c < -1) | (src_c > C)
        assert np.all(got_r[far] == 0)
        # the tilt correction of apply is the rotation by -tilt
        np.testing.assert_array_equal(O.preprocess_apply(rows, tilt=deg),
                                      O.rotate_projections(rows, -deg))
    img = np.zeros((41, 1, 41))
    img[20, 0, 30] = 1.0
    rot = O.rotate_projections(img, 90.0)[:, 0, :]
    assert rot[10, 20] == pytest.approx(1.0, abs=1e-12) and rot[30, 20] == pytest.approx(0.0)
    np.testing.assert_allclose(O.rotate_projections(img, 0.0), img)


@pytest.mark.parametrize("sigma", [2.0, 2.5, 3.7])
def test_gaussian_blur_of_a_delta_is_the_sampled_gaussian(sigma):
    """H(f) = exp(-2 (pi f sigma)^2) is the Fourier transform of the unit-area Gaussian
    G(x) = exp(-x^2 / (2 sigma^2)) / (sigma sqrt(2 pi)) (preprocess.py:23).  For sigma >= 2
    H is below 1e-8 at the Nyquist frequency, so the FFT blur of an interior delta equals
    the sampled Gaussian, separably in rows and columns; a wrong sigma convention (a
    factor 2 in the exponent, sigma in other units) or a shift of the crop fails."""
    R, C = 61, 75
    r0, c0 = 29, 40
    img = np.zeros((R, 1, C))
    img[r0, 0, c0] = 1.0
    got = O.gaussian_blur(img, sigma)[:, 0, :]

    def G(x):
        return np.exp(-x ** 2 / (2 * sigma ** 2)) / (sigma * np.sqrt(2 * np.pi))

    want = G(np.arange(R) - r0)[:, None] * G(np.arange(C) - c0)[None, :]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * want.max())
    assert abs(got.sum() - 1.0) < 1e-9


def test_gaussian_blur_constant_and_ramp():
    """Edge padding leaves a constant block unchanged; the symmetric kernel keeps a
    linear ramp unchanged away from the edges (no shift); sigma <= 0 is the identity."""
    block = np.full((30, 2, 40), 0.375)
    np.testing.assert_allclose(O.gaussian_blur(block, 1.7), block, atol=1e-13)
    ramp = np.broadcast_to(np.arange(60, dtype=float)[None, None, :] * 0.5
                           + np.arange(44, dtype=float)[:, None, None] * -0.25,
                           (44, 2, 60)).copy()
    out = O.gaussian_blur(ramp, 2.0)
    m = 15                                       # edge effects decay as exp(-m^2 / 8)
    np.testing.assert_allclose(out[m:-m, :, m:-m], ramp[m:-m, :, m:-m], atol=1e-9)
    np.testing.assert_array_equal(O.gaussian_blur(ramp, 0.0), ramp)
    # apply: tilt first, then the blur; the two do not commute near the edges
    both = O.preprocess_apply(ramp, gaussian=2.0, tilt=5.0)
    np.testing.assert_allclose(both, O.gaussian_blur(O.rotate_projections(ramp, -5.0), 2.0),
                               atol=1e-12)
    assert np.abs(both - O.rotate_projections(O.gaussian_blur(ramp, 2.0), -5.0)).max() > 1e-3


# ---------------------------------------------------------------- reference composition
def _golden_cone(golden):
    R, C = golden['cone_shape']
    V = golden['cone_vectors']
    s, d, u, v = np.split(V[0], 4)
    mag = np.linalg.norm(d - s) / np.linalg.norm(s)
    box = np.array([C * np.linalg.norm(u) / mag / 2] * 2 + [R * np.linalg.norm(v) / mag / 2])
    rec = O.OracleReconstructor("cone", R, C, box, vectors=V, store_f32=True)
    rec.add_projections(np.arange(len(V)), golden['cone_proj'])
    return rec


def test_oracle_matches_reference_composition_cone(golden):
    rec = _golden_cone(golden)
    for i in range(4):
        a = golden[f'cone_plane{i}_args']
        c = None if np.isnan(a[9]) else a[9:12]
        img, _ = rec.reconstruct_plane(a[0:3], a[3:6], n=int(a[6]), size=a[7] or None,
                                       zoom=a[8], center=c)
        assert rel_l2(img, golden[f'cone_plane{i}']) < 1e-12
    img, _ = rec.reconstruct_plane((0, 0, 0), (1, 0.5, 0.8), n=32, cor_shift=1.25)
    assert rel_l2(img, golden['cone_plane_cor']) < 1e-12
    vol, origin, h = rec.reconstruct_volume(maxdim=12)
    assert rel_l2(vol, golden['cone_preview']) < 1e-6          # reference volume is float32
    np.testing.assert_allclose(origin, golden['cone_preview_origin'])
    scores = rec.cor_scores(golden['cone_cor_shifts'], n=32)
    np.testing.assert_allclose(scores, golden['cone_cor_scores'], rtol=1e-12)


def test_oracle_matches_reference_composition_parallel(golden):
    R, C = golden['par_shape']
    pu, pv = golden['par_pixel']
    box = np.array([C * pu / 2, C * pu / 2, R * pv / 2])
    rec = O.OracleReconstructor("parallel", R, C, box, window="hann", angles=golden['par_angles'],
                                pixel_u=pu, pixel_v=pv, store_f32=True)
    rec.add_projections(np.arange(len(golden['par_angles'])), golden['par_proj'])
    # the reference filters parallel rows in complex64 under numpy >= 2
    img, _ = rec.reconstruct_plane((0, 0, 2.0), (0, 0, 1), n=40)
    assert rel_l2(img, golden['par_plane_axial']) < 1e-6
    img, _ = rec.reconstruct_plane((1.0, 0, 0), (1, 0.5, 0.8), n=36, size=100.0)
    assert rel_l2(img, golden['par_plane_oblique']) < 1e-6
    img, _ = rec.reconstruct_plane((0, 0, 0), (0, 0, 1), n=32, cor_shift=-0.75)
    assert rel_l2(img, golden['par_plane_cor']) < 1e-6


def test_oracle_stream_order_matches_reference(golden):
    rec0 = _golden_cone(golden)
    R, C = golden['cone_shape']
    V = golden['cone_vectors']
    live = O.OracleReconstructor("cone", R, C, rec0.box, vectors=V, store_f32=True)
    order = golden['cone_stream_order']
    np.testing.assert_array_equal(order, O.arrival_order(len(V), "interleaved"))
    for i in range(0, len(V), 7):
        idx = order[i:i + 7]
        live.add_projections(idx, golden['cone_proj'][:, idx, :])
    img, _ = live.reconstruct_plane((0, 0, 0), (0.2, 0.1, 1.0), n=32)
    assert rel_l2(img, golden['cone_stream_plane']) < 1e-12


# ---------------------------------------------------------------- adjoints
def test_bp_cone_adjoint_dot_product():
    K, q, pts = _tiny_cone(7, A=4, R=7, C=11)
    g = np.random.default_rng(8).normal(size=len(pts))
    args = (K['s'], K['d'], K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts, 0.7)
    lhs = O.bp_cone(q, *args) @ g
    rhs = np.sum(q * O.bp_cone_adjoint(g, *args, 7, 11))
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_fdk_filter_adjoint_dot_product():
    rng = np.random.default_rng(9)
    V = O.circular_cone_vectors(3, 400.0, 600.0, 0.3, det_shift_u=0.4)
    K = O.cone_constants(V)
    x = rng.normal(size=(6, 3, 13))
    z = rng.normal(size=(3, 6, 13))
    for window in O.WINDOWS:
        lhs = np.sum(O.fdk_filter_block(x, K, window) * z)
        rhs = np.sum(x * O.fdk_filter_adjoint(z, K, window))
        assert lhs == pytest.approx(rhs, rel=1e-12)


def test_bp_cone_dshift_matches_finite_difference():
    K, q, pts = _tiny_cone(10, A=5, R=8, C=14)
    pts = pts * 0.5
    base = (K['s'], K['d'], K['u'], K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts, 0.3)
    s0, h = 0.31, 1e-6

    def f(sft):
        D = K['d'] - sft * K['u']
        return O.bp_cone(q, K['s'], D, K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts, 0.3)
    fd = (f(s0 + h) - f(s0 - h)) / (2 * h)
    np.testing.assert_allclose(O.bp_cone_dshift(q, *base, s0), fd, rtol=1e-5, atol=1e-7)


# ---------------------------------------------------------------- raw-count correction
@pytest.mark.parametrize("style", ["flexray", "hdf5"])
def test_raw_correct_known_answers(style):
    """The loaders' correction: transmission 1 gives 0, transmission <= 1e-3 is clipped
    to -log(1e-3), a constant frame gives -log((I - D) / (F - D)), and binning, flip
    and crop pick the pixels a brute-force loop picks."""
    rng = np.random.default_rng(5)
    s, H, W = 2, 13, 17                       # trailing odd row/column dropped by binning
    hb, wb = H // s, W // s
    dark = rng.uniform(90, 110, (hb, wb)).astype(np.float32)
    flat = rng.uniform(9e3, 1.1e4, (hb, wb)).astype(np.float32)
    crop = (1, 5, 2, 7)
    flip = style == "flexray"
    # frames whose binned values equal flat and dark exactly (constant blocks)
    def frame_of(img):
        f = np.repeat(np.repeat(img, s, 0), s, 1)
        big = np.zeros((H, W), np.float32)
        big[:hb * s, :wb * s] = f
        return big
    for img, want in ((flat, 0.0), (dark, -np.log(np.float32(1e-3)))):
        src = np.flipud(img) if flip else img
        fr = frame_of(np.round(src)).astype(np.uint16)[None]
        d, fl = np.round(dark).astype(np.float32), np.round(flat).astype(np.float32)
        got = O.raw_correct(fr, d, fl, style, s, flip, crop)
        ref = np.full(got.shape, want, np.float32)
        if want == 0.0:
            # the binned frame equals the flat field exactly
            tgt = np.round(img)
            np.testing.assert_array_equal(fr[0][:hb * s:s, :wb * s:s][::-1 if flip else 1], tgt)
        np.testing.assert_allclose(got, ref, atol=2e-6)
    # constant raw frame: closed form
    I = 5000
    fr = np.full((1, H, W), I, np.uint16)
    got = O.raw_correct(fr, dark, flat, style, s, flip, crop)[:, 0, :]
    r0, r1, c0, c1 = crop
    want = -np.log((I - dark.astype(float)) / (flat.astype(float) - dark))[r0:r1, c0:c1]
    np.testing.assert_allclose(got, want, rtol=1e-5)
    # brute force binning, flip and crop on random counts
    fr = rng.integers(0, 65535, (2, H, W)).astype(np.uint16)
    got = O.raw_correct(fr, dark, flat, style, s, flip, crop)
    for k in range(2):
        for r in range(r1 - r0):
            for c in range(c1 - c0):
                br = r0 + r
                if flip:
                    br = hb - 1 - br
                bc = c0 + c
                x = float(fr[k, br * s:br * s + s, bc * s:bc * s + s].astype(float).mean())
                dd, ff = float(dark[r0 + r, c0 + c]), float(flat[r0 + r, c0 + c])
                y = (x - dd) / (max(ff - dd, 1.0) if style == "flexray" else max(ff - dd, 1e-6))
                assert abs(got[r, k, c] - (-np.log(max(y, 1e-3)))) < 1e-5


# ---------------------------------------------------------------- volume threshold
def test_percentile_restatement_matches_numpy():
    """The order-statistic form of np.percentile that the device implements (float32
    neighbour difference, float64 interpolation) equals numpy bit for bit."""
    rng = np.random.default_rng(9)
    for _ in range(500):
        n = int(rng.integers(1, 3000))
        v = (rng.normal(size=n) * rng.uniform(0.1, 100)).astype(np.float32)
        if rng.random() < 0.3:
            v = np.round(v)                      # ties
        s = np.sort(v)
        q = np.true_divide([0.1, 99.95], np.float32(100))
        want = np.percentile(v, [0.1, 99.95])
        for qq, w in zip(q, want):
            assert O.percentile_from_sorted(s, qq) == float(w)


def test_otsu_known_answers():
    """Two well separated uniform populations: the threshold lies between them (zeros
    ignored); a constant volume gives the first bin centre of numpy's widened range."""
    rng = np.random.default_rng(4)
    a = rng.uniform(0.1, 0.3, 20000)
    b = rng.uniform(0.8, 1.0, 5000)
    vol = np.concatenate([a, b, np.zeros(7000)]).astype(np.float32)
    rng.shuffle(vol)
    t = O.otsu_threshold(vol.reshape(-1, 8))
    # the between-class variance is flat across the gap and argmax takes its first
    # maximum: the centre of the bin holding the largest value of the lower population
    width = (np.percentile(vol[vol != 0], 99.95) - np.percentile(vol[vol != 0], 0.1)) / 512
    assert a.max() - width <= t <= b.min()
    # a constant volume: numpy widens the range to value -+ 0.5, the between-class
    # variance is zero in every bin and argmax returns the first bin centre
    c = np.full(100, 0.25, np.float32)
    assert O.otsu_threshold(c) == (0.25 - 0.5) + 0.5 / 512
