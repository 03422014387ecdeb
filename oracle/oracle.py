"""Float64 NumPy restatement of minirecast's streaming FDK/FBP path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker.  The product (``minirecast`` + ``libminirecast_hip.so``) never calls it.

Every function cites the reference file:line it restates.  Unlike the reference
(which filters parallel-beam rows in complex64 under numpy >= 2 and stores the
filtered projections as float32) the oracle keeps everything in float64: it is
the "float64 NumPy oracle of the same FDK/FBP" against which every GPU result
states its error.  ``OracleReconstructor(store_f32=True)`` rounds the filtered
views to float32 as the reference stores them (recon.py:64, 320), giving
reference-equivalent values.

How each function is pinned (tests/test_oracle.py):
  ramp_response        brute-force O(N^2) DFT of the spatial kernel; ram-lak
                       |f|/tau away from DC; evenness; window closed forms.
  filter_rows          brute-force circular convolution with h on tiny rows;
                       linearity; delta input returns tau*h.
  cone_constants       closed form for circular orbits (sdd, sod, unit normal).
  fdk_filter_block     closed-form cosine weight on a constant block.
  bilinear             exact on integer coordinates; per-tap zero edges;
                       brute force loop on a tiny image.
  bp_cone / bp_parallel  pure-Python scalar loop (recon.py:193-219 order) on
                       tiny inputs; linearity in q; cone COR-shift identity
                       (D - s*u  ==  col + s for u orthogonal to v); analytic
                       phantom reconstruction within the reference selftest
                       tolerances; golden fixtures of the reference's own
                       composition (tests/golden/).
  plane/slice helpers  invariants (points in plane, spacing size/n, mask) and
                       golden fixtures.
  histogram_score      closed form on a hand-made histogram; numpy.histogram.
  arrival_order        known answer [0,8,4,2,6,1,9,5,3,7].
  gaussian_blur / rotate_projections   selftest.py:104-117 invariants.
  line_integrals       closed-form chord of a sphere.
  adjoints            dot-product tests <A x, y> = <x, A^T y> against the forward
                       maps and central finite differences for the shift derivative
                       (tests/test_oracle.py).
  OracleReconstructor  the reference's own outputs on its analytic phantom
                       (tests/golden/reference_composition.npz): bit-exact for
                       cone beam, 1e-6 for parallel beam (the reference filters
                       parallel rows in complex64).
"""
import numpy as np

WINDOWS = ("ram-lak", "shepp-logan", "hann", "cosine")


# --------------------------------------------------------------------------
# Filtering (recon.py:34-104)
# --------------------------------------------------------------------------
def fft_length(n):
    """N = 2**ceil(log2(2n)) (recon.py:37)."""
    return int(2 ** np.ceil(np.log2(2 * n)))


def ramp_kernel(N, tau):
    """Spatial discrete Ram-Lak kernel h (Kak & Slaney), length N (recon.py:38-42)."""
    idx = np.concatenate([np.arange(0, N // 2), np.arange(-N // 2, 0)])
    h = np.zeros(N)
    h[0] = 1.0 / (4.0 * tau * tau)
    odd = idx % 2 == 1
    h[odd] = -1.0 / (np.pi * idx[odd] * tau) ** 2
    return h


def window_response(N, window):
    """Apodisation window on fftfreq(N) (recon.py:44-54)."""
    f = np.fft.fftfreq(N)
    if window in (None, "ram-lak", "none"):
        return np.ones(N)
    if window == "shepp-logan":
        return np.sinc(f)
    if window == "hann":
        return 0.5 * (1.0 + np.cos(2.0 * np.pi * f))
    if window == "cosine":
        return np.cos(np.pi * f)
    raise ValueError(f"unknown filter window {window!r}")


def ramp_response(n, tau, window="shepp-logan"):
    """Windowed frequency response H[N] and N (recon.py:34-55)."""
    N = fft_length(n)
    h = ramp_kernel(N, tau)
    H = np.real(np.fft.fft(h)) * tau
    return H * window_response(N, window), N


def filter_rows(data, tau, window="shepp-logan"):
    """Zero-padded ramp filter along the last axis, float64 throughout
    (recon.py:58-64; the reference casts the result to float32)."""
    data = np.asarray(data, dtype=np.float64)
    n = data.shape[-1]
    H, N = ramp_response(n, tau, window)
    F = np.fft.rfft(data, n=N, axis=-1)
    F *= H[: N // 2 + 1]
    return np.fft.irfft(F, n=N, axis=-1)[..., :n]


def cone_constants(vectors):
    """Per-view cone constants (recon.py:70-82)."""
    V = np.asarray(vectors, dtype=np.float64)
    s, d, u, v = V[:, 0:3], V[:, 3:6], V[:, 6:9], V[:, 9:12]
    n = np.cross(u, v)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    sign = np.sign(np.einsum('ij,ij->i', d - s, n))
    n *= sign[:, None]
    sdd = np.einsum('ij,ij->i', d - s, n)
    sod = np.einsum('ij,ij->i', -s, n)
    uu = u / np.einsum('ij,ij->i', u, u)[:, None]
    vv = v / np.einsum('ij,ij->i', v, v)[:, None]
    return dict(s=s, d=d, u=u, v=v, uu=uu, vv=vv, n=n, sdd=sdd, sod=sod)


def fdk_weights(K, k, R, C):
    """FDK cosine weight w[R, C] of view k and its isocentre pitch (recon.py:90-102)."""
    cc = np.arange(C) - (C - 1) / 2.0
    rr = np.arange(R) - (R - 1) / 2.0
    s, d, u, v, n = K['s'][k], K['d'][k], K['u'][k], K['v'][k], K['n'][k]
    sdd, sod = K['sdd'][k], K['sod'][k]
    p0 = s + sdd * n
    off = d - p0
    un, vn = np.linalg.norm(u), np.linalg.norm(v)
    a = (off @ u) / un + cc * un
    b = (off @ v) / vn + rr * vn
    w = sdd / np.sqrt(sdd ** 2 + a[None, :] ** 2 + b[:, None] ** 2)
    return w, un * sod / sdd


def fdk_filter_block(block, K, window="shepp-logan"):
    """FDK weighting + ramp filter of a [R, k, C] block -> [k, R, C] float64 (recon.py:85-104)."""
    block = np.asarray(block, dtype=np.float64)
    R, A, C = block.shape
    out = np.empty((A, R, C))
    for k in range(A):
        w, tau_iso = fdk_weights(K, k, R, C)
        out[k] = filter_rows(block[:, k, :] * w, tau_iso, window)
    return out


# --------------------------------------------------------------------------
# Back-projection (recon.py:110-219)
# --------------------------------------------------------------------------
def bilinear(img, r, c):
    """Bilinear sample with each tap zero outside the image (recon.py:110-126, 156-173)."""
    R, C = img.shape
    r = np.asarray(r, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    r0 = np.floor(r).astype(np.int64)
    c0 = np.floor(c).astype(np.int64)
    fr = r - r0
    fc = c - c0
    out = np.zeros(r.shape)
    for dr, wr in ((0, 1.0 - fr), (1, fr)):
        for dc, wc in ((0, 1.0 - fc), (1, fc)):
            rr = r0 + dr
            cc = c0 + dc
            ok = (rr >= 0) & (rr < R) & (cc >= 0) & (cc < C)
            vals = np.zeros(r.shape)
            vals[ok] = img[rr[ok], cc[ok]]
            out += wr * wc * vals
    return out


def bp_parallel(q, cos_t, sin_t, pu, pv, pts, scale, col_shift):
    """Parallel-beam back-projection, float64 (recon.py:129-137, 175-191)."""
    q = np.asarray(q)
    A, R, C = q.shape
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    x, y, z = pts[:, 0], pts[:, 1], pts[:, 2]
    row = z / pv + (R - 1) / 2.0
    acc = np.zeros(len(pts))
    for k in range(A):
        col = (-x * sin_t[k] + y * cos_t[k]) / pu + (C - 1) / 2.0 + col_shift
        acc += bilinear(q[k], row, col)
    return acc * scale


def cone_coords(S, D, UU, VV, N, SDD, SOD, pts, R, C):
    """Detector (row, col) and FDK distance weight of points for one view (recon.py:205-216)."""
    rel = pts - S
    L = rel @ N
    t = SDD / L
    X = S + t[:, None] * rel - D
    col = X @ UU + (C - 1) / 2.0
    row = X @ VV + (R - 1) / 2.0
    w = (SOD / L) ** 2
    return row, col, w


def bp_cone(q, S, D, UU, VV, N, SDD, SOD, pts, scale):
    """Cone-beam FDK back-projection, float64 (recon.py:140-152, 193-219)."""
    q = np.asarray(q)
    A, R, C = q.shape
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    acc = np.zeros(len(pts))
    for k in range(A):
        row, col, w = cone_coords(S[k], D[k], UU[k], VV[k], N[k], SDD[k], SOD[k], pts, R, C)
        acc += w * bilinear(q[k], row, col)
    return acc * scale


def bp_cone_scalar(q, S, D, UU, VV, N, SDD, SOD, pts, scale):
    """Pure-Python scalar loop in exactly the order of bp_cone_numba (recon.py:193-219).
    Only for tiny inputs: it pins the vectorised bp_cone."""
    A, R, C = q.shape
    out = np.zeros(len(pts))
    cr, cc = (R - 1) / 2.0, (C - 1) / 2.0
    for i in range(len(pts)):
        px, py, pz = pts[i]
        acc = 0.0
        for k in range(A):
            rx, ry, rz = px - S[k, 0], py - S[k, 1], pz - S[k, 2]
            L = rx * N[k, 0] + ry * N[k, 1] + rz * N[k, 2]
            t = SDD[k] / L
            X0 = S[k, 0] + t * rx - D[k, 0]
            X1 = S[k, 1] + t * ry - D[k, 1]
            X2 = S[k, 2] + t * rz - D[k, 2]
            col = X0 * UU[k, 0] + X1 * UU[k, 1] + X2 * UU[k, 2] + cc
            row = X0 * VV[k, 0] + X1 * VV[k, 1] + X2 * VV[k, 2] + cr
            w = SOD[k] / L
            r0, c0 = int(np.floor(row)), int(np.floor(col))
            fr, fc = row - r0, col - c0
            v = 0.0
            for dr, wr in ((0, 1.0 - fr), (1, fr)):
                for dc, wc in ((0, 1.0 - fc), (1, fc)):
                    rr_, cc_ = r0 + dr, c0 + dc
                    if 0 <= rr_ < R and 0 <= cc_ < C:
                        v += wr * wc * float(q[k, rr_, cc_])
            acc += w * w * v
        out[i] = acc * scale
    return out


# --------------------------------------------------------------------------
# Adjoints (the transposes of the linear maps above), float64
# --------------------------------------------------------------------------
def bilinear_adjoint(g, r, c, R, C):
    """Transpose of ``bilinear``: scatters g with the same per-tap weights into [R, C];
    taps outside the image are dropped."""
    r = np.asarray(r, np.float64)
    c = np.asarray(c, np.float64)
    r0 = np.floor(r).astype(np.int64)
    c0 = np.floor(c).astype(np.int64)
    fr, fc = r - r0, c - c0
    out = np.zeros((R, C))
    for dr, wr in ((0, 1.0 - fr), (1, fr)):
        for dc, wc in ((0, 1.0 - fc), (1, fc)):
            rr, cc = r0 + dr, c0 + dc
            ok = (rr >= 0) & (rr < R) & (cc >= 0) & (cc < C)
            np.add.at(out, (rr[ok], cc[ok]), (wr * wc * g)[ok])
    return out


def bp_cone_adjoint(g, S, D, UU, VV, N, SDD, SOD, pts, scale, R, C):
    """Transpose of bp_cone with respect to q: [A, R, C] from point weights g [M]."""
    pts = np.asarray(pts, np.float64).reshape(-1, 3)
    A = len(S)
    out = np.zeros((A, R, C))
    for k in range(A):
        row, col, w = cone_coords(S[k], D[k], UU[k], VV[k], N[k], SDD[k], SOD[k], pts, R, C)
        out[k] = bilinear_adjoint(scale * w * g, row, col, R, C)
    return out


def bilinear_grad(img, r, c):
    """(d/dr, d/dc) of ``bilinear`` inside a cell (piecewise-linear interpolant)."""
    R, C = img.shape
    r0 = np.floor(r).astype(np.int64)
    c0 = np.floor(c).astype(np.int64)
    fr, fc = r - r0, c - c0

    def tap(dr, dc):
        rr, cc = r0 + dr, c0 + dc
        ok = (rr >= 0) & (rr < R) & (cc >= 0) & (cc < C)
        v = np.zeros(r.shape)
        v[ok] = img[rr[ok], cc[ok]]
        return v
    a, b, cq, d = tap(0, 0), tap(0, 1), tap(1, 0), tap(1, 1)
    ddc = (1 - fr) * (b - a) + fr * (d - cq)
    ddr = (1 - fc) * (cq - a) + fc * (d - b)
    return ddr, ddc


def bp_cone_dshift(q, S, D, U, UU, VV, N, SDD, SOD, pts, scale, shift):
    """d/ds of bp_cone with D -> D - s*u at s = shift (recon.py:420), per point."""
    q = np.asarray(q)
    A, R, C = q.shape
    pts = np.asarray(pts, np.float64).reshape(-1, 3)
    acc = np.zeros(len(pts))
    for k in range(A):
        Dk = D[k] - shift * U[k]
        row, col, w = cone_coords(S[k], Dk, UU[k], VV[k], N[k], SDD[k], SOD[k], pts, R, C)
        ddr, ddc = bilinear_grad(q[k], row, col)
        acc += w * (ddc * (U[k] @ UU[k]) + ddr * (U[k] @ VV[k]))
    return acc * scale


def fdk_filter_adjoint(gq, K, window="shepp-logan"):
    """Transpose of fdk_filter_block: gq [k, R, C] -> [R, k, C].  The zero-padded
    ramp filter is a real even circular convolution cropped to n samples, so it is
    its own transpose; the cosine weight is diagonal."""
    gq = np.asarray(gq, np.float64)
    A, R, C = gq.shape
    out = np.empty((R, A, C))
    for k in range(A):
        w, tau_iso = fdk_weights(K, k, R, C)
        out[:, k, :] = w * filter_rows(gq[k], tau_iso, window)
    return out


# --------------------------------------------------------------------------
# Plane / volume point generation (recon.py:225-283)
# --------------------------------------------------------------------------
def plane_basis(normal):
    """(n, e1, e2): e1 = z x n unless |n.z| >= 0.9, then n x y (recon.py:225-235)."""
    n = np.asarray(normal, dtype=float)
    n = n / np.linalg.norm(n)
    z = np.array([0.0, 0.0, 1.0])
    if abs(n @ z) < 0.9:
        e1 = np.cross(z, n)
    else:
        e1 = np.cross(n, np.array([0.0, 1.0, 0.0]))
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(n, e1)
    return n, e1, e2


def plane_points(origin, normal, size, n):
    """n x n grid, rows along e2 and columns along e1 (recon.py:238-244)."""
    _, e1, e2 = plane_basis(normal)
    t = (np.arange(n) - (n - 1) / 2.0) * (size / n)
    a, b = np.meshgrid(t, t, indexing='xy')
    return np.asarray(origin, float) + a[..., None] * e1 + b[..., None] * e2


def plane_size_for_box(box, normal):
    """Side covering the box cross-section (recon.py:247-250)."""
    _, e1, e2 = plane_basis(normal)
    return 2.0 * float(max(np.abs(e1) @ box, np.abs(e2) @ box))


def slice_points(box, origin, normal, n=256, size=None, zoom=1.0, center=None):
    """Slice points and in-box mask (recon.py:253-265)."""
    n_hat = np.asarray(normal, float) / np.linalg.norm(normal)
    o = np.asarray(origin, float)
    c = np.zeros(3) if center is None else np.asarray(center, float)
    o = c - ((c - o) @ n_hat) * n_hat
    if size is None:
        size = plane_size_for_box(box, normal) / max(float(zoom), 1e-6)
    pts = plane_points(o, normal, size, n)
    mask = np.all(np.abs(pts) <= box, axis=-1)
    return pts, mask


def volume_grid(box, maxdim):
    """(dims, origin, h) of a maxdim grid over the box (recon.py:278-283)."""
    h = 2.0 * float(np.max(box)) / maxdim
    dims = np.maximum(np.ceil(2.0 * np.asarray(box) / h).astype(int), 2)
    origin = -np.asarray(box) + h / 2.0
    return dims, origin, h


def histogram_score(img):
    """sum p ln p over a 128-bin histogram of the non-zero values (recon.py:510-512)."""
    hist, _ = np.histogram(img[img != 0], bins=128)
    prob = hist[hist > 0] / max(hist.sum(), 1)
    return float(np.sum(prob * np.log(prob)))


# --------------------------------------------------------------------------
# Stateful composition (recon.py:289-519) in float64
# --------------------------------------------------------------------------
class OracleReconstructor:
    """Float64 restatement of SliceReconstructor's composition: arrival-order storage,
    dedupe, scale = pi / n, cone COR = D - shift*u, parallel COR = +col shift,
    masked plane points, z-slab volumes and the entropy COR search.

    ``store_f32=True`` rounds the filtered projections to float32 as the reference
    does (recon.py:64, 320), giving reference-equivalent values."""

    def __init__(self, kind, rows, cols, box, window="shepp-logan", vectors=None,
                 angles=None, pixel_u=1.0, pixel_v=1.0, store_f32=False, gaussian=0.0,
                 tilt=0.0):
        self.kind = kind
        self.R, self.C = int(rows), int(cols)
        self.box = np.asarray(box, float)
        self.window = window
        self.store_f32 = store_f32
        self.gaussian, self.tilt = float(gaussian), float(tilt)
        self.pixel_u, self.pixel_v = float(pixel_u), float(pixel_v)
        if kind == "cone":
            self.K_all = cone_constants(vectors)
            self.n_total = len(vectors)
        else:
            self.angles = np.asarray(angles, float)
            self.n_total = len(self.angles)
        self.q = np.zeros((self.n_total, self.R, self.C))
        self.order = np.zeros(self.n_total, dtype=np.int64)
        self.received = np.zeros(self.n_total, dtype=bool)
        self.n = 0
        self.cor_shift = 0.0

    def add_projections(self, indices, block):
        """recon.py:346-376 (float64 preprocessing and filtering)."""
        indices = np.asarray(indices, dtype=np.int64).ravel()
        block = np.asarray(block, dtype=np.float32)
        if block.shape[1] != len(indices):
            raise ValueError("block/indices mismatch")
        keep = ~self.received[indices]
        indices, block = indices[keep], block[:, keep, :]
        if len(indices) == 0:
            return self.n
        block = preprocess_apply(block, self.gaussian, self.tilt)        # recon.py:357
        if self.kind == "parallel":
            filt = np.transpose(filter_rows(block, self.pixel_u, self.window), (1, 0, 2))
        else:
            K = {k: v[indices] for k, v in self.K_all.items()}
            filt = fdk_filter_block(block, K, self.window)
        if self.store_f32:
            filt = filt.astype(np.float32).astype(np.float64)
        n0, n1 = self.n, self.n + len(indices)
        self.q[n0:n1] = filt
        self.order[n0:n1] = indices
        self.received[indices] = True
        self.n = n1
        return self.n

    def reconstruct_points(self, pts, cor_shift=None):
        """recon.py:405-424."""
        pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
        n = self.n
        shift = self.cor_shift if cor_shift is None else float(cor_shift)
        if n == 0 or len(pts) == 0:
            return np.zeros(len(pts))
        scale = np.pi / n
        order = self.order[:n]
        q = self.q[:n]
        if self.kind == "parallel":
            th = self.angles[order]
            return bp_parallel(q, np.cos(th), np.sin(th), self.pixel_u, self.pixel_v,
                               pts, scale, shift)
        K = {k: v[order] for k, v in self.K_all.items()}
        D = K['d'] - shift * K['u'] if shift else K['d']
        return bp_cone(q, K['s'], D, K['uu'], K['vv'], K['n'], K['sdd'], K['sod'], pts, scale)

    def reconstruct_plane(self, origin, normal, n=256, size=None, zoom=1.0, center=None,
                          cor_shift=None):
        """recon.py:426-440."""
        pts, mask = slice_points(self.box, origin, normal, n, size, zoom, center)
        flat = pts.reshape(-1, 3)
        img = np.zeros(len(flat))
        m = mask.ravel()
        if m.any():
            img[m] = self.reconstruct_points(flat[m], cor_shift)
        return img.reshape(n, n), pts

    def reconstruct_volume(self, voxel=None, maxdim=None, native_pixel=None):
        """recon.py:442-470 (single slab; slabs do not change values)."""
        if maxdim is not None:
            dims, origin, h = volume_grid(self.box, maxdim)
        else:
            h = float(voxel or native_pixel)
            dims = np.maximum(np.ceil(2.0 * self.box / h).astype(int), 2)
            origin = -self.box + h / 2.0
        nx, ny, nz = (int(d) for d in dims)
        x = origin[0] + h * np.arange(nx)
        y = origin[1] + h * np.arange(ny)
        z = origin[2] + h * np.arange(nz)
        X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        vol = self.reconstruct_points(pts).reshape(nx, ny, nz)
        return vol, np.asarray(origin, float), h

    def cor_scores(self, shifts, n=192, z=0.0):
        """Scores of recon.py:498-513 (without storing the best shift)."""
        return np.array([histogram_score(self.reconstruct_plane(
            (0.0, 0.0, z), (0.0, 0.0, 1.0), n=n, cor_shift=float(s))[0]) for s in shifts])


# --------------------------------------------------------------------------
# Stream order (stream.py:12-29) and preprocessing (preprocess.py:8-64)
# --------------------------------------------------------------------------
def arrival_order(n, mode="sequential"):
    if mode == "sequential":
        return np.arange(n)
    if mode == "interleaved":
        bits = max(1, int(np.ceil(np.log2(n))))
        idx = np.arange(2 ** bits)
        rev = np.zeros_like(idx)
        for b in range(bits):
            rev |= ((idx >> b) & 1) << (bits - 1 - b)
        return rev[rev < n]
    raise ValueError(f"unknown order {mode!r}")


def gaussian_blur(block, sigma):
    """Edge-padded separable FFT Gaussian, float64 (preprocess.py:8-28)."""
    if sigma <= 0:
        return np.asarray(block, np.float64)
    out = np.asarray(block, np.float64)
    pad = int(np.ceil(3 * sigma))
    for axis in (0, 2):
        n = out.shape[axis]
        widths = [(0, 0)] * 3
        widths[axis] = (pad, pad)
        padded = np.pad(out, widths, mode="edge")
        N = padded.shape[axis]
        f = np.fft.rfftfreq(N)
        H = np.exp(-2.0 * (np.pi * f * sigma) ** 2)
        shape = [1, 1, 1]
        shape[axis] = len(H)
        F = np.fft.rfft(padded, axis=axis) * H.reshape(shape)
        res = np.fft.irfft(F, n=N, axis=axis)
        out = np.take(res, np.arange(pad, pad + n), axis=axis)
    return out


def rotate_projections(block, degrees):
    """Bilinear inverse-map rotation about the detector centre, float64 (preprocess.py:31-56)."""
    block = np.asarray(block, np.float64)
    if abs(degrees) < 1e-9:
        return block
    R, A, C = block.shape
    t = np.deg2rad(degrees)
    rr, cc = np.meshgrid(np.arange(R) - (R - 1) / 2.0, np.arange(C) - (C - 1) / 2.0,
                         indexing="ij")
    rs = np.cos(t) * rr + np.sin(t) * cc + (R - 1) / 2.0
    cs = -np.sin(t) * rr + np.cos(t) * cc + (C - 1) / 2.0
    out = np.zeros((A, R, C))
    for k in range(A):
        out[k] = bilinear(block[:, k, :], rs, cs)
    return np.transpose(out, (1, 0, 2))


def preprocess_apply(block, gaussian=0.0, tilt=0.0):
    """preprocess.py:59-64."""
    block = np.asarray(block, np.float64)
    if tilt:
        block = rotate_projections(block, -tilt)
    if gaussian:
        block = gaussian_blur(block, gaussian)
    return block


# --------------------------------------------------------------------------
# Raw-count correction of the loaders (flexray.py:111-116, 174-184; hdf5.py:24-31, 77-78)
# --------------------------------------------------------------------------
def bin_frames(im, sample):
    """float32 means of sample x sample blocks of the last two axes, trailing rows and
    columns dropped (flexray.py:111-115 _read_binned, hdf5.py:24-31 _bin)."""
    im = np.asarray(im).astype(np.float32)
    if sample <= 1:
        return im
    *lead, h, w = im.shape
    h2, w2 = (h // sample) * sample, (w // sample) * sample
    im = im[..., :h2, :w2].reshape(*lead, h2 // sample, sample, w2 // sample, sample)
    return im.mean(axis=(-1, -3))


def raw_correct(frames, dark, flat, style, sample, flip, crop):
    """uint16 frames [k, H, W] -> line integrals [R, k, C] float32 in the loaders'
    float32 arithmetic: flexray.py:174-184 corrected() (gain = 1/max(flat-dark, 1),
    multiply, clip, log, negate) or hdf5.py:77-78 (divide by max(flat-dark, 1e-6));
    flexray's np.flipud (flexray.py:116) when flip; crop (r0, r1, c0, c1) as
    flexray.py:201."""
    dark = np.asarray(dark, np.float32)
    flat = np.asarray(flat, np.float32)
    r0, r1, c0, c1 = crop
    out = []
    for f in frames:
        im = bin_frames(f, sample)
        if flip:
            im = np.flipud(im)
        if style == "flexray":
            gain = 1.0 / np.maximum(flat - dark, 1.0)
            im = (im - dark) * gain
            np.clip(im, 1e-3, None, out=im)
            np.log(im, out=im)
            np.negative(im, out=im)
        else:
            trans = (im - dark) / np.maximum(flat - dark, 1e-6)
            im = -np.log(np.clip(trans, 1e-3, None)).astype(np.float32)
        out.append(im[r0:r1, c0:c1])
    return np.ascontiguousarray(np.stack(out, axis=1))


# --------------------------------------------------------------------------
# Volume threshold (model.py:18-29)
# --------------------------------------------------------------------------
def otsu_threshold(vol, bins=512):
    """Otsu's threshold on the non-zero voxels, restating model.py:18-29 with numpy."""
    v = vol[vol != 0]
    lo, hi = np.percentile(v, [0.1, 99.95])
    hist, edges = np.histogram(np.clip(v, lo, hi), bins=bins, range=(lo, hi))
    centers = 0.5 * (edges[1:] + edges[:-1])
    w0 = np.cumsum(hist)
    w1 = w0[-1] - w0
    m0 = np.cumsum(hist * centers) / np.maximum(w0, 1)
    m1 = (np.sum(hist * centers) - np.cumsum(hist * centers)) / np.maximum(w1, 1)
    between = w0 * w1 * (m0 - m1) ** 2
    return float(centers[int(np.argmax(between))])


def percentile_from_sorted(s, q):
    """np.percentile(x, 100 q) (method "linear") of float32 data from its sorted copy s:
    virtual index (n - 1) q, gamma its fractional part, the neighbours' difference in
    float32, the interpolation in float64 from b when gamma >= 0.5."""
    n = len(s)
    vi = (n - 1) * float(q)
    if vi >= n - 1:
        i = j = n - 1
    else:
        i = int(np.floor(vi))
        j = i + 1
    g = vi - np.floor(vi)
    a, b = s[i], s[j]
    d = np.float32(b - a)
    return (float(b) - float(d) * (1.0 - g)) if g >= 0.5 else (float(a) + float(d) * g)


# --------------------------------------------------------------------------
# Analytic ellipsoid phantom (phantom.py:25-98), generalised to 3-D rotations
# --------------------------------------------------------------------------
def line_integrals(centers, mats, vals, starts, dirs):
    """Exact chord-length integrals of constant ellipsoids (phantom.py:56-70).
    mats[k] maps (x - centers[k]) to the unit sphere."""
    starts = np.asarray(starts, np.float64)
    dirs = np.asarray(dirs, np.float64)
    out = np.zeros(np.broadcast_shapes(starts.shape, dirs.shape)[:-1])
    for v, c, M in zip(vals, centers, mats):
        p = (starts - c) @ M.T
        q = dirs @ M.T
        qq = np.einsum('...i,...i->...', q, q)
        pq = np.einsum('...i,...i->...', p, q)
        pp = np.einsum('...i,...i->...', p, p)
        disc = pq * pq - qq * (pp - 1.0)
        out += v * 2.0 * np.sqrt(np.clip(disc, 0.0, None)) / qq
    return out


def evaluate(centers, mats, vals, points):
    """Phantom value at points (phantom.py:47-54)."""
    pts = np.asarray(points, np.float64)
    out = np.zeros(pts.shape[:-1])
    for v, c, M in zip(vals, centers, mats):
        y = (pts - c) @ M.T
        out += v * (np.einsum('...i,...i->...', y, y) <= 1.0)
    return out


def project_cone(centers, mats, vals, vectors, R, C):
    """[R, A, C] float64 cone projections (phantom.py:86-98)."""
    cc = np.arange(C) - (C - 1) / 2.0
    rr = np.arange(R) - (R - 1) / 2.0
    out = np.empty((R, len(vectors), C))
    for k, vec in enumerate(np.asarray(vectors, np.float64)):
        s, d, u, v = vec[0:3], vec[3:6], vec[6:9], vec[9:12]
        pix = d + cc[None, :, None] * u + rr[:, None, None] * v
        dirs = pix - s
        dirs /= np.linalg.norm(dirs, axis=-1, keepdims=True)
        out[:, k, :] = line_integrals(centers, mats, vals, s[None, None, :], dirs)
    return out


def project_parallel(centers, mats, vals, angles, R, C, pu, pv, far):
    """[R, A, C] float64 parallel projections (phantom.py:73-83); `far` = start distance."""
    u = (np.arange(C) - (C - 1) / 2.0) * pu
    v = (np.arange(R) - (R - 1) / 2.0) * pv
    out = np.empty((R, len(angles), C))
    for k, th in enumerate(angles):
        d = np.array([np.cos(th), np.sin(th), 0.0])
        eu = np.array([-np.sin(th), np.cos(th), 0.0])
        ez = np.array([0.0, 0.0, 1.0])
        starts = u[None, :, None] * eu + v[:, None, None] * ez - far * d
        out[:, k, :] = line_integrals(centers, mats, vals, starts, d)
    return out


def circular_cone_vectors(n_angles, sod, sdd, pixel, angular_range=2 * np.pi, det_shift_u=0.0):
    """cone_vec rows of an ideal circular orbit (geometry.py:60-72)."""
    angles = np.arange(n_angles) * angular_range / n_angles
    vecs = np.zeros((n_angles, 12))
    for k, a in enumerate(angles):
        ray = np.array([np.cos(a), np.sin(a), 0.0])
        tan = np.array([-np.sin(a), np.cos(a), 0.0])
        vecs[k, 0:3] = -sod * ray
        vecs[k, 3:6] = (sdd - sod) * ray + det_shift_u * pixel * tan
        vecs[k, 6:9] = pixel * tan
        vecs[k, 9:12] = np.array([0.0, 0.0, pixel])
    return vecs
