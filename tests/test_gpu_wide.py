"""GPU parity of wide detector rows: the ramp filter transforms rows of N = 2^ceil(log2 2n)
points (recon.py:37) for any n up to 131072 columns.  Rows of up to 8192 columns take one
workgroup's LDS (N <= 16384); wider rows take the four-step path (16-point column
transforms around an N/16-point transform pair), in chunks of row pairs and views.

Tolerances as in the other parity tests: relative L2 <= 1e-5 for float32 results against
the float64 oracle, 1e-3 for f16 storage, 4e-3 for bf16 (one bf16 rounding, 2^-8)."""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


# columns -> FFT length: 4097..8192 -> 2^14 (LDS), 8193.. -> 2^15..2^18 (four-step)
@pytest.mark.parametrize("n", [4097, 6000, 8192, 8193, 12000, 16385, 33000, 65537, 131072])
def test_filter_rows_wide_vs_oracle(n):
    from minirecast import recon
    print(f"[wide] filter_rows n={n}", flush=True)
    rng = np.random.default_rng(n)
    x = rng.normal(size=(3, n)).astype(np.float32)
    for window in O.WINDOWS:
        got = recon.filter_rows(x, 0.7, window)
        want = O.filter_rows(x, 0.7, window)
        assert got.shape == x.shape and np.isfinite(got).all()
        assert rel_l2(got, want) < TOL, (n, window, rel_l2(got, want))


def test_filter_rows_wide_pair_chunks_and_odd_rows():
    """259 rows (130 row pairs, the last one single) of 70000 columns (N = 2^18): the
    four-step scratch holds 128 slots of 2 MiB, so the pairs go in two chunks."""
    from minirecast import recon
    print("[wide] filter_rows 259 x 70000", flush=True)
    rng = np.random.default_rng(5)
    x = rng.normal(size=(259, 70000)).astype(np.float32)
    got = recon.filter_rows(x, 1.3, "hann")
    want = O.filter_rows(x, 1.3, "hann")
    assert rel_l2(got, want) < TOL
    for r in (0, 127, 128, 255, 256, 257, 258):      # chunk edges and the single last row
        assert rel_l2(got[r], want[r]) < TOL, r


def test_filter_rows_width_limit():
    from minirecast import recon
    with pytest.raises(ValueError, match="131072"):
        recon.filter_rows(np.zeros((1, 131073), np.float32), 1.0)


def _wide_cone(R, A, C, pixel):
    from minirecast.geometry import circular_cone_vectors
    return circular_cone_vectors(A, 400.0, 600.0, R, C, pixel, det_shift_u=2.5)


@pytest.mark.parametrize("R,A,C", [(9, 4, 6000), (200, 12, 9000)])
def test_wide_cone_ingest_vs_oracle(R, A, C):
    """Stored views of every storage against the float64 oracle.  200 rows x 12 views at
    9000 columns (N = 2^15): 100 row pairs per chunk, views in chunks of 10 and 2."""
    from minirecast import recon
    print(f"[wide] cone ingest R={R} A={A} C={C}", flush=True)
    g = _wide_cone(R, A, C, 0.05)
    rng = np.random.default_rng(C)
    block = rng.uniform(0.0, 2.0, size=(R, A, C)).astype(np.float32)
    block[:, 1, :] *= 40.0                               # views of different scales
    q64 = O.fdk_filter_block(block, O.cone_constants(g.vectors), "shepp-logan")
    imgs = {}
    for storage, tol in (("f32", TOL), ("f16", 1e-3), ("bf16", 4e-3)):
        rec = recon.SliceReconstructor(block, g, verbose=False, storage=storage)
        q = rec.filtered_views()
        assert np.isfinite(q).all()
        for v in range(A):
            assert rel_l2(q[v], q64[v]) < tol, (storage, v, rel_l2(q[v], q64[v]))
        imgs[storage], _ = rec.reconstruct_plane((0, 0, 0), (1, 0.4, 0.7), n=48)
    box = g.volume_box()
    orc = O.OracleReconstructor("cone", R, C, box, vectors=g.vectors, store_f32=True)
    orc.add_projections(np.arange(A), block)
    ref, _ = orc.reconstruct_plane((0, 0, 0), (1, 0.4, 0.7), n=48)
    assert rel_l2(imgs["f32"], ref) < TOL
    assert rel_l2(imgs["f16"], ref) < 1e-3
    assert rel_l2(imgs["bf16"], ref) < 4e-3


@pytest.mark.parametrize("C", [6000, 9000])
def test_wide_filter_adjoint(C):
    """Adjoint of the cosine weight + filter at wide rows: dot-product test and the float64
    adjoint."""
    from minirecast import _hip
    print(f"[wide] filter adjoint C={C}", flush=True)
    R, A = 7, 3
    g = _wide_cone(R, A, C, 0.05)
    ctx = _hip.Context(0)
    ctx.set_cone(g.vectors, R, C, "shepp-logan")
    rng = np.random.default_rng(C + 1)
    x = rng.normal(size=(R, A, C)).astype(np.float32)
    ctx.push(np.arange(A), x)
    fx = ctx.read_views(0, A, R, C)
    z = rng.normal(size=(A, R, C)).astype(np.float32)
    ftz = ctx.filter_adjoint(np.arange(A), z)
    lhs = float(np.sum(fx.astype(np.float64) * z))
    rhs = float(np.sum(x.astype(np.float64) * ftz))
    assert abs(lhs - rhs) <= 1e-5 * np.linalg.norm(fx) * np.linalg.norm(z)
    K = O.cone_constants(g.vectors)
    assert rel_l2(ftz, O.fdk_filter_adjoint(z, K, "shepp-logan")) < TOL
    assert rel_l2(fx, O.fdk_filter_block(x, K, "shepp-logan")) < TOL
