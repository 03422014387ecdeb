"""Ring copies as caches: kernels 6 and 7 read a row-pair and a band copy of the ring,
built on first use.  Under a budget (or without the device memory) a request runs without
a copy it has no room for, or frees the copy it does not read first; results stay those
of the unconstrained context to float32 rounding (relative L2 <= 1e-5 between two kernels,
each a float32 evaluation of the same sums; equality where the same kernel runs again),
and captured graphs (COR scoring, streaming updates) are captured again after a copy they
read was freed."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu

A, R, C = 96, 64, 96
XTOL = 1e-5


@pytest.fixture(scope="module")
def setup():
    from minirecast import _hip, recon
    from minirecast.geometry import circular_cone_vectors
    g = circular_cone_vectors(A, 400.0, 600.0, R, C, 0.5)
    box = g.volume_box()
    rng = np.random.default_rng(12)
    block = rng.uniform(0.0, 1.0, size=(R, A, C)).astype(np.float32)
    slices, off = [], 0
    for k, nrm in enumerate(((1.0, 0.5, 0.8), (-0.3, 1.0, 0.4), (0.0, 0.0, 1.0))):
        slices.append(recon.plane_task(box, (0.0, 0.0, 0.0), nrm, 64, cor_shift=0.5 * k - 0.5,
                                       out_offset=off))
        off += 64 * 64
    slices = np.array(slices, dtype=_hip.TASK_DTYPE)
    dims, origin, h = recon.volume_grid(box, 32)
    vol = np.array([recon.volume_task(origin, h, dims, 0, int(dims[2]))], dtype=_hip.TASK_DTYPE)
    nvox = int(np.prod(dims))
    pitch = (C + 3 + 63) // 64 * 64
    sizes = dict(band=2 * ((R + 1) // 3 + 1) * pitch * 4 * A, pair=2 * 2 * (R + 2) * pitch * A)

    def make(budget=None):
        ctx = _hip.Context(0)
        ctx.set_cone(g.vectors, R, C, "shepp-logan", "f16")
        ctx.set_ring_copy_budget(budget)
        ctx.push(np.arange(A), block)
        return ctx

    ref = make()
    want = dict(slices=ref.bp_tasks(slices, A, off), vol=ref.bp_tasks(vol, A, nvox))
    return dict(g=g, block=block, slices=slices, out_len=off, vol=vol, nvox=nvox, sizes=sizes,
                make=make, ref=ref, want=want)


def _held(ctx):
    b = ctx.device_bytes()
    return b["band_copy"], b["pair_copy"]


def test_unconstrained_context_keeps_both_copies(setup):
    s = setup
    ctx = s["ref"]
    band, pair = _held(ctx)
    assert band == s["sizes"]["band"] and pair == s["sizes"]["pair"]
    assert ctx.device_bytes()["ring"] == A * (R + 3) * ((C + 3 + 63) // 64 * 64) * 2


def test_no_budget_for_copies_runs_on_the_ring(setup):
    s = setup
    print("[memory] budget 0", flush=True)
    ctx = s["make"](0)
    got = ctx.bp_tasks(s["slices"], A, s["out_len"])
    assert _held(ctx) == (0, 0)
    assert rel_l2(got, s["want"]["slices"]) < XTOL
    got = ctx.bp_tasks(s["vol"], A, s["nvox"])
    assert _held(ctx) == (0, 0)
    assert rel_l2(got, s["want"]["vol"]) < XTOL
    # the work-queue loop of a plan on the ring itself
    from minirecast import _hip
    p = _hip.Plan(ctx, s["slices"], s["out_len"])
    pref = _hip.Plan(s["ref"], s["slices"], s["out_len"])
    assert rel_l2(p.run(), pref.run()) < XTOL
    np.testing.assert_allclose(p.cor_scores(), pref.cor_scores(), rtol=1e-5, atol=1e-9)
    assert _held(ctx) == (0, 0)
    # a budget set later frees what exceeds it
    big = s["make"]()
    big.bp_tasks(s["slices"], A, s["out_len"])
    big.bp_tasks(s["vol"], A, s["nvox"])
    assert _held(big) == (s["sizes"]["band"], s["sizes"]["pair"])
    big.set_ring_copy_budget(s["sizes"]["band"])
    assert _held(big) == (s["sizes"]["band"], 0)
    big.set_ring_copy_budget(0)
    assert _held(big) == (0, 0)


def test_one_copy_budget_swaps_copies_and_recaptures_graphs(setup):
    s = setup
    from minirecast import _hip
    print("[memory] budget of one copy", flush=True)
    band, pair = s["sizes"]["band"], s["sizes"]["pair"]
    ctx = s["make"](pair + (1 << 20))
    plan = _hip.Plan(ctx, s["slices"], s["out_len"])
    scores0 = plan.cor_scores()                       # graph captured over the band copy
    assert _held(ctx) == (band, 0)
    got = ctx.bp_tasks(s["vol"], A, s["nvox"])        # frees the band copy for the pair copy
    assert _held(ctx) == (0, pair)
    assert rel_l2(got, s["want"]["vol"]) < XTOL
    scores1 = plan.cor_scores()                       # band copy rebuilt, graph captured again
    assert _held(ctx) == (band, 0)
    np.testing.assert_array_equal(scores1, scores0)
    got = ctx.bp_tasks(s["slices"], A, s["out_len"])
    assert rel_l2(got, s["want"]["slices"]) < XTOL
    # a budget below the pair copy: volumes run on the band copy instead of freeing it
    ctx.set_ring_copy_budget(band + (1 << 20))
    got = ctx.bp_tasks(s["vol"], A, s["nvox"])
    assert _held(ctx) == (band, 0)
    assert rel_l2(got, s["want"]["vol"]) < XTOL


def test_streaming_plan_recaptures_after_a_copy_is_freed(setup):
    s = setup
    from minirecast import _hip
    print("[memory] streaming plan across copy swaps", flush=True)
    g, block = s["g"], s["block"]
    outs = {}
    for name, budget in (("free", None), ("tight", s["sizes"]["pair"] + (1 << 20))):
        ctx = _hip.Context(0)
        ctx.set_cone(g.vectors, R, C, "shepp-logan", "f16")
        ctx.set_ring_copy_budget(budget)
        plan = _hip.Plan(ctx, s["slices"], s["out_len"])
        res = []
        for b in range(4):
            idx = np.arange(b * 24, (b + 1) * 24)
            res.append(plan.stream(idx, np.ascontiguousarray(block[:, idx])).copy())
            if b == 1:
                ctx.bp_tasks(s["vol"], int(ctx.num_views()), s["nvox"])   # swaps the copies
        outs[name] = res
        if name == "tight":
            assert _held(ctx)[1] == 0
    for a, b in zip(outs["free"], outs["tight"]):
        assert rel_l2(b, a) < XTOL
    assert rel_l2(outs["tight"][-1], s["want"]["slices"]) < XTOL
