"""Batched stepping (tile table + hipGraph replay) equals every member's own
run, bit for bit (gpu)."""
import numpy as np
import pytest

from lgca_amd import Batch, GpuLattice
import workloads as W

pytestmark = pytest.mark.gpu


def member(w, h, rest, seed, bands, inlet, plate):
    g = GpuLattice(w, h, rest, seed, bands)
    g.set_inlet(inlet, 0.25, (0.2, 0.0))
    if plate:
        g.add_solid_box(*W.plate_box(w, h, w / 5.0, max(4, h // 6)))
    g.init_equilibrium(0.25, 0.0, 0.0)
    return g


SPECS = [  # uneven sizes, ragged tails, bands, inflow on/off, plates
    (256, 128, True, 1, 3, 8, True),
    (192, 64, True, 2, 1, 0, False),
    (100, 36, False, 3, 5, 8, True),
    (1024, 320, True, 4, 16, 8, True),
    (65, 10, True, 5, 2, 70, False),
    (640, 512, True, 6, 7, 8, True),
]


def build_all():
    return [member(*s) for s in SPECS]


# (kernels per step, collide cells per lane, draw cells per lane)
PIPELINES = [(2, 64, 64), (2, 32, 32), (2, 64, 32), (2, 32, 64), (1, 64, 64), (1, 32, 64)]


@pytest.mark.parametrize("pipeline", PIPELINES)
@pytest.mark.parametrize("shape", [None, (128, 0, 8), (256, 1024, 2), (128, 4096, 16)])
def test_batch_equals_solo_runs(shape, pipeline):
    print(f"batch shape={shape} pipeline={pipeline}", flush=True)
    solo = build_all()
    for g in solo:
        g.advance(45)
    batched = build_all()
    b = Batch(batched)
    b.set_passes(pipeline[0])
    b.set_units(pipeline[1])
    b.set_draw_unit(pipeline[2])
    assert b.pipeline() == pipeline
    if shape:
        b.set_shape(*shape)
    b.advance(20)          # graph replays + remainder
    b.advance(25)
    for s, g in zip(solo, batched):
        assert g.steps == s.steps == 45
        assert np.array_equal(g.store(), s.store())


@pytest.mark.parametrize("passes,waves", [(1, 2), (1, 3), (1, 4), (2, 4), (2, 5), (2, 6), (1, 6)])
def test_register_targets_keep_results_exact(passes, waves):
    """Every waves-per-SIMD instance of the batch step kernels (and a target
    the pipeline has no instance for, which runs its default) gives the
    members' own runs."""
    print(f"batch waves passes={passes} waves={waves}", flush=True)
    solo = build_all()
    for g in solo:
        g.advance(21)
    batched = build_all()
    b = Batch(batched)
    b.set_passes(passes)
    b.set_units(64)
    b.set_draw_unit(64)
    b.set_waves(waves)
    default = 3 if passes == 1 else 5
    assert b.waves() == (waves if (passes, waves) in ((1, 2), (1, 3), (1, 4), (2, 4), (2, 6))
                         else default)
    b.set_shape(128, 0, 8)
    b.advance(21)
    for s, g in zip(solo, batched):
        assert np.array_equal(g.store(), s.store())


def test_graph_replays_with_eager_work_between():
    """Each replay checked against eager stepping, with the members' own
    eager calls (store, solo advance) between replays."""
    print("graph replays", flush=True)
    ref = build_all()
    bat = build_all()
    b = Batch(bat)
    b.set_shape(256, 0, 4)
    for rep in range(4):
        b.advance(4)                 # exactly one replay
        for r in ref:
            r.advance(4)
        for r, g in zip(ref, bat):
            assert np.array_equal(g.store(), r.store())
        bat[rep].advance(3)          # a member steps on its own in between
        ref[rep].advance(3)


def test_autotune_keeps_results_exact():
    print("autotune", flush=True)
    ref = build_all()
    bat = build_all()
    b = Batch(bat)
    used = b.autotune(budget_steps=0)  # every candidate: pipelines change within the call
    assert used > 0 and b.shape()[0] in (128, 256)
    b.advance(30)
    for r in ref:
        r.advance(used + 30)
    for r, g in zip(ref, bat):
        assert g.steps == r.steps
        assert np.array_equal(g.store(), r.store())


def test_limited_advance_and_batched_snapshot():
    """Members with fewer steps left stop early inside the same launches; the
    batched snapshot equals each member's own sample + read_field."""
    print("limited advance + snapshot", flush=True)
    ref = build_all()
    bat = build_all()
    for g in ref + bat:
        g.attach_field(16, 16)
    b = Batch(bat)
    limits = [40, 7, 40, 33, 1, 26]
    for rep in range(3):                  # eager first call, then graph replays
        b.advance_limited(40, limits)
        snaps = b.snapshot(1.0)
        for r, g, lim, snap in zip(ref, bat, limits, snaps):
            r.advance(min(40, lim))
            r.sample_field(1.0)
            assert g.steps == r.steps
            assert np.array_equal(g.store(), r.store())
            for x, y in zip(snap, r.read_field()):
                assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_pipeline_changes_between_calls():
    """Fused and two-pass calls interleaved on one batch (each call starts
    its own digit stream) give the members' own runs."""
    print("pipeline changes", flush=True)
    solo = build_all()
    for g in solo:
        g.advance(5 + 8 + 3 + 10)
    bat = build_all()
    b = Batch(bat)
    for passes, n in ((2, 5), (1, 8), (2, 3), (1, 10)):
        b.set_passes(passes)
        b.advance(n)
    for s_, g in zip(solo, bat):
        assert g.steps == s_.steps
        assert np.array_equal(g.store(), s_.store())


def test_destroying_a_member_waits_for_its_batch():
    """lg_destroy of a member only marks it: the batch's calls fail with a
    clear error, and the batch frees it when it is destroyed; the other
    member is untouched and still equals its own run."""
    import ctypes as C
    from lgca_amd import LgcaError, load_library
    L = load_library()
    a, b = member(256, 128, True, 1, 3, 8, True), member(192, 64, True, 2, 1, 0, False)
    ref = member(192, 64, True, 2, 1, 0, False)
    bt = Batch([a, b])
    bt.advance(6)
    raw = a._p                               # still allocated: the batch holds it
    a.close()
    with pytest.raises(LgcaError, match="destroyed"):
        bt.advance(6)
    with pytest.raises(LgcaError, match="destroyed"):
        bt.autotune(8)
    assert not L.lg_batch_create((C.c_void_p * 2)(b._p, raw), 2)
    assert "destroyed" in L.lg_last_error().decode()
    bt.close()                               # frees a
    b.advance(4)
    ref.advance(10)
    assert np.array_equal(b.store(), ref.store())
    with pytest.raises(LgcaError, match="twice"):
        Batch([b, b])
    for g in (b, ref):
        g.close()
