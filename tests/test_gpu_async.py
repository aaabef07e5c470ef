"""Non-blocking advance (lg_advance_async, lg_batch_advance_async): the same
cells and fields as the blocking calls, step counters that count the queued
steps, and later calls queued behind the work (gpu)."""
import numpy as np
import pytest

from lgca_amd import Batch, GpuLattice
import workloads as W

pytestmark = pytest.mark.gpu


def lattice(w, h, seed, bands, inlet=8, plate=True):
    g = GpuLattice(w, h, True, seed, bands)
    g.set_inlet(inlet, 0.25, (0.2, 0.0))
    if plate:
        g.add_solid_box(*W.plate_box(w, h, w / 5.0, max(4, h // 6)))
    g.init_equilibrium(0.25, 0.0, 0.0)
    return g


def same_field(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a.read_field(), b.read_field()))


# (kernels per step, draw-ahead steps): two-pass, fused, two-pass drawn ahead
@pytest.mark.parametrize("passes,ahead", [(2, 0), (1, 0), (2, 4)])
def test_async_advance_equals_advance(passes, ahead):
    a, b = lattice(512, 256, 3, 5), lattice(512, 256, 3, 5)
    for g in (a, b):
        g.set_passes(passes)
        g.set_draw_ahead(ahead)
        g.attach_field(16, 16)
    a.advance_sampling(300, 100, 0.5)
    b.advance_async(100, 100, 0.5)           # two calls queued back to back
    b.advance_async(200, 100, 0.5)
    assert b.steps == 300                    # counted when queued
    cells = b.store()                        # waits for the queued work
    assert b.poll()
    assert np.array_equal(cells, a.store())
    assert same_field(a, b)
    b.advance_async(7)
    b.wait()
    assert b.poll()
    a.advance(7)
    assert np.array_equal(b.store(), a.store()) and b.total_particles() == a.total_particles()
    for g in (a, b):
        g.close()


def test_batch_async_advance_orders_member_calls():
    specs = [(256, 128, 1, 3), (192, 64, 2, 1), (100, 36, 3, 5)]
    solo = [lattice(*s) for s in specs]
    mem = [lattice(*s) for s in specs]
    for g in solo + mem:
        g.attach_field(4, 4)
    bt = Batch(mem)
    limits = [40, 25, 40]
    bt.advance_async(40, limits)
    assert [g.steps for g in mem] == limits  # counted when queued
    # a member's own call queues behind the batch's steps
    mem[1].advance_async(5)
    for g, lim in zip(solo, limits):
        g.advance(lim)
    solo[1].advance(5)
    bt.advance_async(10)                     # waits for the members' own work first
    bt.wait()
    assert bt.poll() and all(g.poll() for g in mem)
    for g in solo:
        g.advance(10)
    for s, m in zip(solo, mem):
        assert s.steps == m.steps
        assert np.array_equal(s.store(), m.store())
    snaps = bt.snapshot(1.0)
    for s, snap in zip(solo, snaps):
        s.sample_field(1.0)
        assert all(np.array_equal(x, y) for x, y in zip(s.read_field(), snap))
    bt.close()
    for g in solo + mem:
        g.close()


def test_batch_sampling_advance_equals_advance_then_snapshot():
    from lgca_amd import LgcaError
    specs = [(256, 128, 1, 3), (192, 64, 2, 1), (100, 36, 3, 5), (640, 512, 6, 7)]
    a, b = [lattice(*s) for s in specs], [lattice(*s) for s in specs]
    for g in a[:3] + b[:3]:                  # the last member has no field
        g.attach_field(8, 8)
    ba, bb = Batch(a), Batch(b)
    limits = [30, 12, 30, 30]                # member 1 stops early in every call
    for _ in range(3):
        ba.advance_limited(30, limits)
        sa = ba.snapshot(0.5)
        bb.advance_sampling_async(30, limits, 0.5)
        sb = bb.read_snapshot()
        assert sa[3] is None and sb[3] is None
        for x, y in zip(sa[:3], sb[:3]):
            assert all(np.array_equal(u, v) for u, v in zip(x, y))
    for x, y in zip(a, b):
        assert x.steps == y.steps and np.array_equal(x.store(), y.store())
    bb.advance_sampling_async(0, None, 1.0)  # the sample alone
    sb, sa = bb.read_snapshot(), ba.snapshot(1.0)
    for x, y in zip(sa[:3], sb[:3]):
        assert all(np.array_equal(u, v) for u, v in zip(x, y))
    bb.advance_sampling_async(2, None, 1.0)
    bb.advance(1)                            # another batch call discards the sample
    with pytest.raises(LgcaError, match="sampling advance"):
        bb.read_snapshot()
    for bt in (ba, bb):
        bt.close()
    for g in a + b:
        g.close()


def test_two_sampling_advances_queued_before_reading():
    """Two sampling advances queued back to back (and a plain async advance
    between them) are read back oldest first, each equal to an advance and a
    snapshot taken in turn; a third unread one is refused."""
    from lgca_amd import LgcaError
    specs = [(256, 128, 1, 3), (192, 64, 2, 1), (640, 512, 6, 7)]
    a, b = [lattice(*s) for s in specs], [lattice(*s) for s in specs]
    for g in a + b:
        g.attach_field(8, 8)
    ba, bb = Batch(a), Batch(b)
    limits = [30, 12, 30]
    want = []
    for n, lim in ((30, limits), (20, None)):
        ba.advance_limited(n, lim if lim is not None else [n] * 3)
        want.append(ba.snapshot(1.0))
    ba.advance(5)
    bb.advance_sampling_async(30, limits, 1.0)
    bb.advance_sampling_async(20, None, 1.0)
    with pytest.raises(LgcaError, match="two snapshots"):
        bb.advance_sampling_async(1, None, 1.0)
    bb.advance_async(5)                      # keeps the unread snapshots
    for w in want:
        got = bb.read_snapshot()
        for x, y in zip(w, got):
            assert all(np.array_equal(u, v) for u, v in zip(x, y))
    with pytest.raises(LgcaError, match="unread snapshot"):
        bb.read_snapshot()
    for x, y in zip(a, b):
        assert x.steps == y.steps and np.array_equal(x.store(), y.store())
    for bt in (ba, bb):
        bt.close()
    for g in a + b:
        g.close()


def test_quantum_on_another_batch_queued_before_the_last_ends():
    """A serving loop's order: a quantum queued on batch A, the next queued at
    once on batch B of another membership that shares members with A (B's
    queue waits for A on the device), then both snapshots read; every
    member, and every snapshot, equals its own run's."""
    specs = [(256, 128, 1, 3), (192, 64, 2, 1), (100, 36, 3, 5), (384, 256, 4, 2)]
    solo, mem = [lattice(*s) for s in specs], [lattice(*s) for s in specs]
    for g in solo + mem:
        g.attach_field(4, 4)
    for rep in range(2):
        ba = Batch(mem[:3])
        bb = Batch(mem[1:])
        ba.advance_sampling_async(40, [40, 17, 40], 1.0)
        bb.advance_sampling_async(40, None, 1.0)   # members 1, 2 follow A's steps
        sa, sb = ba.read_snapshot(), bb.read_snapshot()
        for g, lim in zip(solo[:3], (40, 17, 40)):
            g.advance(lim)
        snaps_a = []
        for g in solo[:3]:
            g.sample_field(1.0)
            snaps_a.append(g.read_field())
        for g in solo[1:]:
            g.advance(40)
        snaps_b = []
        for g in solo[1:]:
            g.sample_field(1.0)
            snaps_b.append(g.read_field())
        for want, got in zip(snaps_a + snaps_b, list(sa) + list(sb)):
            assert all(np.array_equal(u, v) for u, v in zip(want, got)), f"round {rep}"
        ba.close()
        bb.close()
        for s, m in zip(solo, mem):
            assert s.steps == m.steps and np.array_equal(s.store(), m.store()), f"round {rep}"
    for g in solo + mem:
        g.close()


def test_destroy_with_async_work_queued():
    """Destroying a batch or a lattice right after an async advance (no wait)
    lets the queued steps finish first: the members hold the advanced state."""
    specs = [(256, 128, 1, 3), (100, 36, 3, 5)]
    solo, mem = [lattice(*s) for s in specs], [lattice(*s) for s in specs]
    bt = Batch(mem)
    bt.advance_async(20)
    bt.close()                               # no wait
    for s, m in zip(solo, mem):
        s.advance(20)
        assert m.poll() and m.steps == 20
        assert np.array_equal(s.store(), m.store())
    g = lattice(512, 256, 9, 4)
    g.advance_async(50)
    g.close()                                # queued steps end before the free
    for x in solo + mem:
        x.close()
