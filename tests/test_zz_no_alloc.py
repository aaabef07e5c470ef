"""The allocation-free batch snapshot (lg_batch_snapshot_noalloc) gives what
lg_batch_snapshot gives, bit for bit, through the C ABI (gpu): two calls in a
row, the second on a smaller batch, and a batch whose fields exceed its
static staging array is refused with an error."""
import numpy as np
import pytest

from lgca_amd import Batch, GpuLattice, LgcaError, load_library
import workloads as W

pytestmark = pytest.mark.gpu


def member(w, h, seed, bands, block):
    g = GpuLattice(w, h, True, seed, bands)
    g.set_inlet(8, 0.25, (0.2, 0.0))
    g.add_solid_box(*W.plate_box(w, h, w / 5.0, max(4, h // 6)))
    g.init_equilibrium(0.25, 0.0, 0.0)
    g.attach_field(block, block)
    return g


def twins(specs, steps):
    """Two identical sets of lattices, each advanced `steps` steps."""
    out = []
    for _ in range(2):
        ms = [member(*s) for s in specs]
        for g in ms:
            g.advance(steps)
        out.append(ms)
    return out


def same(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for u, v in zip(x, y):
            assert u.dtype == v.dtype == np.float32
            assert np.array_equal(u.view(np.uint32), v.view(np.uint32))


def test_noalloc_snapshot_equals_the_allocating_one():
    big = [(640, 512, 1, 4, 16), (1024, 320, 2, 16, 32), (256, 128, 3, 3, 8)]
    small = [(192, 64, 4, 1, 16), (128, 64, 5, 2, 8)]
    print(f"noalloc snapshot: {len(big)} members then {len(small)}", flush=True)
    # first call: the larger batch
    ref, dut = twins(big, 30)
    ba, bb = Batch(ref), Batch(dut)
    same(ba.snapshot(0.5), bb.snapshot_noalloc(0.5))
    # the field blends into its running average: a second pair of calls on
    # the same batches agrees too
    ba.advance(7)
    bb.advance(7)
    same(ba.snapshot(0.25), bb.snapshot_noalloc(0.25))
    ba.close()
    bb.close()
    # second call: a smaller batch, reusing the static array
    ref2, dut2 = twins(small, 11)
    ba, bb = Batch(ref2), Batch(dut2)
    same(ba.snapshot(1.0), bb.snapshot_noalloc(1.0))
    ba.close()
    bb.close()
    for g in ref + dut + ref2 + dut2:
        g.close()


def test_noalloc_snapshot_refuses_fields_beyond_its_array():
    cap = int(load_library().lg_batch_snapshot_noalloc_max())
    assert cap == 2 << 20
    # 2048 x 2048 in 2 x 2 blocks: 1024 x 1024 blocks, 4 floats each > cap
    print("noalloc snapshot: oversized field", flush=True)
    g = GpuLattice(2048, 2048, True, 9, 1)
    g.init_equilibrium(0.2, 0.0, 0.0)
    g.attach_field(2, 2)
    b = Batch([g])
    with pytest.raises(LgcaError, match="exceed"):
        b.snapshot_noalloc(1.0)
    # the allocating entry point still serves it, and the batch stays usable
    out = b.snapshot(1.0)
    assert out[0][0].size == 1024 * 1024
    b.close()
    g.close()
