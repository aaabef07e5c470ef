"""Stream tables kept across calls (gpu): a lattice builds the tables of the
next run of steps once and later calls whose steps fall inside take them, so
runs of short calls that cross the end of a built run, calls on another
pipeline, a batch that rewrites the tables in between and a restart with a new
seed must all step the same cells as the oracle."""
import numpy as np
import pytest

from lgca_amd import Batch, GpuLattice

pytestmark = pytest.mark.gpu

# call lengths crossing the 256-step runs of tables a 16-band lattice builds
CALLS = [1, 1, 2, 3, 1, 5, 250, 7, 1, 2, 1, 300, 1, 2]


def pair(oracle, w, h, inlet, seed, plate=True):
    lat = oracle.Lattice(w, h, 0.3, (0.2, 0.0), inlet, True, 16, seed)
    if plate:
        lat.add_solid_box(w / 4, w / 4 + 3, h / 3, 2 * h / 3)
    lat.init_equilibrium()
    g = GpuLattice(w, h, True, seed, 16)
    g.set_inlet(inlet, 0.3, (0.2, 0.0))
    g.load(lat.cells)
    return lat, g


def check(lat, g, what):
    got, ref = g.store(), lat.cells
    assert g.steps == lat.steps, what
    if not np.array_equal(got, ref):
        bad = np.nonzero(got != ref)[0]
        pytest.fail(f"{what}: {bad.size} cells differ, first at index {int(bad[0])}")


# (kernels per step, collide cells per lane, draw-ahead steps); None: the size rule
@pytest.mark.parametrize("pipeline", [None, (1, 64, 0), (1, 32, 0), (2, 64, 0), (2, 32, 4)])
@pytest.mark.parametrize("inlet", [0, 8])
def test_short_calls_across_table_runs(oracle, pipeline, inlet):
    print(f"tables: pipeline {pipeline} inlet {inlet}", flush=True)
    lat, g = pair(oracle, 200, 64, inlet, 0x7AB1E5 + inlet)
    if pipeline:
        passes, units, ahead = pipeline
        g.set_passes(passes)
        g.set_work_unit(units)
        g.set_draw_ahead(ahead)
    for i, k in enumerate(CALLS):
        lat.advance(k)
        if i % 3 == 2:
            g.advance_async(k)
        else:
            g.advance(k)
        if i in (5, 6, 11, len(CALLS) - 1):
            check(lat, g, f"after call {i} ({k} steps)")
    g.close()


def test_pipeline_changes_batch_and_restart_between_calls(oracle):
    """Calls on different pipelines share the built tables; a batch rewrites
    them (the member's next call builds its own again); a restart with a new
    seed drops them."""
    lat, g = pair(oracle, 256, 96, 0, 0xC0FFEE)
    seq = [(None, 1), ((2, 64, 0), 3), ((1, 32, 0), 2), (None, 40), ((2, 64, 4), 9), (None, 1)]
    for pipe, k in seq:
        if pipe:
            g.set_passes(pipe[0])
            g.set_work_unit(pipe[1])
            g.set_draw_ahead(pipe[2])
        else:
            g.set_passes(0)
        g.advance(k)
        lat.advance(k)
    check(lat, g, "mixed pipelines")
    b = Batch([g])
    b.advance(20)
    b.close()
    lat.advance(20)
    g.set_passes(0)
    for k in (1, 2, 1):
        g.advance(k)
        lat.advance(k)
    check(lat, g, "after a batch")
    cells = lat.cells
    lat.restart(cells, 0xBEEF)
    g.restart(cells, 0xBEEF)
    for k in (1, 1, 30):
        g.advance(k)
        lat.advance(k)
    check(lat, g, "after a restart")
    # resume at an earlier step inside the built run: the oracle cannot start
    # at step 5, so a fresh handle set to step 5 is the reference here
    cells5 = g.store()
    g.steps = 5
    g.advance(2)
    g2 = GpuLattice(256, 96, True, 0xBEEF, 16)
    g2.restart(cells5, 0xBEEF)
    g2.steps = 5
    g2.set_passes(2)
    g2.advance(2)
    assert np.array_equal(g.store(), g2.store())
    for x in (g, g2):
        x.close()
