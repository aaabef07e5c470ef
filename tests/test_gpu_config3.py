"""The transport-measure workload (periodic shear wave, no obstacle): the
engine against the oracle (gpu)."""
import numpy as np
import pytest

from lgca_amd import GpuLattice
import workloads as W
from oracle.workloads_cpu import config_oracle

pytestmark = pytest.mark.gpu


def column_mode(uy_cols, w):
    """The column_uy projection onto the first mode (src/transport.rs:253-262)."""
    k = np.float32(np.float32(2.0) * np.float32(np.pi) / np.float32(w))
    p = (k * np.arange(w, dtype=np.float32)).astype(np.float32)
    a = float(np.sum((uy_cols * np.cos(p)).astype(np.float64)))
    b = float(np.sum((uy_cols * np.sin(p)).astype(np.float64)))
    return 2 * a / w, 2 * b / w


def test_reduced_config3_every_step_and_the_decay(oracle):
    """1024 x 512 with the config's statistics: cells bit-exact every 50
    steps for 600 steps, and the decaying mode equal on both paths."""
    cfg = W.config3(bands=8, w=1024, h=512)
    print(f"config3 reduced {cfg['w']}x{cfg['h']}", flush=True)
    lat = config_oracle(oracle, cfg)
    g = W.config_gpu(cfg)
    assert np.array_equal(g.store(), lat.cells)
    amps = []
    for _ in range(12):
        lat.advance(50)
        g.advance(50)
        cells = g.store()
        assert np.array_equal(cells, lat.cells)
        amps.append(column_mode(oracle.column_uy(cells, cfg["w"], cfg["h"]), cfg["w"]))
    a0 = np.hypot(*amps[0])
    assert a0 > 0.01 and np.hypot(*amps[-1]) < a0     # the wave decays


def test_full_config3_prefix_and_conservation(oracle):
    """32768 x 16384 (537M cells): device init equals the CPU's, two steps
    equal the CPU replay, and mass and momentum stay exact over 200 steps."""
    cfg = W.config3(bands=16)
    print("config3 full size", flush=True)
    g = W.config_gpu(cfg)
    lat = config_oracle(oracle, cfg)
    lat.set_parallel(True)
    c0 = g.store()
    assert np.array_equal(c0, lat.cells)
    inv0 = W.exact_invariants(c0)
    lat.advance(2)
    g.advance(2)
    assert np.array_equal(g.store(), lat.cells)
    del lat
    g.advance(198)
    assert W.exact_invariants(g.store()) == inv0
    assert g.total_particles() == inv0[0]
