"""transport::measure on the engine equals the CPU path's, bit for bit, at the
same thread count (gpu)."""
import numpy as np
import pytest

import lgca_amd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("density,rest,threads,seed,size", [
    (0.22, True, 4, 0x5EED, 64),
    (0.30, False, 3, 7, 96),
])
def test_small_box_measure_matches_the_oracle(oracle, density, rest, threads, seed, size):
    print(f"measure d={density} rest={rest} threads={threads} size={size}", flush=True)
    nu, g = lgca_amd.measure_transport(density, rest, threads, seed, size)
    assert np.float32(nu) == np.float32(oracle.viscosity(density, rest, threads, seed, size))
    assert np.float32(g) == np.float32(oracle.advection_factor(density, rest, threads, seed, size))


def test_reference_measure_on_the_engine(oracle):
    """The reference's own --measure configuration: 0.2989 / 0.4405
    (docs/transport.md:14-15), bit-identical with the oracle."""
    print("measure 256x256 threads=12", flush=True)
    nu, g = lgca_amd.measure_transport(0.22, True, 12, 0x5EED, 256)
    assert f"{nu:.4f}" == "0.2989" and f"{g:.4f}" == "0.4405"
    assert np.float32(nu) == np.float32(oracle.viscosity(0.22, True, 12, 0x5EED, 256))
    assert np.float32(g) == np.float32(oracle.advection_factor(0.22, True, 12, 0x5EED, 256))
