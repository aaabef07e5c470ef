"""The oracle's transport::measure against the reference's documented values
(docs/transport.md:14-15: density 0.22 with rest particles, 256 x 256, seed
0x5EED, 12 CPU threads: nu 0.2989, g 0.4405)."""
import pytest


@pytest.mark.slow
def test_measure_reproduces_the_documented_coefficients(oracle):
    nu = oracle.viscosity(0.22, True, 12, 0x5EED, 256)
    g = oracle.advection_factor(0.22, True, 12, 0x5EED, 256)
    assert f"{nu:.4f}" == "0.2989"
    assert f"{g:.4f}" == "0.4405"


def test_advection_factor_follows_the_fhp_prediction(oracle):
    """g ~ (7/12)(1-2d)/(1-d) within ten percent (docs/transport.md:35-40),
    on a small box so it stays quick."""
    d = 0.22
    g = oracle.advection_factor(d, True, 4, 0x5EED, 128)
    pred = 7.0 / 12.0 * (1 - 2 * d) / (1 - d)
    assert abs(g - pred) / pred < 0.15
