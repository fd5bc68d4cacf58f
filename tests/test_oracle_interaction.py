"""The channel algebra of the local and non-local interaction against the reference's own
expectations (tests/test_interaction.py:93-108 for U, :166-200 for V)."""
import numpy as np
import pytest

from oracle import dga_oracle as O

RNG = np.random.default_rng(0)
U_LOC = RNG.random((2, 2, 2, 2))


def _adcb(u):
    return np.einsum("abcd->adcb", u)


@pytest.mark.parametrize("channel,expected", [
    ("dens", 2 * U_LOC - _adcb(U_LOC)),
    ("magn", -_adcb(U_LOC)),
    ("sing", U_LOC + _adcb(U_LOC)),
    ("trip", U_LOC - _adcb(U_LOC)),
])
def test_local_interaction_channels(channel, expected):
    assert np.allclose(O.u_as_channel(U_LOC, channel), expected, rtol=1e-12, atol=0)


@pytest.mark.parametrize("n", [1, 2, 3])
@pytest.mark.parametrize("channel,factor", [("dens", 2.0), ("magn", 0.0), ("sing", 1.0), ("trip", 1.0)])
def test_nonlocal_interaction_channels(n, channel, factor):
    v = RNG.random((4, n, n, n, n))
    assert np.allclose(O.v_as_channel(v, channel), factor * v, rtol=1e-12, atol=0)


def test_product_channels_match_oracle():
    """The product's host-side channel algebra (the ladder's U_r) equals the oracle's."""
    from dgamore_amd.pipeline import u_as_channel

    for ch in ("dens", "magn", "sing", "trip"):
        assert np.abs(u_as_channel(U_LOC, ch) - O.u_as_channel(U_LOC, ch)).max() == 0.0
