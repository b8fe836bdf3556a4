"""The GPU tests' parity criteria themselves (tests/parity.py), on the host."""
import numpy as np
import pytest

from tests import parity


def test_clamp_kinks_marks_pixels_next_to_a_clamp_bound():
    img = np.full((2, 3, 4, 5), 0.3)
    img[1, 2, 0, 1] = 1e-6 * (1 + 1e-6)        # next to the lower bound
    img[0, 0, 3, 4] = 1.0 - 1e-6                # next to the upper bound
    img[0, 1, 2, 2] = 1e-6 * (1 + 1e-4)         # not next to it
    img[1, 0, 1, 1] = 1.0 + 1e-3                # saturated, not a kink
    m = parity.clamp_kinks(img)
    assert m.shape == (4, 5)
    assert sorted(zip(*np.nonzero(m))) == [(0, 1), (3, 4)]


def test_check_gradient_leaves_out_kink_pixels_only_up_to_a_fraction():
    rng = np.random.default_rng(0)
    g64 = rng.standard_normal((9, 200, 200))
    g32 = g64 * (1 + 1e-7)
    g = g64.copy()
    g[3, 10, 10] += 100.0                       # a jump at one pixel
    with pytest.raises(AssertionError):
        parity.check_gradient(g, g32, g64)
    ex = np.zeros((200, 200), bool)
    ex[10, 10] = True
    parity.check_gradient(g, g32, g64, exclude=ex)
    ex[:, :5] = True                            # 1000 pixels: beyond 1e-4 of 40,000
    with pytest.raises(AssertionError, match="kink pixels"):
        parity.check_gradient(g, g32, g64, exclude=ex)
