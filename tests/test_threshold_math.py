"""The closed form the staging kernels use for the square-root step of a confocal threshold
(bpx_kernels.hip, stable_threshold): for d in [2^-60, 2^60], the smallest float s with
sqrt_rn(s) >= d is m^2 rounded up to a float, m the midpoint between d and the float below it
(one float higher when m^2 is a float and the tie rounds to the float below d).  Checked here
against numpy's correctly rounded float32 square root, on random and power-of-two distances."""
import numpy as np


def closed_form(d):
    d = np.asarray(d, np.float32)
    below = (d.view(np.uint32) - 1).view(np.float32)
    m = 0.5 * (d.astype(np.float64) + below.astype(np.float64))
    msq = m * m
    s = msq.astype(np.float32)
    up = s.astype(np.float64) < msq
    s[up] = (s[up].view(np.uint32) + 1).view(np.float32)
    tie = (s.astype(np.float64) == msq) & ((d.view(np.uint32) & 1) == 1)
    s[tie] = (s[tie].view(np.uint32) + 1).view(np.float32)
    return s


def test_sqrt_threshold_closed_form():
    rng = np.random.default_rng(0)
    d = np.concatenate([rng.uniform(1e-3, 1e3, 100000), rng.uniform(0.3, 3.0, 100000),
                        2.0 ** rng.integers(-60, 60, 4000), [2.0 ** -60, 2.0 ** 60, 1.0]]).astype(np.float32)
    s = closed_form(d)
    prev = (s.view(np.uint32) - 1).view(np.float32)
    assert np.all(np.sqrt(s) >= d)            # s reaches d ...
    assert np.all(np.sqrt(prev) < d)          # ... and the float below it does not


def test_sqrt_threshold_closed_form_detects_a_wrong_rounding():
    """Rounding m^2 to nearest instead of up breaks the threshold for some distances."""
    rng = np.random.default_rng(1)
    d = rng.uniform(0.3, 3.0, 20000).astype(np.float32)
    below = (d.view(np.uint32) - 1).view(np.float32)
    m = 0.5 * (d.astype(np.float64) + below.astype(np.float64))
    s = (m * m).astype(np.float32)          # nearest, not up
    assert not np.all(np.sqrt(s) >= d)
