"""The flag-free dataflow sweep marks not-yet-written X words with a
sentinel and treats any other bit pattern as a final value.  That is safe
only if no arithmetic result can carry the sentinel's bits: IEEE 754
arithmetic never returns a signalling NaN (an operation on one returns a
quiet NaN), so the sentinel must be a signalling NaN.  These checks read
the constant from the kernel source and pin that property with numpy's
IEEE arithmetic (CPU only)."""
import os
import re

import numpy as np

SRC = os.path.join(os.path.dirname(__file__), "..", "flowpsi_amd", "csrc", "kernels.hip")


def _sentinel():
    m = re.search(r"X_SENT\s*=\s*0x([0-9A-Fa-f]+)ull", open(SRC).read())
    assert m, "sentinel constant not found"
    return int(m.group(1), 16)


def _bits(a):
    return np.asarray(a, np.float64).view(np.uint64)


def test_sentinel_is_a_signalling_nan():
    s = _sentinel()
    exp = (s >> 52) & 0x7FF
    quiet = (s >> 51) & 1
    frac = s & ((1 << 52) - 1)
    assert exp == 0x7FF and frac != 0, "not a NaN"
    assert quiet == 0, "a quiet NaN could be produced by arithmetic"


def test_arithmetic_never_returns_the_sentinel():
    s = _sentinel()
    x = np.array([s], np.uint64).view(np.float64)
    ones = np.ones(1)
    with np.errstate(invalid="ignore"):
        results = [x + 1.0, x * 2.0, x - x, x / 3.0, np.sqrt(x), -x * ones, x * 0.0, x * 2.0 + 1.0, 1.0 / x]
    for r in results:
        b = int(_bits(r)[0])
        assert np.isnan(r).all()
        assert b != s and (b >> 51) & 1 == 1  # quieted
    # and ordinary values and the quiet NaNs arithmetic makes differ from it
    with np.errstate(invalid="ignore", divide="ignore"):
        made = np.array([np.nan, np.inf - np.inf, np.sqrt(-1.0), 0.0, -0.0, 1e308 * 10])
    assert not np.any(_bits(made) == np.uint64(s))
