"""cfd2d_flux_noalloc (no device allocation: the rows are staged through the
library's fixed device array) against cfd2d_flux (staging allocated per
call), through the C ABI: the same kernel on the same rows, so the same
bits; two calls in a row, the second smaller, so a stale tail or counter of
the first would show; the Newton-cap counter starts at zero on every call;
more rows than the array holds are refused with an error."""
import math

import numpy as np
import pytest

from cfd2d import meshgen, native

pytestmark = pytest.mark.gpu


def rows(n, seed):
    i = np.arange(n)
    u = [0.5 * (meshgen.uniform(seed, 80 + k, i) + 1.0) for k in range(9)]
    ang = 2 * math.pi * u[8]
    return np.stack([0.05 + 2 * u[0], 0.05 + 2 * u[1], 3 * (u[2] - 0.5), 3 * (u[3] - 0.5),
                     0.05 + 2 * u[4], 0.05 + 2 * u[5], 3 * (u[6] - 0.5), 3 * (u[7] - 0.5),
                     np.cos(ang), np.sin(ang)], axis=1)


@pytest.mark.parametrize("kind", ["GODUNOV", "LAX", "HLLC", "ROE"])
def test_noalloc_equals_allocating_entry(kind):
    big = rows(native.NOALLOC_MAX_ROWS, 1)  # the largest call allowed
    small = rows(777, 2)
    # the first call leaves the array full of the big call's rows and results
    # (five of the big call's rows, strong near-vacuum rarefactions, run the
    # exact solver's Newton loop to its cap, as the oracle's does)
    for r in (big, small):
        F, h = native.flux(kind, r, max_newton=500)
        G, g = native.flux(kind, r, max_newton=500, noalloc=True)
        assert np.array_equal(F, G) and h == g
        assert np.all(np.isfinite(G))
        if kind != "GODUNOV" or r is small:
            assert h == 0


def test_noalloc_counter_restarts_and_size_limit():
    bad = np.array([[-0.1, -0.2, 0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 1.0, 0.0]] * 3)  # never meets the stop test
    _, h = native.flux("GODUNOV", bad, max_newton=30, noalloc=True)
    assert h == 3
    F, h = native.flux("GODUNOV", rows(50, 3), max_newton=30, noalloc=True)
    assert h == 0
    with pytest.raises(native.Cfd2dError) as e:
        native.flux("LAX", rows(native.NOALLOC_MAX_ROWS + 1, 4), noalloc=True)
    assert e.value.code == 1 and "rows per call" in str(e.value)
    # still usable after the refusal
    F2, _ = native.flux("GODUNOV", rows(50, 3), max_newton=30)
    G2, _ = native.flux("GODUNOV", rows(50, 3), max_newton=30, noalloc=True)
    assert np.array_equal(F2, G2)
