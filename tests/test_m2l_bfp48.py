"""The 48-bit block floating-point M2L storage's rounding rule, on the CPU:
bfp48_round (the test suite's statement of what the GPU conversion must
produce) is pinned by what the format means, independently of how it is
computed: integer mantissas that fit 48-bit two's complement, the smallest
exponent that allows that, each entry the nearest grid point, and an error
of at most half a grid step (2^-46 of the block's largest entry)."""
import numpy as np

from conftest import bfp48_round


def blocks():
    rng = np.random.default_rng(11)
    for _ in range(300):
        r, c = rng.integers(1, 60, size=2)
        scale = 10.0 ** rng.uniform(-20, 20)
        a = rng.standard_normal((r, c)) * scale
        if rng.uniform() < 0.3:  # wide dynamic range inside the block
            a *= 10.0 ** rng.uniform(-12, 0, size=(r, c))
        yield a
    yield np.zeros((3, 4))
    yield np.array([[1.0, -1.0], [0.5, 2.0 ** -60]])
    yield np.array([[np.nextafter(1.0, 0.0)]])  # max just below a power of two


def test_bfp48_round_is_the_format():
    for a in blocks():
        r, s = bfp48_round(a)
        mx = np.abs(a).max()
        if mx == 0.0:
            np.testing.assert_array_equal(r, 0.0)
            continue
        m = np.ldexp(r, -s)
        np.testing.assert_array_equal(m, np.rint(m))  # integer mantissas
        assert np.abs(m).max() <= 2.0 ** 46  # fit 48-bit two's complement
        assert mx * 2.0 ** -s >= 2.0 ** 45  # the smallest such exponent
        q = 2.0 ** s
        err = np.abs(a - r)
        assert np.all(err <= q / 2) and np.all(err <= 2.0 ** -46 * mx)
        # nearest grid point: no neighbour closer (a - r and a - r -+ q are
        # exact here, both operands within a few grid steps)
        assert np.all(err <= np.abs(a - (r - q)))
        assert np.all(err <= np.abs(a - (r + q)))


def test_bfp48_round_planted_mistakes_fail():
    """A truncating or one-exponent-too-large variant breaks the pins
    above: this is what they guard."""
    rng = np.random.default_rng(3)
    a = rng.standard_normal((40, 40))
    r, s = bfp48_round(a)
    trunc = np.ldexp(np.trunc(np.ldexp(a, -s)), s)
    assert np.any(np.abs(a - trunc) > np.abs(a - r))
    coarse = np.ldexp(np.rint(np.ldexp(a, -(s + 1))), s + 1)
    assert np.abs(a).max() * 2.0 ** -(s + 1) < 2.0 ** 45
    assert np.abs(a - coarse).max() > 2.0 ** (s - 1)


def test_bfp48_storage_rejects_bad_arguments(engine):
    L = engine.lib()
    assert L.afmm2d_set_m2l_storage(None, 1) == engine.ERR_INVALID_ARGUMENT
