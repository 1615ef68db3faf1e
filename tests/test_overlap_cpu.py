"""The identities overlapped rounds rely on, in IEEE binary32 round-to-nearest
(the arithmetic of the GPU's v_add_f32 with denormals kept, which the kernels
are compiled for): a colour buffer filled with -0.0 receives one path colour c
per pixel, so it holds -0.0 + c, which is c bit for bit; a pixel no path
touched stays -0.0, and adding -0.0 to the accumulation leaves it unchanged
bit for bit; so accum + (-0.0 + c) has the bits of accum + c.  Checked over
zeros, subnormals, normals across the exponent range and infinities."""
import numpy as np


def _values():
    rng = np.random.default_rng(7)
    special = np.array([0.0, -0.0, np.inf, -np.inf, 1.0, -1.0, np.finfo(np.float32).max,
                        -np.finfo(np.float32).max, np.finfo(np.float32).tiny, -np.finfo(np.float32).tiny],
                       np.float32)
    sub = (rng.integers(1, 1 << 23, 2000, dtype=np.uint32)).view(np.float32)           # positive subnormals
    bits = rng.integers(0, 0x7f800000, 20000, dtype=np.uint32).view(np.float32)       # positive finite
    vals = np.concatenate([special, sub, -sub, bits, -bits]).astype(np.float32)
    return vals


def _bits(x):
    return np.asarray(x, np.float32).view(np.uint32)


def test_negative_zero_fill_copies_the_colour_bits():
    c = _values()
    got = np.float32(-0.0) + c
    assert np.array_equal(_bits(got), _bits(c))


def test_untouched_pixel_adds_nothing():
    a = _values()
    got = a + np.float32(-0.0)
    assert np.array_equal(_bits(got), _bits(a))


def test_merge_equals_direct_accumulation():
    rng = np.random.default_rng(11)
    a = _values()
    c = rng.permutation(_values())
    with np.errstate(invalid="ignore", over="ignore"):
        direct = a + c
        merged = a + (np.float32(-0.0) + c)
    same = _bits(direct) == _bits(merged)
    nan = np.isnan(direct) & np.isnan(merged)   # inf + -inf: both NaN
    assert np.all(same | nan)


def test_positive_zero_fill_would_not_copy_negative_zero():
    """Why -0.0 and not +0.0: +0.0 + (-0.0) is +0.0, so a +0.0 fill would
    change a -0.0 colour's bits."""
    assert _bits(np.float32(0.0) + np.float32(-0.0)) != _bits(np.float32(-0.0))
    assert _bits(np.float32(-0.0) + np.float32(-0.0)) == _bits(np.float32(-0.0))
