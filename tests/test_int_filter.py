"""The march filters 8-bit channels with 1/256 sub-texel weights in integer arithmetic: per level a
16-bit dot product of a texel pair with (256 - fk, fk), then one conversion, RN(N) * 2^-24. That is
bit-identical to the float chain lerp(a, b, t) = fma(t, b - a, a) of the restatement because the first
two levels are exact in float32 and the third rounds once. This checks the identity on the CPU (FMA
emulated in float64, where every product and sum here is exact), and that it would fail for weights
that are not multiples of 1/256."""
import numpy as np


def _fma32(t, d, a):
    # float32 inputs: t * d + a is exact in float64 whenever the float32 FMA result needs <= 53 bits
    return (t.astype(np.float64) * d.astype(np.float64) + a.astype(np.float64)).astype(np.float32)


def _lerp32(a, b, t):
    return _fma32(t, (b - a).astype(np.float32), a)


def _float_chain(tex, fx, fy, fz):
    """filt3 as the restatement computes it: tex (n, 8) bytes in sample order x, then y, then z."""
    c = tex.astype(np.float32)
    a0, b0 = _lerp32(c[:, 0], c[:, 1], fx), _lerp32(c[:, 2], c[:, 3], fx)
    a1, b1 = _lerp32(c[:, 4], c[:, 5], fx), _lerp32(c[:, 6], c[:, 7], fx)
    return _lerp32(_lerp32(a0, b0, fy), _lerp32(a1, b1, fy), fz)


def _int_chain(tex, kx, ky, kz):
    t = tex.astype(np.int64)
    l1 = [t[:, 2 * i] * (256 - kx) + t[:, 2 * i + 1] * kx for i in range(4)]
    m0 = l1[0] * (256 - ky) + l1[1] * ky
    m1 = l1[2] * (256 - ky) + l1[3] * ky
    n = m0 * (256 - kz) + m1 * kz
    assert n.max() < 2**32 and max(m0.max(), m1.max()) < 2**24 and max(x.max() for x in l1) < 2**16
    return n.astype(np.float32) * np.float32(2.0**-24)


def test_integer_trilinear_equals_the_float_chain():
    rng = np.random.default_rng(5)
    n = 400_000
    tex = rng.integers(0, 256, (n, 8))
    tex[: n // 8] = rng.choice([0, 255], (n // 8, 8))  # extremes
    k = rng.integers(0, 256, (3, n))
    f = [(kk.astype(np.float32) * np.float32(1 / 256)) for kk in k]
    want = _float_chain(tex, *f)
    got = _int_chain(tex, *k)
    assert np.array_equal(got, want)
    # the scaled forms the kernel returns: RN(RN(N) * (s * 2^-24)) = RN(RN(N / 2^24) * s)
    s = np.float32(1 / 255)
    assert np.array_equal((got * s).astype(np.float32), (want * s).astype(np.float32))


def test_integer_bilinear_is_exact():
    rng = np.random.default_rng(6)
    tex = rng.integers(0, 256, (200_000, 4))
    kx, ky = rng.integers(0, 256, (2, 200_000))
    fx, fy = kx.astype(np.float32) / 256, ky.astype(np.float32) / 256
    c = tex.astype(np.float32)
    want = _lerp32(_lerp32(c[:, 0], c[:, 1], fx), _lerp32(c[:, 2], c[:, 3], fx), fy)
    m = (tex[:, 0] * (256 - kx) + tex[:, 1] * kx) * (256 - ky) + (tex[:, 2] * (256 - kx) + tex[:, 3] * kx) * ky
    assert np.array_equal(m.astype(np.float32) * np.float32(2.0**-16), want)


def test_identity_needs_the_1_256_weights():
    """With unrounded float weights the third level no longer reduces to one integer sum: the check
    above must be able to fail."""
    rng = np.random.default_rng(7)
    tex = rng.integers(0, 256, (50_000, 8))
    f = rng.random((3, 50_000)).astype(np.float32)
    k = np.floor(f * 256).astype(np.int64)
    assert not np.array_equal(_int_chain(tex, *k), _float_chain(tex, *f))
