"""Pins the surrogate oracle (CPU): bfloat16 rounding against torch's conversion and exact ties,
the float64 gradient against central finite differences of the float64 loss."""
import numpy as np
import pytest

import surrogate_mlp as S


def _rows(rng, n):
    x = rng.normal(0.0, 1.0, (n, 17))
    x[:, 16] = rng.uniform(0.0, 4.0, n)  # density along the light
    return x.astype(np.float32)


def test_bf16_rounding_ties_and_torch():
    torch = pytest.importorskip("torch")
    one = 1.0
    assert S.bf16(np.array([one]))[0] == 1.0
    assert S.bf16(np.array([1.0 + 2.0 ** -8]))[0] == 1.0                    # tie -> even (mantissa 0)
    assert S.bf16(np.array([1.0 + 3 * 2.0 ** -8]))[0] == 1.0 + 2.0 ** -6    # tie -> even (mantissa 2)
    assert S.bf16(np.array([-(1.0 + 3 * 2.0 ** -8)]))[0] == -(1.0 + 2.0 ** -6)
    rng = np.random.default_rng(3)
    v = np.concatenate([rng.normal(0, 10, 5000), rng.normal(0, 1e-3, 5000)]).astype(np.float32)
    ref = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(S.bf16(v), ref)


def test_param_layout():
    assert S.N_PARAMS == 9538
    flat = np.arange(S.N_PARAMS, dtype=np.float64)
    p = S.unpack(flat)
    assert p["W1"].shape == (16, 64) and p["W1"][1, 0] == 64  # W[k_in][unit] row-major
    assert p["b4"].shape == (2,) and p["b4"][1] == S.N_PARAMS - 1
    assert np.array_equal(S.pack(p), flat)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(7)
    flat = S.init_params(5)
    stats = np.concatenate([rng.normal(0, 0.3, 16), rng.uniform(0.5, 2.0, 16)]).astype(np.float32)
    rows = _rows(rng, 6)
    g, L = S.gradient(flat, stats, rows)
    assert np.isclose(L, S.loss(flat, stats, rows))
    # |r| is held fixed by the loss's second term: differentiate with it frozen
    o, _ = S.forward(flat, S.normalise(rows, stats))
    abs_r = np.abs(np.exp(o[:, 0]) - np.exp(-rows[:, 16].astype(np.float64)))
    n = len(rows)
    # every bias and a sample of weights from each layer
    off = {}
    o_ = 0
    for name, shape in S.SHAPES:
        off[name] = (o_, int(np.prod(shape)))
        o_ += int(np.prod(shape))
    idx = []
    for name, (a, m) in off.items():
        idx += list(a + rng.choice(m, size=min(m, 12), replace=False))
    h = 1e-6
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        fd = (S.loss_fixed_abs(fp, stats, rows, abs_r) - S.loss_fixed_abs(fm, stats, rows, abs_r)) / (2 * h) / n
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_planted_transpose_is_caught():
    """A transposed hidden weight (a plausible indexing slip) changes the gradient visibly."""
    rng = np.random.default_rng(2)
    flat = S.init_params(9)
    stats = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    rows = _rows(rng, 64)
    p = S.unpack(flat)
    p["W2"] = p["W2"].T.copy()
    g0, _ = S.gradient(flat, stats, rows)
    g1, _ = S.gradient(S.pack(p), stats, rows)
    assert np.max(np.abs(g0 - g1)) > 1e-3 * np.max(np.abs(g0))


def test_bf16_gradient_close_to_float64():
    rng = np.random.default_rng(11)
    flat = S.init_params(3)
    stats = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    rows = _rows(rng, 256)
    g64, _ = S.gradient(flat, stats, rows)
    g16, _ = S.gradient(flat, stats, rows, bf16_ops=True)
    assert np.max(np.abs(g64 - g16)) <= 0.05 * np.max(np.abs(g64))


def test_hard_silu_activation():
    """Hidden units: hard SiLU z * clamp(z / 6 + 1/2, 0, 1) -- zero below -3, the identity above 3,
    z^2 / 6 + z / 2 between -- and its derivative against central differences away from the kinks."""
    z = np.array([-10.0, -3.0, -1.5, 0.0, 1.5, 3.0, 10.0])
    assert np.allclose(S.hswish(z), [0.0, 0.0, -0.375, 0.0, 1.125, 3.0, 10.0])
    zz = np.linspace(-6, 6, 1201)
    zz = zz[np.abs(np.abs(zz) - 3.0) > 1e-3]
    fd = (S.hswish(zz + 1e-6) - S.hswish(zz - 1e-6)) / 2e-6
    assert np.allclose(S.hswish_grad(zz), fd, atol=1e-6)


def test_f16_inference_rounding():
    """The FP16 inference restatement: binary16 rounding against numpy / torch conversions (ties to
    even, overflow), the packed hard SiLU within its three roundings of the exact activation, FP16
    holding BF16 weights exactly, and the whole FP16 forward pass close to the float64 network (a
    transposed weight or a dropped bias fails)."""
    import torch

    one = 1.0
    assert S.f16(np.array([one + 2.0 ** -11]))[0] == 1.0                    # tie -> even
    assert S.f16(np.array([one + 3 * 2.0 ** -11]))[0] == 1.0 + 2.0 ** -9     # tie -> even
    assert S.f16(np.array([70000.0]))[0] == np.inf
    v = np.random.default_rng(3).normal(0, 30, 10000).astype(np.float32)  # FP32 values, as the kernels round
    assert np.array_equal(S.f16(v), torch.from_numpy(v).half().double().numpy())
    z = np.linspace(-8, 8, 4001)
    ex = S.hswish(z)
    assert np.all(np.abs(S.hswish_f16(z) - ex) <= 3 * 2.0 ** -11 * (np.abs(z) + 0.5) + 2.0 ** -24)
    assert S.hswish_f16(np.array([-4.0]))[0] == 0.0 and S.hswish_f16(np.array([4.0]))[0] == 4.0
    rng = np.random.default_rng(11)
    flat = S.init_params(7)
    stats = np.concatenate([np.zeros(16), np.ones(16)]).astype(np.float32)
    rows = rng.uniform(-1, 1, (512, 17)).astype(np.float32)
    ref = S.predict(flat, stats, rows)
    got = S.predict(flat, stats, rows, f16_ops=True)
    scale = np.max(np.abs(ref), 0)
    good = np.max(np.abs(got - ref))
    assert np.all(np.max(np.abs(got - ref), 0) <= 5e-3 * scale)
    # FP16 holds the BF16 weights exactly (magnitudes in [2^-17, 65504]): the inference fragments
    # carry the training pass's weights
    w = np.concatenate([S.unpack(flat)[k].ravel() for k in ("W1", "W2", "W3", "W4")])
    w = w[np.abs(w) >= 2.0 ** -17]
    assert np.array_equal(S.f16(S.bf16(w)), S.bf16(w))
    ws = np.array([2.0 ** -17 * 1.00390625, 3.0e-5, 0.7, -65280.0])  # 8-bit values across the range
    assert np.array_equal(S.f16(S.bf16(ws)), S.bf16(ws))
    p = S.unpack(flat)
    p["W2"] = p["W2"].T.copy()
    bad = S.predict(S.pack(p), stats, rows, f16_ops=True)
    assert np.max(np.abs(bad - ref)) > 0.05 * scale.max()
    p = S.unpack(flat)
    p["b2"] = np.zeros_like(p["b2"])
    bad = S.predict(S.pack(p), stats, rows, f16_ops=True)
    assert np.max(np.abs(bad - ref)) > 3 * good
