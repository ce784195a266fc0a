"""CPU oracle of the learned light-cone surrogate (test infrastructure only: imported by tests/,
the smoke check and the CPU leg of the benchmark, never by the product path).

The surrogate has no counterpart in the reference: it stands in for the six-sample light cone of
compute-clouds.comp:440-453 (density along the sun, D, then Beer's law). This module restates the
network the HIP kernels implement -- a 16-64-64-64-2 MLP with hard-SiLU units, outputs (o0, o1) with
o0 ~ -D and softplus(o1) = e an error bound in transmittance units -- its training loss and the
exact gradient of that loss, in float64.

Loss per row (T = exp(-D), p = exp(o0), r = p - T):  r^2 + (softplus(o1) - |r|)^2, with |r| held
fixed in the second term (the error head regresses the size of the current residual).

``f16_ops=True`` in predict restates the FP16 inference kernels instead (forward_f16).

``bf16=True`` rounds, like the MFMA training kernels, the operands of every matrix product to bfloat16
(round to nearest even): the weights, the normalised inputs, the hidden activations, and the
back-propagated dZ; sums and element-wise math stay in float64.

Pinning: the float64 gradient is checked against central finite differences of the float64 loss
(tests/test_oracle_surrogate.py); the bf16 rounding against numpy's float32 -> bfloat16 bit
pattern rule on hand-picked ties.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 16, 64, 2
SHAPES = [("W1", (N_IN, N_HID)), ("b1", (N_HID,)), ("W2", (N_HID, N_HID)), ("b2", (N_HID,)),
          ("W3", (N_HID, N_HID)), ("b3", (N_HID,)), ("W4", (N_HID, N_OUT)), ("b4", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 9538


def unpack(flat: np.ndarray) -> dict:
    """Split the flat parameter vector (device layout, W[k_in][unit_out] row-major) into arrays."""
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = np.asarray(flat[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1) for name, _ in SHAPES])


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest, ties to even) via float32, returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _q(x, on):
    return bf16(x) if on else x


def hswish(z):
    """Hidden-unit activation: hard SiLU (hard-swish), z * clamp(z / 6 + 1/2, 0, 1)."""
    return z * np.clip(z / 6.0 + 0.5, 0.0, 1.0)


def hswish_grad(z):
    return np.where(z <= -3.0, 0.0, np.where(z >= 3.0, 1.0, z / 3.0 + 0.5))


def softplus(x):
    return np.where(x > 20.0, x, np.log1p(np.exp(np.minimum(x, 20.0))))


def normalise(rows: np.ndarray, stats: np.ndarray) -> np.ndarray:
    """(x - mean) * inv_std in float32, as the kernels compute it."""
    x = np.asarray(rows, np.float32)[:, :N_IN]
    m = np.asarray(stats[:16], np.float32)
    s = np.asarray(stats[16:32], np.float32)
    return ((x - m) * s).astype(np.float64)


def forward(flat: np.ndarray, xn: np.ndarray, bf16_ops: bool = False):
    """Outputs (n, 2) and the pre-activations of the three hidden layers."""
    p = unpack(flat)
    a = _q(xn, bf16_ops)
    zs = []
    for l in (1, 2, 3):
        z = a @ _q(p[f"W{l}"], bf16_ops) + p[f"b{l}"]
        zs.append(z)
        a = _q(hswish(z), bf16_ops)
    o = a @ _q(p["W4"], bf16_ops) + p["b4"]
    return o, zs


def f16(x: np.ndarray) -> np.ndarray:
    """Round to IEEE binary16 (nearest, ties to even), returned as float64; beyond the FP16 range the
    result is ±inf, as the hardware's conversion gives."""
    with np.errstate(over="ignore"):
        return np.asarray(x, np.float64).astype(np.float16).astype(np.float64)


def hswish_f16(z):
    """The inference kernels' hard SiLU in packed FP16: v = FP16(z), u = clamp(FP16(fma(v, FP16(1/6), 1/2)),
    0, 1), FP16(v u). Each FP16 operation is exact in float64 before its one rounding."""
    v = f16(z)
    u = np.clip(f16(v * f16(1.0 / 6.0) + 0.5), 0.0, 1.0)
    return f16(v * u)


def forward_f16(flat: np.ndarray, xn: np.ndarray) -> np.ndarray:
    """Outputs (n, 2) of the FP16 inference path (the march's light-cone batches and
    mm_surrogate_predict): normalised inputs rounded to FP16, the BF16 weights of the training pass
    (held in FP16), each hidden bias as the sum of its FP16 rounding and the FP16 rounding of the
    remainder, FP32-accumulated sums (here float64), activations by hswish_f16, output biases added
    in FP32."""
    p = unpack(flat)
    a = f16(xn)
    for l in (1, 2, 3):
        b = np.asarray(p[f"b{l}"], np.float32).astype(np.float64)
        hi = f16(b)
        z = a @ f16(bf16(p[f"W{l}"])) + (hi + f16(b - hi))
        a = hswish_f16(z)
    return a @ f16(bf16(p["W4"])) + p["b4"]


def predict(flat, stats, rows, bf16_ops: bool = False, f16_ops: bool = False) -> np.ndarray:
    """(o0, softplus(o1)) per row, like mm_surrogate_predict (f16_ops: its FP16 inference rounding)."""
    xn = normalise(rows, stats)
    o = forward_f16(flat, xn) if f16_ops else forward(flat, xn, bf16_ops)[0]
    return np.stack([o[:, 0], softplus(o[:, 1])], 1)


def loss(flat, stats, rows, bf16_ops: bool = False) -> float:
    """Summed per-row loss."""
    o, _ = forward(flat, normalise(rows, stats), bf16_ops)
    T = np.exp(-np.asarray(rows, np.float64)[:, N_IN])
    r = np.exp(o[:, 0]) - T
    e = softplus(o[:, 1])
    return float(np.sum(r * r + (e - np.abs(r)) ** 2))


def loss_fixed_abs(flat, stats, rows, abs_r: np.ndarray) -> float:
    """The loss with |r| replaced by given constants (what the gradient differentiates)."""
    o, _ = forward(flat, normalise(rows, stats))
    T = np.exp(-np.asarray(rows, np.float64)[:, N_IN])
    r = np.exp(o[:, 0]) - T
    e = softplus(o[:, 1])
    return float(np.sum(r * r + (e - abs_r) ** 2))


def gradient(flat, stats, rows, bf16_ops: bool = False):
    """Mean gradient over the rows (flat vector) and the summed loss, like mm_surrogate_gradient."""
    p = unpack(flat)
    n = len(rows)
    xn = normalise(rows, stats)
    o, zs = forward(flat, xn, bf16_ops)
    T = np.exp(-np.asarray(rows, np.float64)[:, N_IN])
    pe = np.exp(o[:, 0])
    r = pe - T
    ar = np.abs(r)
    e = softplus(o[:, 1])
    sig = 1.0 / (1.0 + np.exp(-o[:, 1]))
    L = float(np.sum(r * r + (e - ar) ** 2))
    dout = np.stack([2.0 * r * pe, 2.0 * (e - ar) * sig], 1) / n
    q = lambda v: _q(v, bf16_ops)
    acts = [q(xn)] + [q(hswish(z)) for z in zs]  # inputs of layers 1..4
    g = {}
    dz = q(dout)
    g["W4"] = acts[3].T @ dz
    g["b4"] = dz.sum(0)
    da = dz @ q(p["W4"]).T
    for l in (3, 2, 1):
        dz = q(da * hswish_grad(zs[l - 1]))
        g[f"W{l}"] = acts[l - 1].T @ dz
        g[f"b{l}"] = dz.sum(0)
        if l > 1:
            da = dz @ q(p[f"W{l}"]).T
    return pack(g), L


def init_params(seed: int) -> np.ndarray:
    """Glorot-uniform weights and small random biases (a test initialisation, not the device's)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            lim = np.sqrt(6.0 / (shape[0] + shape[1]))
            p[name] = rng.uniform(-lim, lim, shape)
        else:
            p[name] = rng.uniform(-0.1, 0.1, shape)
    return pack(p).astype(np.float32).astype(np.float64)
