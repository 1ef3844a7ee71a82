"""CPU restatement of the learned closure's network, in numpy FP64.

Test infrastructure only: the tests and the smoke check import it to check
the GPU kernels; nothing in the product calls it.

The network (no reference counterpart; the learned closure is a project
extension): four 5x5 convolutions with zero padding 2 on an ny x nx grid,
channels 4 -> 32 -> 32 -> 32 -> 3 (stored padded as 16 -> 32 -> 32 -> 32 -> 16),
GELU in tanh form after the first three, mean squared error over the three
output channels. Parameters follow the GPU layout: per layer weights
[25 taps][cout][cin] (tap = 5 dy + dx) then bias [cout].

With ``quant=True`` the restatement rounds to bf16 (round to nearest even)
where the GPU stores bf16: weights, inputs, hidden activations,
pre-activations and backpropagated gradients; everything else is FP64. With
``quant=False`` it is plain FP64 and its gradient is checked against finite
differences.
"""
from __future__ import annotations

import numpy as np

TAPS = 25
LAYERS = ((16, 32), (32, 32), (32, 32), (32, 16))   # (cin, cout) as stored
USED_IN, USED_OUT = 4, 3


def layer_offsets() -> list[int]:
    off, out = 0, []
    for cin, cout in LAYERS:
        out.append(off)
        off += TAPS * cin * cout + cout
    out.append(off)
    return out


def param_count() -> int:
    return layer_offsets()[-1]


def unpack(params: np.ndarray) -> list[tuple[np.ndarray, np.ndarray]]:
    offs = layer_offsets()
    out = []
    for l, (cin, cout) in enumerate(LAYERS):
        o = offs[l]
        w = params[o:o + TAPS * cin * cout].reshape(TAPS, cout, cin)
        b = params[o + TAPS * cin * cout:o + TAPS * cin * cout + cout]
        out.append((w.astype(np.float64), b.astype(np.float64)))
    return out


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bf16 (nearest even) via FP32, returned as FP64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


GELU_K = np.sqrt(2.0 / np.pi)
GELU_C = 0.044715


def gelu(x):
    """GELU, tanh form: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))."""
    return 0.5 * x * (1.0 + np.tanh(GELU_K * (x + GELU_C * x ** 3)))


def gelu_grad(x):
    t = np.tanh(GELU_K * (x + GELU_C * x ** 3))
    return 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_K * (1.0 + 3.0 * GELU_C * x * x)


def conv(x: np.ndarray, w: np.ndarray, b: np.ndarray | None) -> np.ndarray:
    """out[y, x, co] = b[co] + sum_{dy, dx, ci} xpad[y+dy, x+dx, ci] w[5dy+dx, co, ci]."""
    ny, nx, _ = x.shape
    xp = np.pad(x, ((2, 2), (2, 2), (0, 0)))
    out = np.zeros((ny, nx, w.shape[1]))
    for tap in range(TAPS):
        dy, dx = divmod(tap, 5)
        out += xp[dy:dy + ny, dx:dx + nx, :] @ w[tap].T
    if b is not None:
        out += b
    return out


def conv_input_grad(d: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Gradient of conv's input: din[y, x, ci] = sum d[y-dy+2, x-dx+2, co] w[tap, co, ci]."""
    ny, nx, _ = d.shape
    dp = np.pad(d, ((2, 2), (2, 2), (0, 0)))
    out = np.zeros((ny, nx, w.shape[2]))
    for tap in range(TAPS):
        dy, dx = divmod(tap, 5)
        out += dp[4 - dy:4 - dy + ny, 4 - dx:4 - dx + nx, :] @ w[tap]
    return out


def conv_weight_grad(d: np.ndarray, x: np.ndarray) -> np.ndarray:
    """dw[tap, co, ci] = sum_{y, x} d[y, x, co] xpad[y+dy, x+dx, ci]."""
    ny, nx, _ = d.shape
    xp = np.pad(x, ((2, 2), (2, 2), (0, 0)))
    g = np.zeros((TAPS, d.shape[2], x.shape[2]))
    dd = d.reshape(-1, d.shape[2])
    for tap in range(TAPS):
        dy, dx = divmod(tap, 5)
        g[tap] = dd.T @ xp[dy:dy + ny, dx:dx + nx, :].reshape(-1, x.shape[2])
    return g


def _q(x, quant):
    return bf16(x) if quant else x


def forward(params: np.ndarray, x4: np.ndarray, quant: bool = True):
    """x4: normalised inputs (ny, nx, 4). Returns (output (ny, nx, 3), cache)."""
    layers = unpack(params)
    ny, nx, _ = x4.shape
    h = np.zeros((ny, nx, LAYERS[0][0]))
    h[..., :USED_IN] = x4
    h = _q(h, quant)
    cache = {"h": [h], "pre": []}
    for l in range(3):
        w, b = layers[l]
        pre = conv(h, _q(w, quant), b)
        # the GPU takes gelu of the accumulated pre-activation and stores it
        # in bf16; the pre-activation kept for the backward pass is bf16 too
        cache["pre"].append(_q(pre, quant))
        h = _q(gelu(pre), quant)
        cache["h"].append(h)
    w, b = layers[3]
    out = conv(h, _q(w, quant), b)[..., :USED_OUT]
    return out, cache


def loss_and_grad(params: np.ndarray, x4: np.ndarray, target: np.ndarray,
                  quant: bool = True, weights=(1.0, 1.0, 1.0)) -> tuple[float, np.ndarray]:
    """Weighted mean squared error over the 3 output channels,
    sum_c w_c sum_p d_c^2 / (3 n), and its gradient in the parameter layout."""
    layers = unpack(params)
    out, cache = forward(params, x4, quant)
    ny, nx, _ = out.shape
    n = ny * nx
    diff = out - target
    w = np.asarray(weights, np.float64)
    loss = float(np.sum(w * diff * diff) / (3 * n))
    d = np.zeros((ny, nx, LAYERS[3][1]))
    d[..., :USED_OUT] = 2.0 * w * diff / (3 * n)
    d = _q(d, quant)
    grads = np.zeros(param_count())
    offs = layer_offsets()
    for l in range(3, -1, -1):
        cin, cout = LAYERS[l]
        w, _ = layers[l]
        gw = conv_weight_grad(d, cache["h"][l])
        gb = d.reshape(-1, cout).sum(axis=0)
        o = offs[l]
        grads[o:o + TAPS * cin * cout] = gw.ravel()
        grads[o + TAPS * cin * cout:o + TAPS * cin * cout + cout] = gb
        if l > 0:
            din = conv_input_grad(d, _q(w, quant))
            d = _q(din * gelu_grad(cache["pre"][l - 1]), quant)
    return loss, grads
