"""CPU oracle of the subgrid-closure network and the box filter (test
infrastructure only: imported by tests/ and nothing else).

The reference has no learned closure and no filter; these restatements are
pinned by mathematics instead of reference vectors:
  * conv3d: a unit impulse input returns the (flipped) kernel, linearity,
    and agreement with a direct quadruple loop on tiny inputs;
  * the backward pass: central finite differences of the FP64 forward;
  * GELU: its closed form 0.5 x (1 + erf(x / sqrt 2)) and derivative;
  * the box filter: the mean of f x f x f blocks, brute force;
  * the training loss and Adam: closed forms (the MSE of a constant offset;
    Adam's first step moves every parameter by lr * g / (|g| + eps), and a
    constant gradient keeps that size at every step).

Network: 3 layers of 3x3x3 convolutions, cin -> width -> width -> cout,
GELU after layers 1 and 2.  Arrays are channel-last, x[z][y][x][c];
periodic in x and y, zero padding in z.  Weights W[co][tap][ci] with
tap = 9 (dz+1) + 3 (dy+1) + (dx+1), biases b[co].  With bf16=True the
inputs, weights and hidden activations are rounded to bfloat16 (round to
nearest even) exactly where the GPU kernels store them, so the oracle and
the kernels differ only by FP32 vs FP64 accumulation.
"""

from __future__ import annotations

import numpy as np
from scipy.special import erf

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def shift(x, dz, dy, dx):
    """s[z,y,x] = x[z+dz, y+dy, x+dx] (periodic y, x; zero outside z)."""
    s = np.roll(np.roll(x, -dy, axis=1), -dx, axis=2)
    out = np.zeros_like(s)
    nz = x.shape[0]
    if dz == 0:
        out[:] = s
    elif dz > 0:
        out[: nz - dz] = s[dz:]
    else:
        out[-dz:] = s[: nz + dz]
    return out


def conv(x, w, b=None):
    """y[v, co] = b[co] + sum_tap sum_ci w[co, tap, ci] x[v + off(tap), ci]."""
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, (dz, dy, dx) in enumerate(TAPS):
        y += shift(x, dz, dy, dx) @ w[:, t, :].T
    if b is not None:
        y += b
    return y


def conv_backward(x, w, g):
    """Gradients of sum(g * conv(x, w, b)) w.r.t. x, w and b."""
    dx = np.zeros_like(x)
    dw = np.zeros_like(w)
    for t, (dz, dy, ddx) in enumerate(TAPS):
        xs = shift(x, dz, dy, ddx)
        dw[:, t, :] = np.einsum("zyxo,zyxi->oi", g, xs)
        dx += shift(g @ w[:, t, :], -dz, -dy, -ddx)
    return dx, dw, g.sum(axis=(0, 1, 2))


def conv_loop(x, w, b):
    """Direct loops (tiny inputs only)."""
    nz, ny, nx, _ = x.shape
    co = w.shape[0]
    y = np.zeros((nz, ny, nx, co))
    for z in range(nz):
        for j in range(ny):
            for i in range(nx):
                acc = b.astype(np.float64).copy()
                for t, (dz, dy, dx) in enumerate(TAPS):
                    zz = z + dz
                    if 0 <= zz < nz:
                        acc += w[:, t, :] @ x[zz, (j + dy) % ny, (i + dx) % nx, :]
                y[z, j, i] = acc
    return y


class Net:
    """params: dict W1,b1,W2,b2,W3,b3 (float64 arrays)."""

    def __init__(self, params, bf16_emulation=False):
        self.p = {k: np.asarray(v, dtype=np.float64) for k, v in params.items()}
        self.bf = bf16_emulation

    def _r(self, a):
        return bf16(a) if self.bf else a

    def forward(self, x, keep=False):
        p = self.p
        a0 = self._r(x)
        z1 = conv(a0, self._r(p["W1"]), p["b1"])
        a1 = self._r(gelu(z1))
        z2 = conv(a1, self._r(p["W2"]), p["b2"])
        a2 = self._r(gelu(z2))
        y = conv(a2, self._r(p["W3"]), p["b3"])
        if keep:
            self.cache = (a0, z1, a1, z2, a2)
        return y

    def backward(self, x, dy):
        """Gradients of sum(dy * forward(x)) w.r.t. the parameters (and x)."""
        p = self.p
        self.forward(x, keep=True)
        a0, z1, a1, z2, a2 = self.cache
        g3 = dy
        da2, dW3, db3 = conv_backward(a2, self._r(p["W3"]), g3)
        g2 = self._r(da2 * gelu_grad(z2)) if self.bf else da2 * gelu_grad(z2)
        da1, dW2, db2 = conv_backward(a1, self._r(p["W2"]), g2)
        g1 = self._r(da1 * gelu_grad(z1)) if self.bf else da1 * gelu_grad(z1)
        dx, dW1, db1 = conv_backward(a0, self._r(p["W1"]), g1)
        return {"W1": dW1, "b1": db1, "W2": dW2, "b2": db2, "W3": dW3, "b3": db3}, dx


def init_params(cin, width, cout, seed=0, scale=1.0):
    """He-style initial weights (the same generator the library's Python
    wrapper uses for its defaults)."""
    rng = np.random.default_rng(seed)
    out = {}
    for n, (ci, co) in enumerate([(cin, width), (width, width), (width, cout)], start=1):
        out[f"W{n}"] = rng.standard_normal((co, 27, ci)) * scale * np.sqrt(2.0 / (27 * ci))
        out[f"b{n}"] = rng.standard_normal(co) * 0.01
    return out


def flatten(params):
    return np.concatenate([params[k].ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])


def unflatten(vec, cin, width, cout):
    shapes = [("W1", (width, 27, cin)), ("b1", (width,)), ("W2", (width, 27, width)), ("b2", (width,)),
              ("W3", (cout, 27, width)), ("b3", (cout,))]
    out, o = {}, 0
    for k, s in shapes:
        n = int(np.prod(s))
        out[k] = np.asarray(vec[o:o + n], dtype=np.float64).reshape(s)
        o += n
    return out


def box_filter(a, f):
    """Mean of f x f x f blocks of a[z][y][x] (sizes divisible by f)."""
    nz, ny, nx = a.shape
    return a.reshape(nz // f, f, ny // f, f, nx // f, f).mean(axis=(1, 3, 5))


def box_filter_vertex(a, f):
    """Horizontal mean of f x f blocks; vertically coarse plane K on fine
    plane f K, trapezoid weights over the planes within f/2 (ends 1/2),
    boundary planes as they are (a has f (nzc-1) + 1 planes, f even)."""
    nz, ny, nx = a.shape
    h = a.reshape(nz, ny // f, f, nx // f, f).mean(axis=(2, 4))
    nzc = (nz - 1) // f + 1
    out = np.zeros((nzc,) + h.shape[1:])
    out[0], out[-1] = h[0], h[-1]
    w = np.ones(f + 1)
    w[0] = w[-1] = 0.5
    for K in range(1, nzc - 1):
        out[K] = np.tensordot(w, h[f * K - f // 2:f * K + f // 2 + 1], axes=(0, 0)) / f
    return out


def mse_loss(y, t, cout):
    """Training loss: mean over voxels and the first cout channels of
    (y - t)^2, with y - t formed in float32 and squared and summed in
    float64; and its gradient 2 (y - t) / (nvox * cout) as the kernels pass
    it on (float32 scale times the float32 difference, rounded to bfloat16)."""
    y = np.asarray(y, dtype=np.float32)[..., :cout]
    t = np.asarray(t, dtype=np.float32)[..., :cout]
    d = (y - t).astype(np.float32)
    nvox = int(np.prod(d.shape[:-1]))
    loss = float(np.sum(d.astype(np.float64) ** 2) / (nvox * cout))
    gscale = np.float32(2.0 / (nvox * cout))
    return loss, bf16((gscale * d).astype(np.float32))


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in float32: m, v moments, bias corrections
    c1 = 1 - b1^step, c2 = 1 - b2^step, p -= lr * (m / c1) / (sqrt(v / c2) + eps).
    Returns the new (p, m, v)."""
    f = np.float32
    p, g, m, v = (np.asarray(a, dtype=np.float32) for a in (p, g, m, v))
    b1, b2, lr, eps = f(b1), f(b2), f(lr), f(eps)
    m = (b1 * m + (f(1) - b1) * g).astype(np.float32)
    v = (b2 * v + (f(1) - b2) * g * g).astype(np.float32)
    c1 = f(1) - f(np.power(b1, f(step), dtype=np.float32))
    c2 = f(1) - f(np.power(b2, f(step), dtype=np.float32))
    p = (p - lr * (m / c1) / (np.sqrt(v / c2) + eps)).astype(np.float32)
    return p, m, v

