"""Pins of the closure network's CPU oracle (oracle/net_oracle.py): direct
loops, impulse response, finite differences, GELU closed form, bf16
rounding, and the box filter against brute force."""

import os
import sys

import numpy as np
import pytest

import helpers as H

sys.path.insert(0, os.path.join(H.REPO, "oracle"))
import net_oracle as N  # noqa: E402


def test_conv_matches_direct_loops():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((4, 5, 6, 3))
    w = rng.standard_normal((2, 27, 3))
    b = rng.standard_normal(2)
    assert np.allclose(N.conv(x, w, b), N.conv_loop(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv_impulse_response_is_the_flipped_kernel():
    nz, ny, nx = 5, 6, 7
    x = np.zeros((nz, ny, nx, 1))
    x[2, 3, 3, 0] = 1.0
    w = np.arange(27, dtype=float).reshape(1, 27, 1) + 1.0
    y = N.conv(x, w)
    for t, (dz, dy, dx) in enumerate(N.TAPS):
        # y[v] picks x[v + off]: the impulse at c is seen from v = c - off
        assert y[2 - dz, 3 - dy, 3 - dx, 0] == w[0, t, 0]
    assert np.count_nonzero(y) == 27


def test_conv_periodic_in_x_and_y_zero_in_z():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((3, 4, 5, 2))
    w = rng.standard_normal((2, 27, 2))
    y = N.conv(x, w)
    ys = N.conv(np.roll(np.roll(x, 1, axis=1), 2, axis=2), w)
    assert np.allclose(np.roll(np.roll(y, 1, axis=1), 2, axis=2), ys)
    # zero padding in z: the top plane sees no plane above it
    x2 = x.copy()
    x2[0] = 0.0
    w2 = np.zeros_like(w)
    w2[:, :9, :] = w[:, :9, :]  # dz = -1 taps only
    assert np.allclose(N.conv(x2, w2)[0], 0.0)


def test_gelu_closed_form_and_derivative():
    x = np.linspace(-6, 6, 1001)
    assert np.allclose(N.gelu(0.0), 0.0) and abs(N.gelu(1.0) - 0.8413447460685429) < 1e-15
    h = 1e-6
    fd = (N.gelu(x + h) - N.gelu(x - h)) / (2 * h)
    assert np.allclose(N.gelu_grad(x), fd, atol=1e-8)


def test_network_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    cin, width, cout = 3, 4, 2
    p = N.init_params(cin, width, cout, seed=4)
    x = rng.standard_normal((3, 4, 4, cin))
    dy = rng.standard_normal((3, 4, 4, cout))
    net = N.Net(p)
    g, gx = net.backward(x, dy)
    vec = N.flatten(p)
    gvec = N.flatten(g)
    h = 1e-6
    for idx in rng.choice(vec.size, 40, replace=False):
        vp, vm = vec.copy(), vec.copy()
        vp[idx] += h
        vm[idx] -= h
        fp = np.sum(dy * N.Net(N.unflatten(vp, cin, width, cout)).forward(x))
        fm = np.sum(dy * N.Net(N.unflatten(vm, cin, width, cout)).forward(x))
        assert abs((fp - fm) / (2 * h) - gvec[idx]) <= 1e-6 * max(1.0, abs(gvec[idx])), idx
    for idx in [(0, 0, 0, 0), (2, 3, 1, 2), (1, 2, 3, 1)]:
        xp, xm = x.copy(), x.copy()
        xp[idx] += h
        xm[idx] -= h
        fd = (np.sum(dy * net.forward(xp)) - np.sum(dy * net.forward(xm))) / (2 * h)
        assert abs(fd - gx[idx]) <= 1e-6 * max(1.0, abs(gx[idx]))


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -3.140625, 1e-3])
    r = N.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == 1.0  # tie -> even
    assert r[2] == 1.0 + 2 ** -7
    assert r[3] == -3.140625
    assert abs(r[4] - 1e-3) <= 2 ** -9 * 1e-3


@pytest.mark.parametrize("f", [2, 4])
def test_box_filter_brute_force(f):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((2 * f, 3 * f, 4 * f))
    got = N.box_filter(a, f)
    for K in range(2):
        for J in range(3):
            for I in range(4):
                assert abs(got[K, J, I] - a[K * f:(K + 1) * f, J * f:(J + 1) * f, I * f:(I + 1) * f].mean()) < 1e-14


def test_box_filter_vertex_brute_force():
    rng = np.random.default_rng(6)
    f = 4
    a = rng.standard_normal((f * 3 + 1, 2 * f, 3 * f))
    got = N.box_filter_vertex(a, f)
    assert got.shape == (4, 2, 3)
    hm = lambda k, J, I: a[k, J * f:(J + 1) * f, I * f:(I + 1) * f].mean()
    for J in range(2):
        for I in range(3):
            assert abs(got[0, J, I] - hm(0, J, I)) < 1e-14 and abs(got[3, J, I] - hm(12, J, I)) < 1e-14
            for K in (1, 2):
                c = f * K
                ref = (0.5 * hm(c - 2, J, I) + hm(c - 1, J, I) + hm(c, J, I) + hm(c + 1, J, I) + 0.5 * hm(c + 2, J, I)) / 4
                assert abs(got[K, J, I] - ref) < 1e-14
    # a field linear in the plane index is reproduced at the coarse planes
    lin = np.broadcast_to(np.arange(f * 3 + 1, dtype=float)[:, None, None], a.shape)
    assert np.allclose(N.box_filter_vertex(lin, f)[:, 0, 0], [0, 4, 8, 12])


def test_mse_loss_of_a_constant_offset():
    """y = t + c on the first cout channels: loss c^2, gradient 2c/(n cout);
    channels beyond cout do not count."""
    rng = np.random.default_rng(1)
    t = rng.standard_normal((3, 4, 5, 8)).astype(np.float32)
    y = t + np.float32(0.25)
    y[..., 7] += 100.0
    loss, g = N.mse_loss(y, t, 7)
    assert loss == pytest.approx(0.0625, rel=1e-6)
    assert g.shape == (3, 4, 5, 7)
    assert np.allclose(g, 2 * 0.25 / (60 * 7), rtol=1e-2)


def test_adam_first_steps_closed_form():
    """Adam's first step is lr * g / (|g| + eps) for every parameter, and a
    constant gradient keeps the step size lr * |g| / (|g| + eps)."""
    rng = np.random.default_rng(2)
    p0 = rng.standard_normal(50).astype(np.float32)
    g = rng.standard_normal(50).astype(np.float32)
    z = np.zeros(50, np.float32)
    p1, m, v = N.adam_step(p0, g, z, z, 1, 1e-3)
    want = 1e-3 * g / (np.abs(g) + 1e-8)
    tol = 4 * np.finfo(np.float32).eps * np.abs(p0).max()  # float32 rounding of p
    assert np.allclose(p0 - p1, want, rtol=0, atol=tol)
    p2, _, _ = N.adam_step(p1, g, m, v, 2, 1e-3)
    assert np.allclose(p1 - p2, want, rtol=0, atol=tol)

