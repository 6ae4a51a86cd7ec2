"""Pins of the numpy network oracle (oracle/unet_oracle.py): the convolution
against scipy.signal.correlate, every backward pass against central finite
differences, pooling and upsampling as adjoints, and the whole U-Net's
parameter gradient against finite differences on a narrow 8^3 network."""
import numpy as np
import pytest
from scipy.signal import correlate

from oracle import unet_oracle as U


@pytest.mark.parametrize("ks", [1, 3])
def test_conv_matches_scipy_correlate(ks):
    rng = np.random.default_rng(ks)
    x = rng.standard_normal((3, 5, 6, 7))
    w = rng.standard_normal((4, 3, ks ** 3))
    b = rng.standard_normal(4)
    y = U.conv3d(x, w, b, ks)
    for o in range(4):
        ref = b[o] + sum(correlate(x[c], w[o, c].reshape(ks, ks, ks), mode="same")
                         for c in range(3))
        assert np.abs(y[o] - ref).max() < 1e-12


def fd_check(f, x, g, idx, h=1e-6):
    for i in idx:
        xp, xm = x.copy(), x.copy()
        xp.flat[i] += h
        xm.flat[i] -= h
        assert abs((f(xp) - f(xm)) / (2 * h) - g.flat[i]) < 1e-6 * (1 + abs(g.flat[i]))


def test_conv_backward_finite_differences():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 5, 3))
    w = rng.standard_normal((3, 2, 27))
    b = rng.standard_normal(3)
    r = rng.standard_normal((3, 4, 5, 3))
    dx, dw, db = U.conv3d_backward(x, w, r, 3)
    idx = rng.choice(x.size, 12, replace=False)
    fd_check(lambda v: np.sum(r * U.conv3d(v, w, b, 3)), x, dx, idx)
    fd_check(lambda v: np.sum(r * U.conv3d(x, v, b, 3)), w, dw, rng.choice(w.size, 12, False))
    fd_check(lambda v: np.sum(r * U.conv3d(x, w, v, 3)), b, db, range(3))


def test_pool_and_upsample_are_adjoint():
    rng = np.random.default_rng(1)
    f = rng.standard_normal((2, 4, 6, 8))
    c = rng.standard_normal((2, 2, 3, 4))
    assert abs(np.sum(U.pool2(f) * c) - np.sum(f * U.pool2_backward(c))) < 1e-12
    assert abs(np.sum(U.up2(c) * f) - np.sum(c * U.up2_backward(f))) < 1e-12
    assert np.allclose(U.pool2(U.up2(c)), c)


def test_unet_gradient_finite_differences():
    rng = np.random.default_rng(2)
    specs = U.layer_specs(cin=2, widths=(2, 3, 3, 4))
    n = sum(co * ci * ks ** 3 + co for _, ci, co, ks in specs)
    params = 0.3 * rng.standard_normal(n)
    x = rng.standard_normal((2, 8, 8, 8))
    r = rng.standard_normal((1, 8, 8, 8))
    y, A = U.unet_forward(x, params, specs)
    g = U.unet_backward(r, params, specs, A)
    assert g.shape == params.shape
    # a sample from every layer
    idx, o = [], 0
    for _, ci, co, ks in specs:
        nw = co * ci * ks ** 3
        idx += [o + int(rng.integers(nw)), o + nw + int(rng.integers(co))]
        o += nw + co
    fd_check(lambda v: np.sum(r * U.unet_forward(x, v, specs)[0]), params, g, idx)
