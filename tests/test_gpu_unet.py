"""The network layers of csrc/unet.hip and the whole U-Net's forward and
backward passes against the numpy oracle (oracle/unet_oracle.py), through the
C ABI.  FP32 on the device against FP64 in the oracle: 1e-5 of the largest
entry for single layers, 1e-4 for the whole network."""
import ctypes

import numpy as np
import pytest

from oracle import unet_oracle as U
from ttdft_public_amd import engine as E
from ttdft_public_amd import learned as Ln

pytestmark = pytest.mark.gpu


def f32(ctx, a):
    return E.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float32))


def dims(d):
    return (ctypes.c_int32 * 3)(*d)


def close(a, b, tol):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-30)


@pytest.mark.parametrize("ci,co,ks,shape", [(3, 37, 3, (5, 6, 7)), (16, 16, 3, (8, 8, 8)),
                                            (48, 16, 3, (4, 8, 8)), (16, 1, 1, (6, 6, 6)),
                                            (130, 5, 3, (3, 4, 5))])
def test_conv_forward_and_backward(ctx, ci, co, ks, shape):
    rng = np.random.default_rng(ci + co)
    x = rng.standard_normal((ci,) + shape)
    w = rng.standard_normal((co, ci, ks ** 3)) / np.sqrt(ci * ks ** 3)
    b = rng.standard_normal(co)
    dy = rng.standard_normal((co,) + shape)
    L, h = E.lib(), ctx.h
    dx_, dw_, db_, ddy = f32(ctx, x), f32(ctx, w), f32(ctx, b), f32(ctx, dy)
    P = int(np.prod(shape))
    print(f"conv3d ci={ci} co={co} ks={ks} shape={shape}", flush=True)
    y = E.DeviceArray(ctx, (co * P,), np.float32)
    E._check(L.ttdft_nn_conv3d_forward(h, dx_.ptr, ci, dims(shape), dw_.ptr, db_.ptr, co, ks,
                                       y.ptr, 0))
    y0 = U.conv3d(x, w, b, ks)
    assert close(y.to_host().reshape(y0.shape), y0, 1e-5)
    E._check(L.ttdft_nn_conv3d_forward(h, dx_.ptr, ci, dims(shape), dw_.ptr, None, co, ks, y.ptr,
                                       1))
    assert close(y.to_host().reshape(y0.shape), 2 * y0 - b[:, None, None, None], 1e-5)
    gx0, gw0, gb0 = U.conv3d_backward(x, w, dy, ks)
    gx = E.DeviceArray(ctx, (ci * P,), np.float32)
    E._check(L.ttdft_nn_conv3d_backward_data(h, ddy.ptr, co, dims(shape), dw_.ptr, ci, ks, gx.ptr,
                                             0))
    assert close(gx.to_host().reshape(gx0.shape), gx0, 1e-5)
    gw = f32(ctx, np.ones(w.size))
    gb = f32(ctx, np.ones(co))
    E._check(L.ttdft_nn_conv3d_backward_weight(h, dx_.ptr, ci, dims(shape), ddy.ptr, co, ks,
                                               gw.ptr, gb.ptr, 1))
    assert close(gw.to_host().reshape(gw0.shape), gw0 + 1, 1e-5)
    assert close(gb.to_host(), gb0 + 1, 1e-5)
    E._check(L.ttdft_nn_conv3d_backward_weight(h, dx_.ptr, ci, dims(shape), ddy.ptr, co, ks,
                                               gw.ptr, None, 0))
    assert close(gw.to_host().reshape(gw0.shape), gw0, 1e-5)


def test_activation_pool_upsample_loss_adam(ctx):
    rng = np.random.default_rng(7)
    L, h = E.lib(), ctx.h
    x = rng.standard_normal((3, 4, 6, 8))
    d = f32(ctx, x)
    E._check(L.ttdft_nn_leaky_relu(h, d.ptr, x.size, U.SLOPE))
    y = U.leaky(x)
    assert close(d.to_host().reshape(x.shape), y, 1e-7)
    g = rng.standard_normal(x.shape)
    dg = f32(ctx, g)
    E._check(L.ttdft_nn_leaky_relu_backward(h, d.ptr, dg.ptr, x.size, U.SLOPE))
    assert close(dg.to_host().reshape(x.shape), U.leaky_backward(y, g), 1e-7)
    # pooling and upsampling on the fine dims (4, 6, 8)
    fine = dims((4, 6, 8))
    dp = E.DeviceArray(ctx, (3 * 2 * 3 * 4,), np.float32)
    dx = f32(ctx, x)
    E._check(L.ttdft_nn_avgpool2(h, dx.ptr, 3, fine, dp.ptr))
    assert close(dp.to_host().reshape(3, 2, 3, 4), U.pool2(x), 1e-6)
    c = rng.standard_normal((3, 2, 3, 4))
    dc = f32(ctx, c)
    dfine = f32(ctx, np.ones(x.size))
    E._check(L.ttdft_nn_avgpool2_backward(h, dc.ptr, 3, fine, dfine.ptr, 1))
    assert close(dfine.to_host().reshape(x.shape), U.pool2_backward(c) + 1, 1e-6)
    E._check(L.ttdft_nn_upsample2(h, dc.ptr, 3, fine, dfine.ptr))
    assert close(dfine.to_host().reshape(x.shape), U.up2(c), 1e-7)
    E._check(L.ttdft_nn_upsample2_backward(h, dx.ptr, 3, fine, dc.ptr, 0))
    assert close(dc.to_host().reshape(c.shape), U.up2_backward(x), 1e-6)
    # loss
    t = rng.standard_normal(x.size)
    dt, dy2 = f32(ctx, t), E.DeviceArray(ctx, (x.size,), np.float32)
    loss = ctypes.c_double()
    E._check(L.ttdft_nn_mse(h, dx.ptr, dt.ptr, x.size, dy2.ptr, ctypes.byref(loss)))
    xr = x.ravel().astype(np.float32).astype(np.float64)
    tr = t.astype(np.float32).astype(np.float64)
    assert abs(loss.value - np.mean((xr - tr) ** 2)) < 1e-6 * loss.value
    assert close(dy2.to_host(), 2 * (xr - tr) / x.size, 1e-6)
    # two Adam steps
    p0, gg = rng.standard_normal(50), rng.standard_normal(50)
    dpp, dgg = f32(ctx, p0), f32(ctx, gg)
    dm, dv = f32(ctx, np.zeros(50)), f32(ctx, np.zeros(50))
    m = v = np.zeros(50)
    p = p0.copy()
    for step in (1, 2):
        E._check(L.ttdft_nn_adam(h, dpp.ptr, dgg.ptr, dm.ptr, dv.ptr, 50, 1e-2, 0.9, 0.999, 1e-8,
                                 step))
        m = 0.9 * m + 0.1 * gg
        v = 0.999 * v + 0.001 * gg ** 2
        p = p - 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    assert close(dpp.to_host(), p, 1e-6)


@pytest.mark.parametrize("n,widths", [(8, (2, 3, 3, 4)), (16, (16, 32, 64, 128))])
def test_unet_forward_backward_match_oracle(ctx, n, widths):
    """Device U-Net (learned.UNet) against the oracle's forward pass and its
    parameter gradient for the loss mean (y - t)^2."""
    rng = np.random.default_rng(n)
    net = Ln.UNet(ctx, n, 2, widths, seed=1)
    specs = U.layer_specs(2, widths)
    assert [s for s in specs] == net.specs
    params = net.get_params().astype(np.float64)
    params[-1] = 0.05  # nonzero head bias
    params += 0.05 * rng.standard_normal(params.size) * (params != 0)
    net.set_params(params)
    params = net.get_params().astype(np.float64)
    x = rng.standard_normal((2, n, n, n)).astype(np.float32)
    t = rng.standard_normal((1, n, n, n)).astype(np.float32)
    net.a["x"].copy_from(x.ravel())
    net.a["t"].copy_from(t.ravel())
    print(f"unet n={n} widths={widths}", flush=True)
    net.forward()
    loss = net.loss_and_grad()
    net.backward()
    y0, A = U.unet_forward(x.astype(np.float64), params, specs)
    assert close(net.a["y"].to_host().reshape(y0.shape), y0, 1e-4)
    assert abs(loss - np.mean((y0 - t) ** 2)) < 1e-4 * loss
    g0 = U.unet_backward(2 * (y0 - t) / y0.size, params, specs, A)
    g = net.grads.to_host()
    for name, ci, co, ks in specs:  # layer by layer, relative to the layer's largest entry
        wo, bo = net.offsets[name]
        sl = slice(wo, bo + co)
        assert close(g[sl], g0[sl], 2e-4), name


@pytest.mark.parametrize("ci,co,ks,shape", [(3, 37, 3, (5, 6, 7)), (48, 16, 3, (8, 8, 8)),
                                            (16, 1, 1, (6, 6, 6))])
def test_conv_bf16_mfma(ctx, ci, co, ks, shape):
    """BF16 MFMA forms (operands rounded to BF16, FP32 accumulation): forward,
    backward data and backward weight within BF16 rounding of the oracle,
    2e-2 of the largest entry."""
    rng = np.random.default_rng(ci * co)
    x = rng.standard_normal((ci,) + shape)
    w = rng.standard_normal((co, ci, ks ** 3)) / np.sqrt(ci * ks ** 3)
    b = rng.standard_normal(co)
    dy = rng.standard_normal((co,) + shape)
    L, h = E.lib(), ctx.h
    P = int(np.prod(shape))
    dx_, dw_, db_, ddy = f32(ctx, x), f32(ctx, w), f32(ctx, b), f32(ctx, dy)
    y = E.DeviceArray(ctx, (co * P,), np.float32)
    gx = E.DeviceArray(ctx, (ci * P,), np.float32)
    gw = E.DeviceArray(ctx, (w.size,), np.float32)
    E._check(L.ttdft_nn_set_math(h, E.NN_BF16))
    try:
        print(f"conv3d bf16 ci={ci} co={co} ks={ks} shape={shape}", flush=True)
        E._check(L.ttdft_nn_conv3d_forward(h, dx_.ptr, ci, dims(shape), dw_.ptr, db_.ptr, co, ks,
                                           y.ptr, 0))
        E._check(L.ttdft_nn_conv3d_backward_data(h, ddy.ptr, co, dims(shape), dw_.ptr, ci, ks,
                                                 gx.ptr, 0))
        E._check(L.ttdft_nn_conv3d_backward_weight(h, dx_.ptr, ci, dims(shape), ddy.ptr, co, ks,
                                                   gw.ptr, None, 0))
    finally:
        E._check(L.ttdft_nn_set_math(h, E.NN_FP32))
    y0 = U.conv3d(x, w, b, ks)
    gx0, gw0, _ = U.conv3d_backward(x, w, dy, ks)
    assert close(y.to_host().reshape(y0.shape), y0, 2e-2)
    assert close(gx.to_host().reshape(gx0.shape), gx0, 2e-2)
    assert close(gw.to_host().reshape(gw0.shape), gw0, 2e-2)
    # and it is not the FP32 path: BF16 rounding shows at the 1e-4 level
    assert not close(y.to_host().reshape(y0.shape), y0, 1e-6)
