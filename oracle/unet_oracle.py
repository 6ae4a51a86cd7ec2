"""numpy restatement of the learned initial-density network's layers and of
the whole network's forward and backward passes (test infrastructure only:
tests compare the device kernels of csrc/unet.hip with it).

Tensors are one sample's [C][D][H][W] arrays.  Convolution weights are
[co][ci][ks^3] with taps z-major (tap = kz ks^2 + ky ks + kx), stride 1,
zero padding ks // 2.  Pinned in tests/test_oracle_unet.py: the convolution
against scipy.signal.correlate, every backward against central finite
differences, pooling / upsampling as each other's adjoints.
"""
import numpy as np

SLOPE = 0.1


def _pad(x, p):
    return np.pad(x, ((0, 0), (p, p), (p, p), (p, p))) if p else x


def conv3d(x, w, b, ks):
    """y[co] = b[co] + sum_{ci, tap} w[co, ci, tap] x[ci] shifted by the tap."""
    ci, D, H, W = x.shape
    co = w.shape[0]
    p = ks // 2
    xp = _pad(x, p)
    y = np.zeros((co, D, H, W), dtype=np.float64)
    w4 = w.reshape(co, ci, ks, ks, ks)
    for kz in range(ks):
        for ky in range(ks):
            for kx in range(ks):
                sl = xp[:, kz:kz + D, ky:ky + H, kx:kx + W]
                y += np.einsum("oc,cdhw->odhw", w4[:, :, kz, ky, kx], sl)
    if b is not None:
        y += b[:, None, None, None]
    return y


def conv3d_backward(x, w, dy, ks):
    """(dx, dw, db) of conv3d."""
    ci, D, H, W = x.shape
    co = w.shape[0]
    p = ks // 2
    xp = _pad(x, p)
    w4 = w.reshape(co, ci, ks, ks, ks)
    dxp = np.zeros_like(xp, dtype=np.float64)
    dw = np.zeros((co, ci, ks, ks, ks))
    for kz in range(ks):
        for ky in range(ks):
            for kx in range(ks):
                sl = xp[:, kz:kz + D, ky:ky + H, kx:kx + W]
                dw[:, :, kz, ky, kx] = np.einsum("odhw,cdhw->oc", dy, sl)
                dxp[:, kz:kz + D, ky:ky + H, kx:kx + W] += np.einsum("oc,odhw->cdhw",
                                                                      w4[:, :, kz, ky, kx], dy)
    dx = dxp[:, p:p + D, p:p + H, p:p + W] if p else dxp
    return dx, dw.reshape(co, ci, ks ** 3), dy.sum(axis=(1, 2, 3))


def leaky(x):
    return np.where(x > 0, x, SLOPE * x)


def leaky_backward(y, dy):
    return np.where(y > 0, dy, SLOPE * dy)


def pool2(x):
    c, D, H, W = x.shape
    return x.reshape(c, D // 2, 2, H // 2, 2, W // 2, 2).mean(axis=(2, 4, 6))


def pool2_backward(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 1), 2, 2), 2, 3) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up2_backward(dy):
    c, D, H, W = dy.shape
    return dy.reshape(c, D // 2, 2, H // 2, 2, W // 2, 2).sum(axis=(2, 4, 6))


def layer_specs(cin=2, widths=(16, 32, 64, 128)):
    """(name, ci, co, ks) of the four-level U-Net, in parameter order."""
    c0, c1, c2, c3 = widths
    return [("enc0a", cin, c0, 3), ("enc0b", c0, c0, 3),
            ("enc1a", c0, c1, 3), ("enc1b", c1, c1, 3),
            ("enc2a", c1, c2, 3), ("enc2b", c2, c2, 3),
            ("bot_a", c2, c3, 3), ("bot_b", c3, c3, 3),
            ("dec2a", c3 + c2, c2, 3), ("dec2b", c2, c2, 3),
            ("dec1a", c2 + c1, c1, 3), ("dec1b", c1, c1, 3),
            ("dec0a", c1 + c0, c0, 3), ("dec0b", c0, c0, 3),
            ("head", c0, 1, 1)]


def unpack(flat, specs):
    """{name: (w, b)} views of a flat parameter vector (w then b per layer)."""
    out, o = {}, 0
    for name, ci, co, ks in specs:
        nw = co * ci * ks ** 3
        out[name] = (flat[o:o + nw].reshape(co, ci, ks ** 3), flat[o + nw:o + nw + co])
        o += nw + co
    return out


def unet_forward(x, params, specs):
    """Output and the activations the backward pass needs."""
    P = unpack(params, specs)
    ks = {n: k for n, _, _, k in specs}
    A = {}

    def cl(name, inp):
        y = leaky(conv3d(inp, *P[name], ks[name]))
        A[name] = y
        return y

    s0 = cl("enc0b", cl("enc0a", x))
    s1 = cl("enc1b", cl("enc1a", pool2(s0)))
    s2 = cl("enc2b", cl("enc2a", pool2(s1)))
    b = cl("bot_b", cl("bot_a", pool2(s2)))
    c2 = np.concatenate([up2(b), s2])
    d2 = cl("dec2b", cl("dec2a", c2))
    c1 = np.concatenate([up2(d2), s1])
    d1 = cl("dec1b", cl("dec1a", c1))
    c0 = np.concatenate([up2(d1), s0])
    d0 = cl("dec0b", cl("dec0a", c0))
    y = conv3d(d0, *P["head"], 1)
    A.update(x=x, c2=c2, c1=c1, c0=c0)
    return y, A


def unet_backward(dy, params, specs, A):
    """Flat gradient of the parameters for the output gradient dy."""
    P = unpack(params, specs)
    ks = {n: k for n, _, _, k in specs}
    G = {}
    inputs = {"enc0a": A["x"], "enc0b": A["enc0a"], "enc1a": pool2(A["enc0b"]),
              "enc1b": A["enc1a"], "enc2a": pool2(A["enc1b"]), "enc2b": A["enc2a"],
              "bot_a": pool2(A["enc2b"]), "bot_b": A["bot_a"], "dec2a": A["c2"],
              "dec2b": A["dec2a"], "dec1a": A["c1"], "dec1b": A["dec1a"], "dec0a": A["c0"],
              "dec0b": A["dec0a"], "head": A["dec0b"]}

    def back(name, g, act=True):
        if act:
            g = leaky_backward(A[name], g)
        dx, dw, db = conv3d_backward(inputs[name], P[name][0], g, ks[name])
        G[name] = (dw, db)
        return dx

    g = back("head", dy, act=False)
    g = back("dec0a", back("dec0b", g))
    c_up = specs[12][1] - specs[0][2]
    g_s0 = g[c_up:]
    g = up2_backward(g[:c_up])
    g = back("dec1a", back("dec1b", g))
    c_up = specs[10][1] - specs[2][2]
    g_s1 = g[c_up:]
    g = up2_backward(g[:c_up])
    g = back("dec2a", back("dec2b", g))
    c_up = specs[8][1] - specs[4][2]
    g_s2 = g[c_up:]
    g = up2_backward(g[:c_up])
    g = back("bot_a", back("bot_b", g))
    g = back("enc2a", back("enc2b", g_s2 + pool2_backward(g)))
    g = back("enc1a", back("enc1b", g_s1 + pool2_backward(g)))
    back("enc0a", back("enc0b", g_s0 + pool2_backward(g)))
    return np.concatenate([np.concatenate([G[n][0].ravel(), G[n][1].ravel()])
                           for n, _, _, _ in specs])
