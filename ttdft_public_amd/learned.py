"""Learned correction to the atomic-superposition initial density.

A four-level 3-D U-Net (widths 16-32-64-128, 1.46 M parameters) maps two
channels on a uniform n^3 grid over the simulation box -- Gaussian splats of
the atom positions (sigma 1 bohr) and the renormalised atomic superposition
-- to the correction rho_converged - rho_superposition.  Every layer runs on
the device through the engine's hand-written kernels (csrc/unet.hip):
forward, backward and the Adam update; the training data are densities that
the SCF loop (scf.py) converges on seeded clusters.

The start handed to the SCF is the superposition on the quadrature grid plus
the predicted correction (trilinear from the uniform grid), clamped at zero
and renormalised to the electron count.  The plain superposition
(InitialGuess::initialize_rho, dft/InitialGuess.cc:683-762, is the start this
corrects) stays as the fallback: if the first SCF residual from the learned
start exceeds the superposition's, the loop restarts from the superposition.
The superposition's first residual is probed only when the learned start's
is not clearly below the smallest seen on the evaluation clusters.
"""
import ctypes
import math
import os
import time

import numpy as np

from . import engine as E
from . import scf as S

SLOPE = 0.1
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                               "unet_initial_density.npz")


def layer_specs(cin=2, widths=(16, 32, 64, 128)):
    """(name, ci, co, ks) of the network's convolutions, in parameter order."""
    c0, c1, c2, c3 = widths
    return [("enc0a", cin, c0, 3), ("enc0b", c0, c0, 3),
            ("enc1a", c0, c1, 3), ("enc1b", c1, c1, 3),
            ("enc2a", c1, c2, 3), ("enc2b", c2, c2, 3),
            ("bot_a", c2, c3, 3), ("bot_b", c3, c3, 3),
            ("dec2a", c3 + c2, c2, 3), ("dec2b", c2, c2, 3),
            ("dec1a", c2 + c1, c1, 3), ("dec1b", c1, c1, 3),
            ("dec0a", c1 + c0, c0, 3), ("dec0b", c0, c0, 3),
            ("head", c0, 1, 1)]


def _c(rc):
    E._check(rc)


class UNet:
    """The network's parameters, Adam state and activations on one device."""

    def __init__(self, ctx, n=64, cin=2, widths=(16, 32, 64, 128), seed=0, arith=E.NN_FP32):
        assert n % 8 == 0
        self.ctx, self.n, self.cin, self.widths = ctx, n, cin, tuple(widths)
        self.arith = arith  # E.NN_FP32, or E.NN_BF16: convolutions on BF16 MFMA
        self.specs = layer_specs(cin, widths)
        self.offsets = {}
        o = 0
        for name, ci, co, ks in self.specs:
            self.offsets[name] = (o, o + co * ci * ks ** 3)
            o += co * ci * ks ** 3 + co
        self.nparams = o
        rng = np.random.default_rng(seed)
        init = np.zeros(o, np.float32)
        for name, ci, co, ks in self.specs:
            wo, _ = self.offsets[name]
            fan = ci * ks ** 3
            std = math.sqrt(2.0 / ((1 + SLOPE ** 2) * fan)) if name != "head" else 1e-2 / fan
            init[wo:wo + co * fan] = rng.standard_normal(co * fan) * std
        self.params = E.DeviceArray.from_host(ctx, init)
        zeros = np.zeros(o, np.float32)
        self.grads = E.DeviceArray.from_host(ctx, zeros)
        self.m = E.DeviceArray.from_host(ctx, zeros)
        self.v = E.DeviceArray.from_host(ctx, zeros)
        self.step_count = 0
        self.in_scale, self.out_scale = 1.0, 1.0
        self._alloc()

    # ---- buffers -----------------------------------------------------------------
    def _alloc(self):
        n, (c0, c1, c2, c3) = self.n, self.widths
        P = [n ** 3 // 8 ** l for l in range(4)]
        self.P = P
        self.dims = [(n >> l,) * 3 for l in range(4)]
        f32 = np.float32
        shapes = {"x": self.cin * P[0], "enc0a": c0 * P[0], "cat0": (c1 + c0) * P[0],
                  "p0": c0 * P[1], "enc1a": c1 * P[1], "cat1": (c2 + c1) * P[1],
                  "p1": c1 * P[2], "enc2a": c2 * P[2], "cat2": (c3 + c2) * P[2],
                  "p2": c2 * P[3], "bot_a": c3 * P[3], "bot_b": c3 * P[3],
                  "dec2a": c2 * P[2], "dec2b": c2 * P[2], "dec1a": c1 * P[1],
                  "dec1b": c1 * P[1], "dec0a": c0 * P[0], "dec0b": c0 * P[0], "y": P[0],
                  "t": P[0]}
        self.a = {k: E.DeviceArray(self.ctx, (v,), f32) for k, v in shapes.items()}
        self.g = {k: E.DeviceArray(self.ctx, (v,), f32) for k, v in shapes.items()
                  if k not in ("x", "t")}

    def _w(self, name):
        wo, bo = self.offsets[name]
        return self.params.ptr + 4 * wo, self.params.ptr + 4 * bo

    def _gw(self, name):
        wo, bo = self.offsets[name]
        return self.grads.ptr + 4 * wo, self.grads.ptr + 4 * bo

    @staticmethod
    def _sub(buf, channel_offset, P):
        return buf.ptr + 4 * channel_offset * P

    def _dims(self, level):
        return (ctypes.c_int32 * 3)(*self.dims[level])

    # ---- layers --------------------------------------------------------------------
    def _conv(self, name, x_ptr, y_ptr, level, act=True):
        _, ci, co, ks = next(s for s in self.specs if s[0] == name)
        w, b = self._w(name)
        _c(E.lib().ttdft_nn_conv3d_forward(self.ctx.h, x_ptr, ci, self._dims(level), w, b, co, ks,
                                           y_ptr, 0))
        if act:
            _c(E.lib().ttdft_nn_leaky_relu(self.ctx.h, y_ptr, co * self.P[level], SLOPE))

    def _conv_back(self, name, x_ptr, y_ptr, gy_ptr, gx_ptr, level, act=True):
        """Gradient of one convolution (+ activation): parameter gradients are
        written (added when self._acc), the input gradient written to gx_ptr
        (if not None)."""
        _, ci, co, ks = next(s for s in self.specs if s[0] == name)
        if act:
            _c(E.lib().ttdft_nn_leaky_relu_backward(self.ctx.h, y_ptr, gy_ptr, co * self.P[level],
                                                    SLOPE))
        gw, gb = self._gw(name)
        d = self._dims(level)
        _c(E.lib().ttdft_nn_conv3d_backward_weight(self.ctx.h, x_ptr, ci, d, gy_ptr, co, ks, gw, gb,
                                                   self._acc))
        if gx_ptr is not None:
            w, _ = self._w(name)
            _c(E.lib().ttdft_nn_conv3d_backward_data(self.ctx.h, gy_ptr, co, d, w, ci, ks, gx_ptr,
                                                     0))

    def _math(self, on):
        _c(E.lib().ttdft_nn_set_math(self.ctx.h, self.arith if on else E.NN_FP32))

    def forward(self):
        """y = net(x) on the buffers a["x"] -> a["y"]."""
        self._math(True)
        try:
            self._forward()
        finally:
            self._math(False)

    def _forward(self):
        a, P, (c0, c1, c2, c3) = self.a, self.P, self.widths
        L, h = E.lib(), self.ctx.h
        s0, s1, s2 = (self._sub(a["cat0"], c1, P[0]), self._sub(a["cat1"], c2, P[1]),
                      self._sub(a["cat2"], c3, P[2]))
        self._conv("enc0a", a["x"].ptr, a["enc0a"].ptr, 0)
        self._conv("enc0b", a["enc0a"].ptr, s0, 0)
        _c(L.ttdft_nn_avgpool2(h, s0, c0, self._dims(0), a["p0"].ptr))
        self._conv("enc1a", a["p0"].ptr, a["enc1a"].ptr, 1)
        self._conv("enc1b", a["enc1a"].ptr, s1, 1)
        _c(L.ttdft_nn_avgpool2(h, s1, c1, self._dims(1), a["p1"].ptr))
        self._conv("enc2a", a["p1"].ptr, a["enc2a"].ptr, 2)
        self._conv("enc2b", a["enc2a"].ptr, s2, 2)
        _c(L.ttdft_nn_avgpool2(h, s2, c2, self._dims(2), a["p2"].ptr))
        self._conv("bot_a", a["p2"].ptr, a["bot_a"].ptr, 3)
        self._conv("bot_b", a["bot_a"].ptr, a["bot_b"].ptr, 3)
        _c(L.ttdft_nn_upsample2(h, a["bot_b"].ptr, c3, self._dims(2), a["cat2"].ptr))
        self._conv("dec2a", a["cat2"].ptr, a["dec2a"].ptr, 2)
        self._conv("dec2b", a["dec2a"].ptr, a["dec2b"].ptr, 2)
        _c(L.ttdft_nn_upsample2(h, a["dec2b"].ptr, c2, self._dims(1), a["cat1"].ptr))
        self._conv("dec1a", a["cat1"].ptr, a["dec1a"].ptr, 1)
        self._conv("dec1b", a["dec1a"].ptr, a["dec1b"].ptr, 1)
        _c(L.ttdft_nn_upsample2(h, a["dec1b"].ptr, c1, self._dims(0), a["cat0"].ptr))
        self._conv("dec0a", a["cat0"].ptr, a["dec0a"].ptr, 0)
        self._conv("dec0b", a["dec0a"].ptr, a["dec0b"].ptr, 0)
        self._conv("head", a["dec0b"].ptr, a["y"].ptr, 0, act=False)

    def backward(self, accumulate=False):
        """Parameter gradients for the output gradient g["y"], written or (with
        accumulate, over the samples of a batch) added to self.grads."""
        self._acc = 1 if accumulate else 0
        self._math(True)
        try:
            self._backward()
        finally:
            self._math(False)

    def _backward(self):
        a, g, P, (c0, c1, c2, c3) = self.a, self.g, self.P, self.widths
        L, h = E.lib(), self.ctx.h
        s0, s1, s2 = (self._sub(a["cat0"], c1, P[0]), self._sub(a["cat1"], c2, P[1]),
                      self._sub(a["cat2"], c3, P[2]))
        gs0, gs1, gs2 = (self._sub(g["cat0"], c1, P[0]), self._sub(g["cat1"], c2, P[1]),
                         self._sub(g["cat2"], c3, P[2]))
        cb = self._conv_back
        cb("head", a["dec0b"].ptr, None, g["y"].ptr, g["dec0b"].ptr, 0, act=False)
        cb("dec0b", a["dec0a"].ptr, a["dec0b"].ptr, g["dec0b"].ptr, g["dec0a"].ptr, 0)
        cb("dec0a", a["cat0"].ptr, a["dec0a"].ptr, g["dec0a"].ptr, g["cat0"].ptr, 0)
        _c(L.ttdft_nn_upsample2_backward(h, g["cat0"].ptr, c1, self._dims(0), g["dec1b"].ptr, 0))
        cb("dec1b", a["dec1a"].ptr, a["dec1b"].ptr, g["dec1b"].ptr, g["dec1a"].ptr, 1)
        cb("dec1a", a["cat1"].ptr, a["dec1a"].ptr, g["dec1a"].ptr, g["cat1"].ptr, 1)
        _c(L.ttdft_nn_upsample2_backward(h, g["cat1"].ptr, c2, self._dims(1), g["dec2b"].ptr, 0))
        cb("dec2b", a["dec2a"].ptr, a["dec2b"].ptr, g["dec2b"].ptr, g["dec2a"].ptr, 2)
        cb("dec2a", a["cat2"].ptr, a["dec2a"].ptr, g["dec2a"].ptr, g["cat2"].ptr, 2)
        _c(L.ttdft_nn_upsample2_backward(h, g["cat2"].ptr, c3, self._dims(2), g["bot_b"].ptr, 0))
        cb("bot_b", a["bot_a"].ptr, a["bot_b"].ptr, g["bot_b"].ptr, g["bot_a"].ptr, 3)
        cb("bot_a", a["p2"].ptr, a["bot_a"].ptr, g["bot_a"].ptr, g["p2"].ptr, 3)
        _c(L.ttdft_nn_avgpool2_backward(h, g["p2"].ptr, c2, self._dims(2), gs2, 1))
        cb("enc2b", a["enc2a"].ptr, s2, gs2, g["enc2a"].ptr, 2)
        cb("enc2a", a["p1"].ptr, a["enc2a"].ptr, g["enc2a"].ptr, g["p1"].ptr, 2)
        _c(L.ttdft_nn_avgpool2_backward(h, g["p1"].ptr, c1, self._dims(1), gs1, 1))
        cb("enc1b", a["enc1a"].ptr, s1, gs1, g["enc1a"].ptr, 1)
        cb("enc1a", a["p0"].ptr, a["enc1a"].ptr, g["enc1a"].ptr, g["p0"].ptr, 1)
        _c(L.ttdft_nn_avgpool2_backward(h, g["p0"].ptr, c0, self._dims(0), gs0, 1))
        cb("enc0b", a["enc0a"].ptr, s0, gs0, g["enc0a"].ptr, 0)
        cb("enc0a", a["x"].ptr, a["enc0a"].ptr, g["enc0a"].ptr, None, 0)

    def loss_and_grad(self):
        """MSE of a["y"] against a["t"]; writes g["y"]."""
        loss = ctypes.c_double()
        _c(E.lib().ttdft_nn_mse(self.ctx.h, self.a["y"].ptr, self.a["t"].ptr, self.P[0],
                                self.g["y"].ptr, ctypes.byref(loss)))
        return loss.value

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        _c(E.lib().ttdft_nn_adam(self.ctx.h, self.params.ptr, self.grads.ptr, self.m.ptr,
                                 self.v.ptr, self.nparams, lr, beta1, beta2, eps,
                                 self.step_count))

    # ---- persistence ---------------------------------------------------------------
    def get_params(self):
        return self.params.to_host()

    def set_params(self, p):
        self.params.copy_from(np.asarray(p, np.float32))

    def save(self, path, **meta):
        np.savez_compressed(path, params=self.get_params().astype(np.float32),
                            widths=np.array(self.widths), cin=self.cin, n=self.n,
                            in_scale=self.in_scale, out_scale=self.out_scale, arith=self.arith,
                            **meta)

    @classmethod
    def load(cls, ctx, path=DEFAULT_WEIGHTS):
        d = np.load(path, allow_pickle=False)
        arith = int(d["arith"]) if "arith" in d.files else E.NN_FP32
        net = cls(ctx, int(d["n"]), int(d["cin"]), tuple(int(w) for w in d["widths"]),
                  arith=arith)
        net.set_params(d["params"])
        net.in_scale, net.out_scale = float(d["in_scale"]), float(d["out_scale"])
        return net


# ---- inputs and targets on the uniform grid --------------------------------------
# The network sees a fixed cube centred on the origin (the cluster's centre),
# whatever the SCF box: 40 bohr, the box of the training clusters.  Larger
# SCF boxes take the correction inside the cube and zero outside it.
NET_BOX = 40.0


def uniform_points(n, box=NET_BOX):
    """n points per axis spanning the cube (point i at -box/2 + i box/(n-1))."""
    return np.linspace(-box / 2, box / 2, n)


def network_input(net, scf, d_sup_uniform=None):
    """Fills net.a["x"]: [splats, superposition * in_scale] on the n^3 grid."""
    n = net.n
    pts = uniform_points(n)
    h = pts[1] - pts[0]
    d_pts = E.DeviceArray.from_host(scf.ctx, pts)
    d_splat = E.DeviceArray(scf.ctx, (n ** 3,))
    atoms = np.column_stack([scf.xyz, np.ones(scf.natoms)])
    scf.ctx.gaussian_splat(atoms, 1.0, (n, n, n), (pts[0],) * 3, (h,) * 3, d_splat.ptr)
    if d_sup_uniform is None:
        d_sup_uniform = E.DeviceArray(scf.ctx, (n ** 3,))
        scf.superposition(d_sup_uniform, points=d_pts, n=n)
    L, hh = E.lib(), scf.ctx.h
    _c(L.ttdft_nn_to_f32(hh, d_splat.ptr, net.a["x"].ptr, n ** 3, 1.0))
    _c(L.ttdft_nn_to_f32(hh, d_sup_uniform.ptr, net.a["x"].ptr + 4 * n ** 3, n ** 3,
                         net.in_scale))
    return d_sup_uniform


def learned_start(net, scf, d_rho):
    """The learned initial density on the SCF's quadrature grid (into d_rho)."""
    n = net.n
    network_input(net, scf)
    net.forward()
    d_corr = E.DeviceArray(scf.ctx, (n ** 3,))
    _c(E.lib().ttdft_nn_to_f64(scf.ctx.h, net.a["y"].ptr, d_corr.ptr, n ** 3,
                               1.0 / net.out_scale))
    scf.superposition(d_rho)
    pts = uniform_points(n)
    h = pts[1] - pts[0]
    scf.ctx.grid_resample(d_corr.ptr, (n, n, n), (pts[0],) * 3, (h,) * 3, scf.p,
                          *(a.d_xq.ptr for a in scf.axes), d_rho.ptr, add=True)
    scf.ctx.grid_clamp(d_rho.ptr, scf.ngrid, 0.0)
    scf.renormalise(d_rho)


def default_probe_below(path=None):
    """A quarter of the smallest superposition first residual over the
    evaluation clusters in the shipped network's report, or None without the
    report."""
    import json
    path = path or os.path.join(os.path.dirname(DEFAULT_WEIGHTS),
                                "unet_initial_density_report.json")
    try:
        with open(path) as fh:
            test = json.load(fh)["test"]
        return 0.25 * min(c["superposition"]["first_residual"] for c in test)
    except (OSError, KeyError, ValueError, TypeError):
        return None


def run_learned(net, scf, callback=None, probe_below="report"):
    """SCF from the learned start with the superposition fallback: the bar for
    the learned start's first residual is the superposition start's first
    residual, measured by a probe (one SCF iteration from the superposition,
    counted in the time and the iterations).  The probe runs only when the
    learned start's first residual is at least probe_below; below it the
    learned start is kept without one.  probe_below defaults to a quarter of
    the smallest superposition first residual on the evaluation clusters
    (default_probe_below); None always probes."""
    t0 = time.perf_counter()
    if probe_below == "report":
        probe_below = default_probe_below()
    d_start = E.DeviceArray(scf.ctx, (scf.ngrid,))
    learned_start(net, scf, d_start)
    probed = []

    def bar(res):
        if probe_below is not None and res < probe_below:
            return None
        saved, saved_bufs = scf.cfg.max_iter, getattr(scf, "_tucker_bufs", {})
        scf.cfg.max_iter = 1  # the probe: one iteration from the superposition,
        scf._tucker_bufs = {}  # with its own Tucker factors
        try:
            probed.append(scf.run().residuals[0])
        finally:
            scf.cfg.max_iter, scf._tucker_bufs = saved, saved_bufs
        return probed[0]

    r = scf.run(d_start, callback, fallback_residual=bar)
    r.seconds = time.perf_counter() - t0
    r.probed = bool(probed)
    r.iterations += r.probed  # the probe iteration
    return r


# ---- training data and training --------------------------------------------------
def random_cluster(rng, nmin=13, nmax=19, jitter=0.3):
    """An FCC cut of nmin..nmax atoms around a random centre, jittered and
    rotated at random."""
    natoms = int(rng.integers(nmin, nmax + 1))
    base = S.fcc_cluster(200, jitter=0.0)
    c = rng.uniform(-1.9, 1.9, 3)
    d = np.linalg.norm(base - c, axis=1)
    xyz = base[np.argsort(d, kind="stable")[:natoms]]
    xyz = xyz - xyz.mean(axis=0)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    return xyz @ q.T + rng.uniform(-jitter, jitter, xyz.shape)


def converged_sample(ctx, xyz, cfg, n):
    """(splat, superposition, target correction) of one cluster on the n^3 grid,
    float64 host arrays, and the SCF result."""
    s = S.Scf(ctx, xyz, 3.0, cfg)
    r = s.run()
    pts = E.DeviceArray.from_host(ctx, uniform_points(n))
    d_sup = E.DeviceArray(ctx, (n ** 3,))
    s.superposition(d_sup, points=pts, n=n)
    d_conv = s.density_on_uniform(n, box=NET_BOX)
    sup, conv = d_sup.to_host(), d_conv.to_host()
    return sup, conv - sup, r


def train(ctx, samples, n=64, epochs=40, lr=1e-3, seed=0, log=print, arith=E.NN_FP32):
    """samples: list of (xyz, superposition, correction) on the n^3 grid.
    Returns the trained network (scales fitted to the data)."""
    net = UNet(ctx, n, seed=seed, arith=arith)
    sup_max = max(float(np.abs(s).max()) for _, s, _ in samples)
    corr_rms = math.sqrt(np.mean([np.mean(c ** 2) for _, _, c in samples]))
    net.in_scale, net.out_scale = 1.0 / sup_max, 1.0 / corr_rms
    dummy_cfg = S.ScfConfig()
    staged = []
    for xyz, sup, corr in samples:
        d_sup = E.DeviceArray.from_host(ctx, sup)
        d_t = E.DeviceArray.from_host(ctx, corr)
        staged.append((xyz, d_sup, d_t))
    rng = np.random.default_rng(seed)

    class _Geo:  # what network_input needs of an Scf for a sample
        def __init__(self, xyz):
            self.ctx, self.xyz, self.natoms, self.cfg = ctx, xyz, len(xyz), dummy_cfg

    inputs = []
    for xyz, d_sup, d_t in staged:
        network_input(net, _Geo(xyz), d_sup)
        inputs.append(E.DeviceArray.from_host(ctx, net.a["x"].to_host()))
    history = []
    for ep in range(epochs):
        order = rng.permutation(len(staged))
        tot = 0.0
        t0 = time.perf_counter()
        for i in order:
            ctx.copy(net.a["x"].ptr, inputs[i].ptr, inputs[i].nbytes)
            _c(E.lib().ttdft_nn_to_f32(ctx.h, staged[i][2].ptr, net.a["t"].ptr, n ** 3,
                                       net.out_scale))
            net.forward()
            tot += net.loss_and_grad()
            net.backward()
            net.adam(lr)
        ctx.synchronize()
        history.append(tot / len(staged))
        log(f"epoch {ep} loss {history[-1]:.5f} ({time.perf_counter() - t0:.2f} s)")
    net.history = history
    return net


# ---- evaluation and the command line ------------------------------------------------
def data_config(nelec):
    """The SCF configuration of the training and evaluation clusters: the
    13-atom test configuration (40-bohr box, 24 quartic elements per axis,
    8-point rule, rank 16) with enough states for the electron count."""
    return S.ScfConfig(nel=24, quad=8, inner_size=12.0, inner_h=1.0, coarsening=1.25, rank=16,
                       rho_rank=24, veff_rank=24, max_iter=80, tol=1e-8,
                       nstates=max(24, nelec // 2 + 10))


def iterations_to(residuals, tol):
    for i, r in enumerate(residuals):
        if r < tol:
            return i + 1
    return None


def compare_starts(ctx, net, xyz, tol=1e-10):
    """SCF from the superposition and from the learned start (with fallback) on
    one cluster: iterations to 1e-6 and to tol, seconds, energies."""
    nelec = 3 * len(xyz)
    cfg = data_config(nelec)
    cfg.tol = tol
    s0 = S.Scf(ctx, xyz, 3.0, cfg)
    r0 = s0.run()
    s1 = S.Scf(ctx, xyz, 3.0, cfg)
    r1 = run_learned(net, s1)
    to1 = iterations_to(r1.residuals, 1e-6)
    if to1 is not None and r1.probed:
        to1 += 1  # the fallback probe's iteration
    return {"natoms": len(xyz),
            "superposition": {"iterations": r0.iterations, "to_1e-6": iterations_to(r0.residuals, 1e-6),
                              "first_residual": r0.residuals[0], "seconds": r0.seconds,
                              "energy": r0.energy, "converged": r0.converged},
            "learned": {"iterations": r1.iterations, "to_1e-6": to1,
                        "first_residual": r1.residuals[0], "seconds": r1.seconds,
                        "energy": r1.energy, "converged": r1.converged,
                        "fell_back": r1.fell_back},
            "energy_diff_per_atom": abs(r1.energy - r0.energy) / len(xyz)}


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="generate SCF densities, train the initial-density "
                                             "network on the device, evaluate it")
    ap.add_argument("--train", type=int, default=48, help="training clusters")
    ap.add_argument("--test", type=int, default=8, help="held-out clusters")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--n", type=int, default=64, help="uniform grid points per axis")
    ap.add_argument("--nmin", type=int, default=13, help="fewest atoms of a cluster")
    ap.add_argument("--nmax", type=int, default=19, help="most atoms of a cluster")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--math", choices=["fp32", "bf16"], default="bf16",
                    help="convolution arithmetic: FP32 vector math or BF16 MFMA")
    ap.add_argument("--out", required=True, help="weights (.npz)")
    ap.add_argument("--report", required=True, help="evaluation (.json)")
    a = ap.parse_args(argv)
    ctx = E.Context(0)
    rng = np.random.default_rng(a.seed)
    t0 = time.perf_counter()
    samples, failed = [], 0
    while len(samples) < a.train:
        xyz = random_cluster(rng, a.nmin, a.nmax)
        sup, corr, r = converged_sample(ctx, xyz, data_config(3 * len(xyz)), a.n)
        if not r.converged:
            failed += 1
            continue
        samples.append((xyz, sup, corr))
        print(f"sample {len(samples)}: {len(xyz)} atoms, {r.iterations} SCF iterations, "
              f"{r.seconds:.2f} s", flush=True)
    t_data = time.perf_counter() - t0
    arith = E.NN_BF16 if a.math == "bf16" else E.NN_FP32
    net = train(ctx, samples, a.n, a.epochs, a.lr, seed=a.seed,
                log=lambda m: print(m, flush=True), arith=arith)
    t_train = time.perf_counter() - t0 - t_data
    net.save(a.out, seed=a.seed, train_clusters=a.train, epochs=a.epochs)
    test_rng = np.random.default_rng(a.seed + 1)
    evals = []
    for i in range(a.test):
        res = compare_starts(ctx, net, random_cluster(test_rng, a.nmin, a.nmax))
        evals.append(res)
        print(json.dumps(res), flush=True)
    sup_it = [e["superposition"]["to_1e-6"] or 80 for e in evals]
    lrn_it = [e["learned"]["to_1e-6"] or 80 for e in evals]
    report = {"train_clusters": a.train, "atoms": [a.nmin, a.nmax], "failed_scf": failed,
              "epochs": a.epochs,
              "math": a.math,
              "data_seconds": t_data, "train_seconds": t_train, "loss_history": net.history,
              "params": net.nparams, "test": evals,
              "mean_iterations_to_1e-6": {"superposition": float(np.mean(sup_it)),
                                          "learned": float(np.mean(lrn_it))},
              "mean_seconds": {"superposition": float(np.mean([e["superposition"]["seconds"]
                                                               for e in evals])),
                               "learned": float(np.mean([e["learned"]["seconds"]
                                                         for e in evals]))},
              "max_energy_diff_per_atom": max(e["energy_diff_per_atom"] for e in evals)}
    with open(a.report, "w") as fh:
        json.dump(report, fh, indent=1)
    print(json.dumps({k: v for k, v in report.items() if k not in ("test", "loss_history")}),
          flush=True)


if __name__ == "__main__":
    main()
