"""The trained initial-density network (ttdft_public_amd/data/
unet_initial_density.npz, trained on the device by ttdft_public_amd.learned on
48 SCF-converged 13-19-atom clusters) against the atomic superposition on
held-out clusters: the learned start's first SCF residual is lower, the SCF
from it converges to the same energy (1e-5 Ha/atom) in fewer iterations, and
the fallback rule sends a bad start back to the superposition."""
import numpy as np
import pytest

from ttdft_public_amd import engine as E
from ttdft_public_amd import learned as Ln
from ttdft_public_amd import scf as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net(ctx):
    return Ln.UNet.load(ctx)


def test_learned_start_beats_superposition(ctx, net):
    rng = np.random.default_rng(777)  # clusters outside the training and evaluation seeds
    gains = []
    for i in range(2):
        xyz = Ln.random_cluster(rng)
        print(f"held-out cluster {i}: {len(xyz)} atoms", flush=True)
        res = Ln.compare_starts(ctx, net, xyz)
        sup, lrn = res["superposition"], res["learned"]
        assert sup["converged"] and lrn["converged"] and not lrn["fell_back"]
        assert lrn["first_residual"] < 0.2 * sup["first_residual"]
        assert res["energy_diff_per_atom"] < 1e-5
        gains.append(sup["iterations"] - lrn["iterations"])
    assert sum(gains) > 0, gains


def test_fallback_to_superposition(ctx, net):
    """A start far from the answer (the learned correction scaled up 40 times)
    fails the first-residual bar and the loop restarts from the
    superposition."""
    xyz = S.cuboctahedron(seed=5)
    cfg = Ln.data_config(3 * len(xyz))
    s = S.Scf(ctx, xyz, 3.0, cfg)
    r0 = S.Scf(ctx, xyz, 3.0, S.ScfConfig(**{**cfg.__dict__, "max_iter": 1})).run().residuals[0]
    bad = E.DeviceArray(ctx, (s.ngrid,))
    saved = net.out_scale
    net.out_scale = saved / 40.0
    try:
        Ln.learned_start(net, s, bad)
    finally:
        net.out_scale = saved
    r = s.run(bad, fallback_residual=r0)
    assert r.fell_back and r.converged
    assert r.residuals[0] > r0
