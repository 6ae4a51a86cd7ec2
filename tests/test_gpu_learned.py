"""The trained initial-density network (ttdft_public_amd/data/
unet_initial_density.npz, trained on the device by ttdft_public_amd.learned on
48 SCF-converged 13-19-atom clusters) against the atomic superposition on
held-out clusters: the learned start's first SCF residual is lower, the SCF
from it converges to the same energy (1e-5 Ha/atom) in fewer iterations, and
the fallback rule sends a start that misses the bar back to the
superposition."""
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
    """The fallback rule: a first residual above the bar restarts the loop from
    the superposition, counting the wasted iteration.  (The superposition's
    own first residual is about 40 times the learned start's, so the bar is
    set below the learned start's residual to trigger the rule.)"""
    xyz = S.cuboctahedron(seed=5)
    cfg = Ln.data_config(3 * len(xyz))
    s = S.Scf(ctx, xyz, 3.0, cfg)
    start = E.DeviceArray(ctx, (s.ngrid,))
    Ln.learned_start(net, s, start)
    r = s.run(start, fallback_residual=1e-3)
    ref = S.Scf(ctx, xyz, 3.0, cfg).run()
    assert r.fell_back and r.converged
    assert r.residuals[0] > 1e-3
    assert abs(r.iterations - (ref.iterations + 1)) <= 1
    assert abs(r.energy - ref.energy) < 1e-5 * len(xyz)


def test_probe_only_when_in_doubt(ctx, net):
    """The superposition probe runs only when the learned start's first
    residual is not below probe_below; forced (probe_below None) it costs one
    counted iteration and leads to the same converged energy."""
    xyz = S.cuboctahedron(seed=6)
    cfg = Ln.data_config(3 * len(xyz))
    bar = Ln.default_probe_below()
    assert bar is not None and bar > 0
    r = Ln.run_learned(net, S.Scf(ctx, xyz, 3.0, cfg))
    assert r.residuals[0] < bar and not r.probed and not r.fell_back
    rp = Ln.run_learned(net, S.Scf(ctx, xyz, 3.0, cfg), probe_below=None)
    assert rp.probed and not rp.fell_back
    assert rp.iterations == r.iterations + 1
    assert rp.residuals[0] == pytest.approx(r.residuals[0], rel=1e-8)
    assert abs(rp.energy - r.energy) < 1e-7 * len(xyz)
