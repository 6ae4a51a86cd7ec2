"""The reference-shaped C++ layer (ChebyShevFilter / EigenSolver / AX /
MultTM over the C ABI) and its ChFSI driver, which follows main.cc:977-1237."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from ttdft_public_amd import engine as E

PKG = os.path.join(os.path.dirname(os.path.dirname(__file__)), "ttdft_public_amd")
DRIVER = os.path.join(PKG, "chfsi_driver")


def parse_driver(out):
    """Per SCF iteration: the Ritz values of its last pass and the Fermi level."""
    scfs = []
    cur = {}
    for ln in out.splitlines():
        m = re.match(r"(\d+)th e-vals:\s+(\S+)", ln)
        if m:
            cur[int(m.group(1))] = float(m.group(2))
        m = re.match(r"0 :Fermi Energy: (\S+)", ln)
        if m:
            scfs.append((np.array([cur[i] for i in range(len(cur))]), float(m.group(1))))
            cur = {}
    return scfs


def test_driver_fails_cleanly_without_device():
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    try:
        ndev = E.device_count()
    except E.TtdftError:
        ndev = 0
    if ndev > 0:
        pytest.skip("a GPU is present")
    r = subprocess.run([DRIVER, "--scf", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "ttdft status" in r.stderr


@pytest.mark.gpu
def test_driver_matches_oracle_chfsi():
    dims, bands, p, k, m, nelec, nscf = (8, 8, 8), (2, 2, 3), 5, 16, 12, 19, 5
    r = subprocess.run([DRIVER, "--scf", str(nscf), "--k", str(k), "--m", str(m), "--nelec",
                        str(nelec)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = parse_driver(r.stdout)
    assert len(got) == nscf
    s = E.SynthHam(*dims, *bands, p)
    H = oracle.Ham.from_synth(s)
    X = oracle.orth(oracle.rng_normal(5, s.n * k).reshape(s.n, k))
    ev = np.zeros(k)
    st = oracle.new_state()
    occ_index = min(int(nelec / 2.0 + 10), k - 1)
    for scf in range(nscf):
        oracle.chfsi_step(H, X, ev, st, m=m, scf_iter=scf, occ_index=occ_index,
                          restart_times=10 if scf == 0 else 3)
        evg, efg = got[scf]
        nocc = (nelec + 1) // 2
        assert np.abs(evg[: nocc + 1] - ev[: nocc + 1]).max() < 1e-9, scf
        assert np.abs(evg - ev).max() < 1e-6, scf
        assert abs(efg - oracle.fermi_energy(ev, nelec)) < 1e-9
