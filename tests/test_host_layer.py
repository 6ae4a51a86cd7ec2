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


def run_driver_against_oracle(args, H, n, k, m, nelec, nscf):
    """Run the driver, then the oracle's ChFSI steps on H from the same start
    block, and compare each SCF iteration's Ritz values and Fermi level."""
    r = subprocess.run([DRIVER, "--scf", str(nscf), "--k", str(k), "--m", str(m), "--nelec",
                        str(nelec)] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = parse_driver(r.stdout)
    assert len(got) == nscf
    X = oracle.orth(oracle.rng_normal(5, n * k).reshape(n, k))
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
    return r.stdout


@pytest.mark.gpu
def test_driver_matches_oracle_chfsi():
    dims, bands, p, k, m, nelec, nscf = (8, 8, 8), (2, 2, 3), 5, 16, 12, 19, 5
    s = E.SynthHam(*dims, *bands, p)
    run_driver_against_oracle([], oracle.Ham.from_synth(s), s.n, k, m, nelec, nscf)


def write_assembly_file(path, hl, proj, inv, trunc, c_trunc):
    """The driver's --assembly input (layout in chfsi_driver.cc)."""
    X, DX, Y, DY, Z, DZ, PX, PY, PZ, G = hl
    cores, mx, my, mz = proj
    with open(path, "wb") as f:
        np.array([X.shape[0], Y.shape[0], Z.shape[0], PX.shape[0], PY.shape[0], PZ.shape[0],
                  len(cores), 0], np.int32).tofile(f)
        np.array([c.shape[0] for c in cores], np.int32).tofile(f)
        np.array([trunc, c_trunc], np.float64).tofile(f)
        for a in (X, DX, Y, DY, Z, DZ, PX, PY, PZ):
            np.ascontiguousarray(a, np.float64).tofile(f)
        np.ascontiguousarray(np.transpose(G, (2, 1, 0))).tofile(f)  # first index fastest
        np.concatenate([np.transpose(c, (2, 1, 0)).reshape(-1) for c in cores]).tofile(f)
        for ms in (mx, my, mz):
            np.concatenate([m.T.reshape(-1) for m in ms]).tofile(f)
        np.asarray(inv, np.float64).tofile(f)


@pytest.mark.gpu
@pytest.mark.parametrize("device", [False, True])
def test_driver_on_assembled_hamiltonian(tmp_path, device):
    """ChFSI on H built by ProjectHamiltonianSparse::Create_Hloc / Create_Cnloc
    (host CSR) or AX::assemble_system (device memory), against the oracle's
    ChFSI on the oracle's own assembly of the same inputs."""
    from test_gpu_assembly import hloc_inputs, proj_inputs

    dims, rv, k, m, nelec, nscf = (9, 8, 7), (3, 4, 2), 16, 12, 19, 4
    hl = hloc_inputs(31, dims, rv, 2, sym=True)
    proj = proj_inputs(32, dims, [2, 3, 1])
    inv = np.array([0.8, -0.4, 1.3])
    path = str(tmp_path / "assembly.bin")
    write_assembly_file(path, hl, proj, inv, 1e-6, 1e-8)
    rp, c, v = oracle.assemble_hloc(*hl, 1e-6)
    crp, cc, cv = oracle.assemble_projectors(*proj, 1e-8)
    n = int(np.prod(dims))
    H = oracle.Ham(n, rp, c, v, crp, cc, cv, inv)
    args = ["--assembly", path] + (["--device-assembly"] if device else [])
    out = run_driver_against_oracle(args, H, n, k, m, nelec, nscf)
    assert ("assembled H on the device" in out) == device


@pytest.mark.gpu
def test_driver_anderson_mixing_matches_oracle(tmp_path):
    """ttdft_host::AndersonMixing over an 8-step density sequence with history
    4 (entries dropped), against the oracle's mixing of the same sequence."""
    rng = np.random.default_rng(21)
    ng, nn, nsteps, hmax, alpha = 5000, 700, 8, 4, 0.5
    jac, wts = rng.uniform(0.5, 2.0, ng), rng.uniform(0.1, 1.0, ng)
    seq = [(rng.uniform(-0.2, 1.5, nn), rng.uniform(-0.2, 1.5, ng),
            rng.uniform(0.5, 1.5, nn), rng.uniform(0.5, 1.5, ng)) for _ in range(nsteps)]
    inp, outp = str(tmp_path / "mix_in.bin"), str(tmp_path / "mix_out.bin")
    with open(inp, "wb") as f:
        np.array([ng, nn], np.int64).tofile(f)
        np.array([nsteps, hmax], np.int32).tofile(f)
        np.array([alpha]).tofile(f)
        jac.tofile(f)
        wts.tofile(f)
        for step in seq:
            for a in step:
                a.tofile(f)
    r = subprocess.run([DRIVER, "--mix", inp, outp], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = np.fromfile(outp).reshape(nsteps, nn + ng)
    hist = []
    for s, (ni, gi, no, go) in enumerate(seq):
        if s == 0:
            en, eg = np.abs(ni), np.abs(gi)
        elif s == 1:
            en = alpha * hist[-1][2] + (1 - alpha) * hist[-1][0]
            eg = alpha * hist[-1][3] + (1 - alpha) * hist[-1][1]
        else:
            c = oracle.anderson_constants(np.stack([e[1] for e in hist]),
                                          np.stack([e[3] for e in hist]), jac, wts)
            en = oracle.anderson_mix(np.stack([e[0] for e in hist]),
                                     np.stack([e[2] for e in hist]), c, alpha)
            eg = oracle.anderson_mix(np.stack([e[1] for e in hist]),
                                     np.stack([e[3] for e in hist]), c, alpha)
        assert np.abs(got[s, :nn] - en).max() <= 1e-11 * np.abs(en).max(), s
        assert np.abs(got[s, nn:] - eg).max() <= 1e-11 * np.abs(eg).max(), s
        hist.append((got[s, :nn], got[s, nn:], no, go))
        hist = hist[-hmax:]
