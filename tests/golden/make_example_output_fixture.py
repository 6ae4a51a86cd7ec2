"""Extract the golden vectors printed in the reference's example/output.

The reference ships no tests; its only recorded results are the stdout of one
converged 13-atom run (example/output).  For every SCF iteration this script
collects the Ritz values that fed the Fermi-level solve, the Fermi level, the
24 occupancies, the band energy, the total energy and the charge (plus the
projector normalisations 1 / C_lm printed once at set-up), and writes
them to ``example_output_scf.json`` next to this file.  The values are data
(inputs and expected outputs), printed by the reference with %.18e.

Run in a container that has the reference checked out:
    python tests/golden/make_example_output_fixture.py /root/reference/example/output
"""
import json
import os
import re
import sys


def parse_c_lm(path):
    """The four "C_lm:" lines printed before the first rank (example/output:48-51):
    1 / C_lm of the s and the three p projectors, atoms/NonLocalPSPData.cc:640-642."""
    out = []
    with open(path) as fh:
        for ln in fh:
            m = re.match(r"C_lm: ([-+0-9.eE]+)", ln)
            if m:
                out.append(float(m.group(1)))
    return out


def parse(path):
    scfs = []
    cur = None
    evals = {}
    with open(path) as fh:
        lines = fh.read().splitlines()
    for ln in lines:
        m = re.match(r"\s*Entering SCF (\d+)", ln)
        if m:
            cur = {"scf": int(m.group(1)), "passes": [], "upper_bounds": []}
            scfs.append(cur)
            evals = {}
            continue
        if cur is None:
            continue
        m = re.match(r"upperBoundUnWantedSpectrum is ([-+0-9.eE]+)\.", ln)
        if m:
            cur["upper_bounds"].append(float(m.group(1)))
            continue
        m = re.match(r"(\d+)th e-vals:\s+([-+0-9.eE]+)", ln)
        if m:
            evals[int(m.group(1))] = float(m.group(2))
            continue
        if ln.startswith("Projected Eigenvalues of Chebyshev Pass"):
            cur["passes"].append([evals[i] for i in range(len(evals))])
            evals = {}
            continue
        m = re.match(r"\d+ :Fermi Energy: ([-+0-9.eE]+)", ln)
        if m:
            cur["fermi"] = float(m.group(1))
            cur["occupancy"] = []
            continue
        m = re.match(r"Occupancy Factor (\d+)\s+([-+0-9.eE]+)", ln)
        if m:
            cur["occupancy"].append(float(m.group(2)))
            continue
        m = re.match(r"band energy: ([-+0-9.eE]+)", ln)
        if m:
            cur["band_energy"] = float(m.group(1))
            continue
        m = re.match(r"ground state energy : ([-+0-9.eE]+)", ln)
        if m:
            cur["total_energy"] = float(m.group(1))
            continue
        m = re.match(r"charge out: ([-+0-9.eE]+)", ln)
        if m:
            cur["charge_out"] = float(m.group(1))
            continue
        m = re.match(r"square L2 diff norm: ([-+0-9.eE]+)", ln)
        if m:
            cur["rho_residual"] = float(m.group(1))
    return scfs


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/example/output"
    scfs = parse(src)
    fixture = {
        "source": "example/output of the reference (13-atom cluster, 24 states, T = 500 K)",
        "nelec": 39,
        "kb": 3.166811429e-06,
        "temperature": 500.0,
        "scf": scfs,
        "inv_c_lm": parse_c_lm(src),
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example_output_scf.json")
    with open(out, "w") as fh:
        json.dump(fixture, fh, indent=1)
    print(f"wrote {len(scfs)} SCF iterations to {out}")


if __name__ == "__main__":
    main()
