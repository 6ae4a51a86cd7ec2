"""Extract the exponential-sum expansion of the Coulomb kernel that the
reference's example ships (example/alphak35_1e8, example/omegak35_1e8, 35
values each, one per line) and its A^2 (example/input.inp, Asquare) into
``kernel_expansion_k35.json`` next to this file.  The values are data.

The reference reads "alpha file" into the exponents and "omega file" into
the weights and scales them by 1/A^2 and 1/A (dft/KSDFTPotential.cc:273-284).
With exponents from alphak35_1e8 and weights from omegak35_1e8 the sum
approximates 1/r to about 7e-6 relative for A <= r <= 1e4 A; the example deck
names the files the other way round.  Both files are stored under their own
names.

Run in a container that has the reference checked out:
    python tests/golden/make_kernel_expansion_fixture.py /root/reference/example
"""
import json
import os
import sys


def read_column(path):
    with open(path) as fh:
        return [float(tok) for tok in fh.read().split()]


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/example"
    asq = None
    with open(os.path.join(src, "input.inp")) as fh:
        for ln in fh:
            if ln.startswith("Asquare="):
                asq = float(ln.split("=", 1)[1])
    fixture = {
        "source": "example/alphak35_1e8, example/omegak35_1e8, example/input.inp (Asquare)",
        "alphak35_1e8": read_column(os.path.join(src, "alphak35_1e8")),
        "omegak35_1e8": read_column(os.path.join(src, "omegak35_1e8")),
        "Asquare": asq,
    }
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kernel_expansion_k35.json")
    with open(out, "w") as fh:
        json.dump(fixture, fh, indent=1)


if __name__ == "__main__":
    main()
