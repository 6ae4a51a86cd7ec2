"""Copies input data files of the reference's example deck (example/) into
tests/golden/example/: the 1-D basis functions of the three axes, the atom file, the
three pseudopotential tables of the atom type and its radial density.  They
are data (numbers the example run reads), used as fixtures for the readers,
the 1-D matrices and the pseudopotential set-up.

    python tests/golden/copy_example_inputs.py /root/reference
"""
import os
import shutil
import sys

FILES = ["basis_x.txt", "basis_y.txt", "basis_z.txt", "shell1", "nlpV_AT0", "nlpWaveFun_AT0",
         "locPotential_AT0", "Density_AT0"]


def main(ref):
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "example")
    os.makedirs(out, exist_ok=True)
    for f in FILES:
        shutil.copyfile(os.path.join(ref, "example", f), os.path.join(out, f))
        os.chmod(os.path.join(out, f), 0o644)
    print("copied", FILES, "to", out)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
