"""The reference's example deck end to end on the device (ttdft_public_amd.example_deck):
13 atoms, tabulated pseudopotentials, the three file bases of rank 40, main.cc's SCF settings
(example/input.inp), against the converged iteration the reference printed (example/output,
SCF 23: Ritz values, Fermi level, band energy, total energy, charge).

What cannot agree exactly, and the bounds set from it:
  * the reference recompresses rho_out by STHOSVD with a 1e-4 tolerance (its charge out is
    38.99965 against our exact grid sum 38.99977), and its H_loc / C_nloc truncations were
    command-line arguments that the output does not record (we use 1e-8);
  * its C_lm come from a rank-25 Tucker form (3e-8 relative to ours).
Observed (converged in 10 iterations): Ritz values within 5.8e-4 Ha (the lowest within 1.6e-4),
the Fermi level within 2.4e-4 Ha, the band energy within 8.9e-3 Ha and the total energy within
6.4e-3 Ha (0.5 mHa per atom).  Bounds: Ritz values and Fermi level 1e-3 Ha, band energy
1.5e-2 Ha, total energy 1e-2 Ha, charge 1e-3."""
import json
import os

import numpy as np
import pytest

from ttdft_public_amd.example_deck import ExampleDeck

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EX = os.path.join(HERE, "golden", "example")


def test_example_deck_matches_reference_output(ctx):
    with open(os.path.join(HERE, "golden", "example_output_scf.json")) as fh:
        ref = json.load(fh)
    last_ref = ref["scf"][-1]
    print("example deck: 13 atoms, rank 40^3, 24 states, to residual 5e-7", flush=True)
    deck = ExampleDeck(ctx, EX)
    assert np.max(np.abs(1.0 / deck.c_lm / np.array(ref["inv_c_lm"]) - 1.0)) < 1e-7

    def show(rec):
        print(f"scf {rec['iter']}: E {rec['total_energy']:.8f} res {rec['residual']:.2e} "
              f"{rec['seconds']:.1f} s", flush=True)

    out = deck.run(show)
    assert out["converged"]
    last = out["records"][-1]
    ev_ref = np.array(last_ref["passes"][-1])
    ev = np.array(last["evals"])
    assert ev.shape == ev_ref.shape == (24,)
    print(f"max |d eps| {np.max(np.abs(ev - ev_ref)):.2e}, d E_F "
          f"{last['fermi'] - last_ref['fermi']:.2e}, d E_band "
          f"{last['band_energy'] - last_ref['band_energy']:.2e}, d E "
          f"{last['total_energy'] - last_ref['total_energy']:.2e}", flush=True)
    assert np.max(np.abs(ev - ev_ref)) < 1e-3
    assert abs(last["fermi"] - last_ref["fermi"]) < 1e-3
    assert abs(last["band_energy"] - last_ref["band_energy"]) < 1.5e-2
    assert abs(last["total_energy"] - last_ref["total_energy"]) < 1e-2
    assert abs(last["charge_out"] - last_ref["charge_out"]) < 1e-3
