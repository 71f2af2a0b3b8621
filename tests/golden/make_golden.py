"""Writes the golden fixtures of this directory from the oracle.

The reference ships no golden vectors for the EAM step, and its sources do
not build here (pot, libcomm and kiwi are absent), so the fixtures are the
oracle's own outputs on small seeded inputs.  They pin the oracle against
later changes (tests/test_golden.py re-runs it bit for bit) and give the
device tests fixed targets.

    python tests/golden/make_golden.py      # rewrites the .npz files
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "misa-md_amd"), os.path.dirname(HERE)]

import oracle_py as O  # noqa: E402
import setfl_gen  # noqa: E402
from inter_cases import crafted_inter_state, oracle_forces_with_inter  # noqa: E402

A = 2.855
F = 5.6 / 2.855
FIELDS = ("id", "type", "x", "v", "f", "rho", "df")


def lattice_case(setfl):
    """4^3 FeCuNi at 600 K: the first force evaluation and 10 steps of 1 fs."""
    sim = O.Sim((4, 4, 4), A, F, setfl)
    sim.build_world(7, 600.0, (90, 6, 4), 7)
    init = sim.atoms().copy()
    sim.prepare()
    first = sim.sub_box().copy()
    sim.run(10)
    last = sim.sub_box().copy()
    out = {"cells": np.array([4, 4, 4]), "init_aos": init}
    for k in FIELDS:
        out["first_" + k] = first[k]
        out["last_" + k] = last[k]
    return out


def inter_case(setfl):
    """6x7x8 FeCuNi with 19 inter atoms (crafted): forces, then 12 steps."""
    cells = (6, 7, 8)
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(23, 300.0, (985, 10, 5), 23)
    sb = sim.sub_box()
    rng = np.random.default_rng(23)
    sb["x"] = sb["x"] + rng.uniform(-0.08, 0.08, sb["x"].shape)
    inter = crafted_inter_state(sim, 23)
    init = sim.atoms().copy()
    oracle_forces_with_inter(sim, inter)
    first, first_inter = sim.sub_box().copy(), sim.inter()
    sim.run(12)
    last, last_inter = sim.sub_box().copy(), sim.inter()
    out = {"cells": np.array(cells), "init_aos": init, "init_inter": inter}
    for k in FIELDS:
        out["first_" + k] = first[k]
        out["last_" + k] = last[k]
        out["first_inter_" + k] = first_inter[k]
        out["last_inter_" + k] = last_inter[k]
    return out


def write_setfl(path):
    setfl_gen.write_setfl(path, ("Fe", "Cu", "Ni"))


def main():
    setfl = os.path.join(HERE, "_fecuni.eam.alloy")
    write_setfl(setfl)
    try:
        np.savez_compressed(os.path.join(HERE, "fecuni_4c_10steps.npz"), **lattice_case(setfl))
        np.savez_compressed(os.path.join(HERE, "fecuni_inter_12steps.npz"), **inter_case(setfl))
    finally:
        os.remove(setfl)


if __name__ == "__main__":
    main()
