"""The device's ownership decisions against the hand-built Wigner-Seitz
scenarios of tests/decide_cases.py (needs an MI355X).

Each scenario's state goes to the device as the reference hands over a state
(ghost-extended AtomElement array and inter list, atom.cpp:21-84 reads
both); the step length is so short (1e-200 ps) that no position moves by
a single rounding step, so the device's first step decides on exactly the
positions of the scenario.  After one step the sites, the vacancies and the
inter list (ids, species, positions, order) must be what BCC geometry and
the reference's rule give, with exchangeInter's periodic shift applied
(inter_particle_packer.cpp:17-121).  The second step's decisions run in the
force pass's fused first half-step; after it the device must still agree
with the oracle stepped the same way, bit for bit.
"""
import numpy as np
import pytest

import misa_eam as M
import oracle_py as O
from decide_cases import A, CASES, CELLS, build, check

pytestmark = pytest.mark.gpu

F = 5.6 / 2.855
DT = 1e-200  # ps: every displacement dt * v underflows below one ulp of x


def _oracle(setfl, case):
    sim = O.Sim(CELLS, A, F, setfl, dt=DT)
    sim.build_world(1, 0.0, (985, 10, 5), 1)
    inter, named = build(sim, case)
    sim.set_inter(inter)
    sim.phase("exchange_atom_first")
    return sim, inter, named


@pytest.mark.parametrize("mode", [M.PREC_FP64, M.PREC_MIXED_KAHAN], ids=["fp64", "mixed"])
@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_device_decide_scenario(setfl_fecuni, case, mode):
    print("decide scenario:", case.name, "precision", mode, flush=True)
    sim, inter, named = _oracle(setfl_fecuni, case)
    dev = M.EamDevice(CELLS)
    try:
        dev.make_offsets()
        dev.set_potential(setfl_fecuni)
        dev.upload(sim.atoms().copy(), inter)
        dev.set_precision(mode)
        dev.set_timestep(DT)
        dev.compute_forces()
        dev.step(1)
        atoms, got_inter = dev.download()
        sub = M.sub_box_view(atoms, dev)
        bad = check(sub, got_inter, case, named, shifted=True)
        assert not bad, "%s (%s): %s" % (case.name, case.why, "; ".join(bad))
        # second step: decisions inside the fused force pass; the oracle steps alike
        dev.step(1)
        atoms, got_inter = dev.download()
        sub = M.sub_box_view(atoms, dev)
        sim.prepare()
        sim.run(2)
        want_sub, want_inter = sim.sub_box(), sim.inter()
        assert np.array_equal(sub["id"], want_sub["id"])
        assert np.array_equal(sub["type"], want_sub["type"])
        live = want_sub["type"] >= 0
        assert np.array_equal(sub["x"][live], want_sub["x"][live])
        assert [int(i) for i in got_inter["id"]] == [int(i) for i in want_inter["id"]]
        assert np.array_equal(got_inter["type"], want_inter["type"])
        assert np.array_equal(got_inter["x"], want_inter["x"])
    finally:
        dev.close()
