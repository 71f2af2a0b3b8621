"""The oracle's decide (src/atom.cpp:21-84) against hand-built Wigner-Seitz
scenarios whose outcomes follow from BCC geometry (tests/decide_cases.py),
and proof that the scenarios catch the plausible mistakes: each mistake is
planted in the oracle's decide in turn and must fail at least one scenario.
"""
import math

import numpy as np
import pytest

import oracle_py as O
from decide_cases import A, CASES, CELLS, build, check

F = 5.6 / 2.855


def _sim(setfl):
    sim = O.Sim(CELLS, A, F, setfl)
    sim.build_world(1, 0.0, (985, 10, 5), 1)
    return sim


def _run_decide(sim, case):
    inter, named = build(sim, case)
    sim.set_inter(inter)
    sim.phase("exchange_atom_first")  # ghost slots as the previous step's halo leaves them
    flag = sim.phase("decide")
    return named, flag


def test_tie_is_exact():
    """The exact-0.2a case needs x*x to equal pow(0.2a, 2) (both sides of
    the reference's comparison, atom.cpp:42)."""
    x = 0.2 * A
    assert x * x == math.pow(0.2 * A, 2.0)


@pytest.mark.parametrize("case", CASES, ids=[c.name for c in CASES])
def test_decide_scenario(setfl_fecuni, case):
    sim = _sim(setfl_fecuni)
    named, flag = _run_decide(sim, case)
    bad = check(sim.sub_box(), sim.inter(), case, named)
    assert not bad, "%s (%s): %s" % (case.name, case.why, "; ".join(bad))
    # decide returns 1 when any lattice atom was taken off (atom.cpp:46)
    moved_far = any(np.dot(o, o) * A * A > math.pow(0.2 * A, 2.0) for o in map(np.asarray, case.moves.values()))
    assert flag == int(moved_far)
    # exchangeInter then moves the atoms outside the box by one box length
    sim.phase("exchange_inter")
    bad = check(sim.sub_box(), sim.inter(), case, named, shifted=True)
    assert not bad, "%s after exchangeInter: %s" % (case.name, "; ".join(bad))


@pytest.mark.parametrize("plant, what", [(1, "threshold 0.2a left unsquared"), (2, ">= in place of >"),
                                         (3, "refill without the box test on the vacancy's position"),
                                         (4, "the last claimant wins"), (5, "refills into ghost sites")])
def test_planted_mistakes_are_caught(setfl_fecuni, plant, what):
    caught = []
    for case in CASES:
        sim = _sim(setfl_fecuni)
        sim.plant(plant)
        named, _ = _run_decide(sim, case)
        if check(sim.sub_box(), sim.inter(), case, named):
            caught.append(case.name)
    assert caught, "planted mistake '%s' passes every scenario" % what
