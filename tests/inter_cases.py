"""Crafted inter-atom states shared by the oracle and device tests.

An inter atom is an atom that left its lattice site (the reference's
InterAtomList).  `crafted_inter_state` empties chosen sites of a built world
the way decide() does (type -1, v = 0, the stale position is where the atom
went) and returns the inter list: atoms displaced from corner, face, edge and
interior sites of the sub-box, some pushed across a box face so that
exchangeInter has to migrate them, and extra interstitials with fresh ids on
octahedral sites next to the faces, so that borderInter's ghost images of
inter atoms are exercised in every direction.
"""
import numpy as np

import oracle_py as O

A = 2.855


def _site_x(i, j, k):
    return np.array([i * 0.5 * A, (j + (i % 2) * 0.5) * A, (k + (i % 2) * 0.5) * A])


def crafted_inter_state(sim, seed, n_random=8, n_octa=4, kick=1.2):
    cells = sim.cells
    sb = sim.sub_box()
    rng = np.random.default_rng(seed)
    n0, n1, n2 = cells
    X = 2 * n0
    chosen = [
        ((0, 0, 0), None),
        ((X - 1, n1 - 1, n2 - 1), None),
        ((0, n1 // 2, n2 // 2), (-1.6, 0.0, 0.0)),     # pushed below x = 0
        ((X - 2, n1 - 1, n2 // 2), (0.0, 1.6, 0.0)),   # pushed above y = L
        ((X // 2, n1 // 2, 0), (0.0, 0.0, -1.6)),      # pushed below z = 0
        ((X // 2 + 1, n1 // 2, n2 // 2), None),
        ((1, 0, n2 - 1), None),
    ]
    used = {c for c, _ in chosen}
    while len(chosen) < 7 + n_random:
        c = (int(rng.integers(X)), int(rng.integers(n1)), int(rng.integers(n2)))
        if c not in used:
            used.add(c)
            chosen.append((c, None))
    inter = []
    for (i, j, k), d in chosen:
        at = sb[k, j, i].copy()
        if d is None:
            u = rng.normal(size=3)
            d = u / np.linalg.norm(u) * rng.uniform(0.7, kick)
        at["x"] = at["x"] + np.asarray(d)
        inter.append(at)
        sb[k, j, i]["type"] = -1
        sb[k, j, i]["v"] = 0.0
        sb[k, j, i]["x"] = at["x"]
    next_id = int(sb["id"].max()) + 1
    octa = [(0, 0, 0, (0.5 * A, 0.0, 0.5 * A)), (X - 2, n1 - 1, 0, (0.5 * A, 0.5 * A, 0.0)),
            (0, n1 - 1, n2 - 1, (0.0, 0.5 * A, 0.5 * A)), (X // 2, 0, n2 // 2, (0.5 * A, 0.0, 0.0))]
    for q in range(n_octa):
        i, j, k, off = octa[q % len(octa)]
        at = np.zeros(1, dtype=O.ATOM_DTYPE)[0]
        at["id"] = next_id + q
        at["type"] = q % 3 if sim_has_alloy(sim) else 0
        at["x"] = _site_x(i, j, k) + np.asarray(off) + rng.uniform(-0.05, 0.05, 3)
        at["v"] = rng.normal(0.0, 3.0, 3)
        inter.append(at)
    out = np.zeros(len(inter), dtype=O.ATOM_DTYPE)
    for q, at in enumerate(inter):
        out[q] = at
    out["f"] = 0.0
    out["rho"] = 0.0
    out["df"] = 0.0
    return out


def sim_has_alloy(sim):
    t = sim.sub_box()["type"]
    return bool(np.any(t == 1) or np.any(t == 2))


def oracle_forces_with_inter(sim, inter):
    """Load `inter` into the oracle and evaluate forces the way the step does
    after decide: exchangeInter, borderInter, exchangeAtom, computeEam."""
    sim.set_inter(inter)
    sim.phase("exchange_atom_first")
    sim.phase("exchange_inter")
    sim.phase("border_inter")
    sim.phase("exchange_atom")
    sim.phase("clear_force")
    sim.phase("compute_eam")
