"""Hand-built ownership scenarios for decide (src/atom.cpp:21-84) on a 4^3-cell
BCC box, shared by the oracle test (CPU) and the device test (GPU).

Every expected outcome below is written out by hand from BCC geometry and
from the rule as the reference states it, not computed by any
implementation:

* an atom is taken off its site when its squared distance from the ideal
  site is strictly greater than (0.2 a)^2 (atom.cpp:40-42);
* the atoms taken off are appended to the inter list in the sub-box loop
  order, z, then y, then doubled x (atom.cpp:28-30);
* then every inter atom, in list order, refills the site of its
  Wigner-Seitz cell when that site is in the sub-box, is vacant, and the
  vacancy's own (stale) position is in the sub-box (atom.cpp:58-83,
  ws_utils.cpp:66-89, 131-152).

The Wigner-Seitz cell of a BCC site is a truncated octahedron: its square
faces lie 0.5 a away along <100> (bisecting the next corner-to-corner
vector) and its hexagonal faces sqrt(3)/4 a = 0.433 a away along <111>
(bisecting the corner-to-centre vector).  So an atom moved by 0.42 a along
[111] or 0.49 a along [100] is still in its own cell, and one moved by
0.44 a along [111] or 0.51 a along [100] is in its neighbour's.

Sites are named (i, j, k) with i the doubled x index: the site sits at
(i a/2, (j + (i%2)/2) a, (k + (i%2)/2) a).  Atoms are named by the site they
start on ("s:i,j,k") or, for interstitials present before the step, "C1",
"C2", ... in list order.
"""
import numpy as np

import oracle_py as O

A = 2.855
CELLS = (4, 4, 4)
R3 = 1.0 / np.sqrt(3.0)


def site_pos(i, j, k):
    return np.array([i * 0.5 * A, (j + (i % 2) * 0.5) * A, (k + (i % 2) * 0.5) * A])


class Case:
    """moves: {site: offset in units of a}; vacant: {site: stale offset or
    None} (sites emptied before the step, their position left at the site
    plus the offset); inter: [(name, site, offset)] interstitials present
    before the step, in list order.  expect_inter: names in list order after
    decide; expect_vacant: the vacant sites after decide; expect_site:
    {site: name} for the sites that end up holding another atom than the
    one they started with.  shift: {name: (sx, sy, sz)} periodic shift in
    box lengths that exchangeInter then applies (atoms leaving the box go
    to the end of the list in the order inter_particle_packer.cpp appends
    them; the cases here move at most one atom)."""

    def __init__(self, name, moves=None, vacant=None, inter=(), expect_inter=(), expect_vacant=(),
                 expect_site=None, shift=None, why=""):
        self.name = name
        self.moves = dict(moves or {})
        self.vacant = dict(vacant or {})
        self.inter = list(inter)
        self.expect_inter = list(expect_inter)
        self.expect_vacant = set(expect_vacant)
        self.expect_site = dict(expect_site or {})
        self.shift = dict(shift or {})
        self.why = why


def s(i, j, k):
    return "s:%d,%d,%d" % (i, j, k)


CASES = [
    Case("below 0.2a stays", moves={(2, 1, 1): (0.19, 0, 0)},
         why="0.19a < 0.2a: never leaves its site"),
    Case("above 0.2a inside its own cell refills at once", moves={(2, 1, 1): (0.21, 0, 0)},
         why="taken off (0.21a > 0.2a), its own site is the vacant site of its cell: put back the same step"),
    Case("a waiting interstitial takes a site left above 0.2a",
         moves={(2, 1, 1): (0.21, 0, 0)}, inter=[("C1", (2, 1, 1), (-0.25, 0.03, 0.02))],
         expect_inter=[s(2, 1, 1)], expect_site={(2, 1, 1): "C1"},
         why="C1 comes first in the list and its cell's site is now vacant; the moved atom then finds it taken"),
    Case("a waiting interstitial finds a site held below 0.2a taken",
         moves={(2, 1, 1): (0.19, 0, 0)}, inter=[("C1", (2, 1, 1), (-0.25, 0.03, 0.02))],
         expect_inter=["C1"],
         why="0.19a: the site is never vacated"),
    Case("exactly 0.2a is not more than 0.2a",
         moves={(0, 0, 0): (0.2, 0, 0)}, inter=[("C1", (0, 0, 0), (-0.25, 0.05, 0.03))],
         expect_inter=["C1"],
         why="the test is strict (dist > (0.2a)^2); the atom at x = 0.2a exactly keeps its site"),
    Case("0.42a along [111] is still its own cell", moves={(2, 1, 1): (0.42 * R3, 0.42 * R3, 0.42 * R3)},
         why="the hexagonal face is 0.433a away: taken off and put back"),
    Case("0.44a along [111] crosses into the occupied body centre's cell",
         moves={(2, 1, 1): (0.44 * R3, 0.44 * R3, 0.44 * R3)},
         expect_inter=[s(2, 1, 1)], expect_vacant=[(2, 1, 1)],
         why="nearest site (3,1,1) is occupied: the atom stays interstitial and its site stays vacant"),
    Case("0.49a along [100] is still its own cell", moves={(2, 1, 1): (0.49, 0, 0)},
         why="the square face is 0.5a away"),
    Case("0.51a along [100] crosses into the occupied corner's cell", moves={(2, 1, 1): (0.51, 0, 0)},
         expect_inter=[s(2, 1, 1)], expect_vacant=[(2, 1, 1)],
         why="nearest site (4,1,1) is occupied"),
    Case("an atom moving into a vacancy's cell fills it",
         moves={(2, 1, 1): (0.6, 0, 0)}, vacant={(4, 1, 1): None},
         expect_vacant=[(2, 1, 1)], expect_site={(4, 1, 1): s(2, 1, 1)},
         why="0.4a from the vacant (4,1,1), 0.6a from its own site"),
    Case("an interstitial beside a vacancy fills it",
         vacant={(3, 2, 2): None}, inter=[("C1", (3, 2, 2), (0.1, -0.05, 0.04))],
         expect_site={(3, 2, 2): "C1"}),
    Case("two claimants: the first in the list wins",
         vacant={(3, 2, 2): None},
         inter=[("C1", (3, 2, 2), (0.1, 0, 0)), ("C2", (3, 2, 2), (0, -0.1, 0.05))],
         expect_inter=["C2"], expect_site={(3, 2, 2): "C1"}),
    Case("two claimants in the other order",
         vacant={(3, 2, 2): None},
         inter=[("C2", (3, 2, 2), (0, -0.1, 0.05)), ("C1", (3, 2, 2), (0.1, 0, 0))],
         expect_inter=["C1"], expect_site={(3, 2, 2): "C2"}),
    Case("a site left through a face is not refilled this step",
         moves={(0, 1, 1): (-0.6, 0, 0)}, inter=[("C1", (0, 1, 1), (0.25, 0.05, 0.0))],
         expect_inter=["C1", s(0, 1, 1)], expect_vacant=[(0, 1, 1)], shift={s(0, 1, 1): (1, 0, 0)},
         why="the vacancy keeps the leaver's position, 0.6a below x = 0 and outside the box, so the box test "
             "on it fails and C1 may not refill; the leaver's own nearest site is a ghost (no sub-box index)"),
    Case("an atom whose nearest site is a ghost stays interstitial",
         moves={(0, 1, 1): (-0.6, 0, 0)}, vacant={(6, 1, 1): (1.1, 0, 0)},
         expect_inter=[s(0, 1, 1)], expect_vacant=[(0, 1, 1), (6, 1, 1)], shift={s(0, 1, 1): (1, 0, 0)},
         why="its nearest site is the ghost image of the vacant (6,1,1), whose previous occupant left through "
             "the +x face, so the ghost's position lies inside the box; ghosts are never refilled, the atom "
             "migrates by +L"),
    Case("atoms taken off in one step join the list in z, y, x site order",
         moves={(5, 2, 2): (0.45 * R3, 0.45 * R3, 0.45 * R3), (2, 0, 1): (0.45 * R3, 0.45 * R3, 0.45 * R3),
                (6, 1, 0): (0.45 * R3, 0.45 * R3, 0.45 * R3), (1, 2, 1): (0.45 * R3, 0.45 * R3, 0.45 * R3)},
         expect_inter=[s(6, 1, 0), s(2, 0, 1), s(1, 2, 1), s(5, 2, 2)],
         expect_vacant=[(5, 2, 2), (2, 0, 1), (6, 1, 0), (1, 2, 1)],
         why="each crosses a hexagonal face into an occupied cell; the list order is k, then j, then i"),
]


def build(sim, case, seed=7):
    """Set up `case` on the oracle `sim` (a 4^3 box): perfect sites, shuffled
    ids, mixed species, zero velocities and forces.  Returns the inter list,
    and the id, position and species of each named atom."""
    sb = sim.sub_box()
    rng = np.random.default_rng(seed)
    n = sb.size
    sb["id"] = (rng.permutation(n) + 1).reshape(sb.shape)
    t = np.zeros(n, dtype=np.int32)
    t[rng.choice(n, size=n // 8, replace=False)] = 1
    t[rng.choice(n, size=n // 16, replace=False)] = 2
    sb["type"] = t.reshape(sb.shape)
    for k in range(sb.shape[0]):
        for j in range(sb.shape[1]):
            for i in range(sb.shape[2]):
                sb[k, j, i]["x"] = site_pos(i, j, k)
    sb["v"] = 0.0
    sb["f"] = 0.0
    ids, pos, typ = {}, {}, {}
    for k in range(sb.shape[0]):
        for j in range(sb.shape[1]):
            for i in range(sb.shape[2]):
                ids[s(i, j, k)] = int(sb[k, j, i]["id"])
                pos[s(i, j, k)] = site_pos(i, j, k)
                typ[s(i, j, k)] = int(sb[k, j, i]["type"])
    for (i, j, k), off in case.moves.items():
        x = site_pos(i, j, k) + np.asarray(off) * A
        sb[k, j, i]["x"] = x
        pos[s(i, j, k)] = x
    for (i, j, k), off in case.vacant.items():
        sb[k, j, i]["type"] = -1
        sb[k, j, i]["v"] = 0.0
        if off is not None:
            sb[k, j, i]["x"] = site_pos(i, j, k) + np.asarray(off) * A
    inter = np.zeros(len(case.inter), dtype=O.ATOM_DTYPE)
    next_id = n + 1
    for q, (name, site, off) in enumerate(case.inter):
        inter[q]["id"] = next_id + int(name[1:])
        inter[q]["type"] = int(name[1:]) % 3
        inter[q]["x"] = site_pos(*site) + np.asarray(off) * A
        ids[name] = int(inter[q]["id"])
        pos[name] = inter[q]["x"].copy()
        typ[name] = int(inter[q]["type"])
    return inter, {"id": ids, "x": pos, "type": typ}


def check(sub, inter, case, named, shifted=False):
    """Compare a sub-box view and an inter list with the case's hand-written
    outcome; returns a list of mismatch messages (empty when it matches)."""
    ids, pos, typ = named["id"], named["x"], named["type"]
    bad = []
    got = [int(x) for x in inter["id"]]
    want = [ids[nm] for nm in case.expect_inter]
    if got != want:
        name_of = {v: k for k, v in ids.items()}
        bad.append("inter list %s, expected %s" % ([name_of.get(g, g) for g in got], case.expect_inter))
    vac = {(i, j, k) for k in range(sub.shape[0]) for j in range(sub.shape[1]) for i in range(sub.shape[2])
           if sub[k, j, i]["type"] < 0}
    if vac != case.expect_vacant:
        bad.append("vacant sites %s, expected %s" % (sorted(vac), sorted(case.expect_vacant)))
    L = np.array(CELLS) * A
    for k in range(sub.shape[0]):
        for j in range(sub.shape[1]):
            for i in range(sub.shape[2]):
                if (i, j, k) in case.expect_vacant:
                    continue
                nm = case.expect_site.get((i, j, k), s(i, j, k))
                if int(sub[k, j, i]["id"]) != ids[nm] or int(sub[k, j, i]["type"]) != typ[nm]:
                    bad.append("site %s holds id %d type %d, expected %s" %
                               ((i, j, k), int(sub[k, j, i]["id"]), int(sub[k, j, i]["type"]), nm))
                elif not np.array_equal(sub[k, j, i]["x"], pos[nm]):
                    bad.append("site %s: position %s, expected %s" % ((i, j, k), sub[k, j, i]["x"], pos[nm]))
    if not bad:
        for q, nm in enumerate(case.expect_inter):
            want_x = pos[nm] + (np.asarray(case.shift.get(nm, (0, 0, 0))) * L if shifted else 0.0)
            if not np.array_equal(inter[q]["x"], want_x) or int(inter[q]["type"]) != typ[nm]:
                bad.append("inter %s at %s type %d, expected %s" % (nm, inter[q]["x"], int(inter[q]["type"]), want_x))
    return bad
