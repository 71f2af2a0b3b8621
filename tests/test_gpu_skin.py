"""The neighbour skin of the density pass (needs an MI355X): between
full scans the pass visits only the offsets that were within cutoff + skin,
which must leave every sum, and so every trajectory, bit-identical to a pass
that scans all offsets every step: at equilibrium with a fixed step, and
through a cascade whose refills force full scans, with an adaptive step."""
import numpy as np
import pytest

import misa_eam as M
import oracle_py as O

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855
FIELDS = ("id", "type", "x", "v", "f", "rho", "df")


def _pair(setfl, cells, seed, tset, ratio, precision):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, tset, ratio, seed)
    atoms = sim.atoms().copy()
    out = []
    for skin in (0.2, 0.0):
        dev = M.EamDevice(cells)
        dev.make_offsets()
        dev.set_potential(setfl)
        dev.set_precision(precision)
        dev.set_neighbour_skin(skin)
        dev.upload(atoms)
        out.append(dev)
    return out


def _same(a, b, what):
    sa, ia = a.download()
    sb, ib = b.download()
    va, vb = M.sub_box_view(sa, a), M.sub_box_view(sb, b)
    for k in FIELDS:
        np.testing.assert_array_equal(va[k], vb[k], err_msg="%s: %s" % (what, k))
    assert len(ia) == len(ib)
    for k in ("id", "x", "v", "f"):
        np.testing.assert_array_equal(ia[k], ib[k], err_msg="%s: inter %s" % (what, k))


@pytest.mark.parametrize("precision", [M.PREC_MIXED_KAHAN, M.PREC_FP64])
def test_skin_at_equilibrium(setfl_fecuni, precision):
    a, b = _pair(setfl_fecuni, (8, 8, 12), 21, 900.0, (900, 70, 30), precision)
    for d in (a, b):
        d.set_timestep(0.001)
        d.compute_forces()
    for k in range(4):
        a.step(15)
        b.step(15)
        _same(a, b, "after %d steps" % (15 * (k + 1)))
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [M.PREC_MIXED_KAHAN, M.PREC_FP64])
def test_skin_through_a_cascade(setfl_fecuni, precision):
    a, b = _pair(setfl_fecuni, (10, 10, 10), 7, 600.0, (985, 10, 5), precision)
    for d in (a, b):
        d.set_timestep(0.0005)
        d.compute_forces()
        d.collision_step((5, 5, 5, 0), (1.0, 3.0, 5.0), 1000.0)
        d.set_adaptive_dt(0.0005, 0.01)
    seen = 0
    for k in range(6):
        a.step(25)
        b.step(25)
        seen = max(seen, a.thermo()["n_inter"])
        _same(a, b, "cascade +%d steps" % (25 * (k + 1)))
    assert seen > 0
    a.close()
    b.close()
