"""Work-balanced XCD slabs (needs an MI355X).

The site passes give each of the 8 XCDs a z-slab of the box.  With balancing
on, the first LDS force pass of every step call measures the cycles spent per
z-plane and the slabs are re-cut so that each holds about an eighth of that
work.  Which XCD walks a site changes nothing of what the site computes (the
same pairs in the same order), so a box stepped with any cut must end bit for
bit where the same box ends with the even cut: through a recoil, adaptive
and fixed steps, and with a cut measured on a cascade whose disordered core
sits in one slab.
"""
import numpy as np
import pytest

import misa_eam as M
import oracle_py as O

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855
FIELDS = ("id", "type", "x", "v", "f", "rho", "df")


def _same(a, b, what):
    sa, ia = a.download()
    sb, ib = b.download()
    va, vb = M.sub_box_view(sa, a), M.sub_box_view(sb, b)
    for k in FIELDS:
        np.testing.assert_array_equal(va[k], vb[k], err_msg="%s: %s" % (what, k))
        np.testing.assert_array_equal(ia[k], ib[k], err_msg="%s: inter %s" % (what, k))
    return len(ia)


def _same_energy(a, b, what):
    """The same terms in another order: FP64 sums agree to rounding."""
    ta, tb = a.thermo(), b.thermo()
    for k in ("e_potential", "virial"):
        assert tb[k] == pytest.approx(ta[k], rel=1e-12), "%s: %s" % (what, k)


def _even(n2):
    return [c * n2 // 8 for c in range(9)]


def _pair(setfl, cells, seed, tset):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, tset, (900, 70, 30), seed)
    atoms = sim.atoms().copy()
    devs = []
    for _ in range(2):
        dev = M.EamDevice(cells)
        dev.make_offsets()
        dev.set_potential(setfl)
        dev.set_precision(M.PREC_MIXED_KAHAN)
        dev.set_lds_tables(M.LDS_ON)
        dev.upload(atoms)
        dev.set_timestep(0.0005)
        dev.compute_forces()
        devs.append(dev)
    devs[0].set_xcd_balance(False)
    return devs


def test_any_cut_gives_one_trajectory(setfl_fecuni):
    cells = (16, 16, 48)
    even, cut = _pair(setfl_fecuni, cells, 7, 600.0)
    assert even.xcd_slabs() == _even(48)
    for bad in ([0, 6, 12, 18, 24, 30, 36, 42, 47], [0, 6, 5, 18, 24, 30, 36, 42, 48],
                [0, 13, 14, 18, 24, 30, 36, 42, 48], [1, 6, 12, 18, 24, 30, 36, 42, 48]):
        with pytest.raises(M.MisaError):
            cut.set_xcd_slabs(bad)
    zlo = [0, 2, 4, 4, 12, 24, 30, 40, 48]  # an empty slab and one of twice the even thickness
    cut.set_xcd_slabs(zlo)
    assert cut.xcd_slabs() == zlo
    print("cut: recoil", flush=True)
    for dev in (even, cut):
        dev.step(5)
        dev.collision_step((8, 8, 20, 0), (1.0, 3.0, 5.0), 1500.0)
        dev.step(25)  # the fixed step: the force pass fused with the next half step
    assert _same(even, cut, "after the recoil") > 0
    _same_energy(even, cut, "after the recoil")
    for dev in (even, cut):
        dev.set_adaptive_dt(0.001, 0.01)
        dev.step(20)
    _same(even, cut, "after adaptive steps")
    assert cut.xcd_slabs() == zlo  # an explicit cut turns balancing off
    for dev in (even, cut):
        dev.close()


def test_measured_cut_thins_the_cascade_slab(setfl_fecuni):
    """A 3 keV recoil from plane 43 of a 64^3 box, heading up the z axis:
    after 0.08 ps the slabs over its core are cut thinner than the even
    eighth, the cut covers the box, and the trajectory is the even cut's."""
    cells = (64, 64, 64)
    even, bal = _pair(setfl_fecuni, cells, 11, 300.0)
    print("measured: recoil at plane 43", flush=True)
    for dev in (even, bal):
        dev.collision_step((32, 32, 43, 0), (1.0, 3.0, 5.0), 3000.0)
        dev.set_adaptive_dt(0.001, 0.01)
    while bal.sim_time()[0] < 0.08:
        for dev in (even, bal):
            dev.step(25)
    assert even.xcd_slabs() == _even(64)
    zlo = bal.xcd_slabs()
    print("measured: slabs", zlo, flush=True)
    assert zlo[0] == 0 and zlo[8] == 64 and all(zlo[c + 1] > zlo[c] for c in range(8))
    # the recoil runs up the z axis from plane 43: the thinned slabs lie there
    thin = [c for c in range(8) if zlo[c + 1] - zlo[c] < 8]
    assert thin, "no slab was cut thinner than the even eighth"
    assert all(zlo[c + 1] > 40 for c in thin), "slabs away from the cascade were thinned"
    assert _same(even, bal, "after the cascade") > 0
    # right after the step call that re-cut the slabs, the energy pass must
    # not take the last density pass's wave masks for the new cut's waves
    _same_energy(even, bal, "after the cascade")
    for dev in (even, bal):
        dev.close()
