"""The force passes (mixed and FP64) in persistent blocks with the majority
species' spline knots in LDS (needs an MI355X).

An LDS pass visits the same pairs, in the same order, with the same values
and rounding as the pass that gathers every segment from global memory, so a
box stepped with it must end bit for bit where it ends without it: in every
precision mode, with either count of gathers in flight, through a recoil
(unlike pairs and close pairs take the global tables), and on a fine grid
whose first knots lie below the LDS range.
"""
import numpy as np
import pytest

import misa_eam as M
import oracle_py as O
import setfl_gen

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855
FIELDS = ("id", "type", "x", "v", "f", "rho", "df")


def _pair(setfl, cells, seed, tset, precision):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, tset, (900, 70, 30), seed)
    atoms = sim.atoms().copy()
    out = []
    for mode in (M.LDS_OFF, M.LDS_ON):
        dev = M.EamDevice(cells)
        dev.make_offsets()
        dev.set_potential(setfl)
        dev.set_precision(precision)
        dev.set_lds_tables(mode)
        dev.upload(atoms)
        dev.set_timestep(0.0005)
        dev.compute_forces()
        out.append(dev)
    return out


def _same(a, b, what):
    sa, ia = a.download()
    sb, ib = b.download()
    va, vb = M.sub_box_view(sa, a), M.sub_box_view(sb, b)
    for k in FIELDS:
        np.testing.assert_array_equal(va[k], vb[k], err_msg="%s: %s" % (what, k))
        np.testing.assert_array_equal(ia[k], ib[k], err_msg="%s: inter %s" % (what, k))
    return len(ia)


@pytest.mark.parametrize("precision", [M.PREC_MIXED_KAHAN, M.PREC_MIXED, M.PREC_MIXED_F32SUM, M.PREC_FP64])
@pytest.mark.parametrize("gathers", [4, 2])
def test_lds_force_pass_gives_one_trajectory(setfl_fecuni, precision, gathers):
    cells = (8, 8, 12)
    print("lds: precision %d gathers %d" % (precision, gathers), flush=True)
    off, on = _pair(setfl_fecuni, cells, 41, 900.0, precision)
    for dev in (off, on):
        dev.set_launch_shape(4, 4, gathers)
    _same(off, on, "first forces")
    # the energy pass also runs in persistent blocks with LDS tables: the
    # same terms, summed in another order
    eo, en = off.thermo(), on.thermo()
    assert en["e_potential"] == pytest.approx(eo["e_potential"], rel=1e-13)
    assert en["virial"] == pytest.approx(eo["virial"], rel=1e-12)
    for dev in (off, on):
        dev.step(10)
        dev.collision_step((4, 4, 6, 0), (1.0, 3.0, 5.0), 600.0)
        dev.step(25)
    ni = _same(off, on, "after a recoil")
    print("lds: inter %d" % ni, flush=True)
    assert ni > 0
    for dev in (off, on):  # the adaptive step: force pass and first half step apart
        dev.set_adaptive_dt(0.001, 0.01)
        dev.step(10)
    _same(off, on, "after adaptive steps")
    off.close()
    on.close()


@pytest.mark.parametrize("precision", [M.PREC_MIXED_KAHAN, M.PREC_FP64])
def test_lds_fine_grid_falls_back_below_the_table(tmp_path, precision):
    """8000 knots: the LDS holds knots from about 2.1 A (mixed) or 3.6 A
    (FP64) on, so the closer pairs of a 1200 K box take the global tables."""
    path = setfl_gen.write_setfl(str(tmp_path / "fine.eam.alloy"), ("Fe", "Cu", "Ni"), nr=8000, nrho=3000,
                                 rho_max=60.0)
    off, on = _pair(path, (7, 6, 8), 43, 1200.0, precision)
    _same(off, on, "first forces")
    for dev in (off, on):
        dev.step(30)
    _same(off, on, "after 30 steps")
    off.close()
    on.close()


def test_lds_mode_is_checked(setfl_fe):
    dev = M.EamDevice((4, 4, 4))
    with pytest.raises(M.MisaError):
        dev.set_lds_tables(3)
    dev.set_lds_tables(M.LDS_AUTO)
    dev.close()
