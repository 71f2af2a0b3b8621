"""Device error flags (needs an MI355X).

The reference's hooks have no error channel (arch_imp.h:17-31); its
development build aborts when the summed force leaves 1e-5
(md_simulation.cpp:150-164).  Here every call that runs kernels checks the
device's flags when it ends.  The states below make values non-finite
without any index leaving its array: a potential whose Fe-Fe pair table
holds 1e308 and 0 in turn around the nearest-neighbour distance (the
interpolation's 3 (y1 - y0) overflows there, so every first-shell pair force
is NaN), and a potential whose F(rho) table ends below the lattice's electron
density.
"""
import numpy as np
import pytest

import misa_eam as M
import setfl_gen

pytestmark = pytest.mark.gpu

CELLS = (6, 6, 6)


def _broken_pair_setfl(path):
    """Fe-Cu-Ni tables with the Fe-Fe r*phi knots of 2.3-2.65 A set to 1e308
    and 0 in turn: every segment there has coefficients of +-inf."""
    setfl_gen.write_setfl(path, ("Fe", "Cu", "Ni"))
    dr, _ = setfl_gen.grids()
    lines = open(path).read().split("\n")
    # header (5 lines), then per element: 1 line + F (1000 lines) + rho (1000 lines)
    first_pair = 5 + 3 * (1 + 2 * 1000)
    vals = " ".join(lines[first_pair:first_pair + 1000]).split()
    assert len(vals) == 5000
    for k in range(int(2.3 / dr), int(2.65 / dr)):
        vals[k] = "1e308" if k % 2 == 0 else "0"
    lines[first_pair:first_pair + 1000] = [" ".join(vals[i:i + 5]) for i in range(0, 5000, 5)]
    open(path, "w").write("\n".join(lines))
    return path


def _device(pot, precision):
    dev = M.EamDevice(CELLS)
    dev.make_offsets()
    dev.set_potential(pot)
    dev.set_precision(precision)
    dev.build_world(466953, 600.0, (985, 10, 5), 1)
    return dev


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED_KAHAN], ids=["fp64", "mixed"])
def test_non_finite_force_is_reported(tmp_path, setfl_fecuni, precision):
    bad = _broken_pair_setfl(str(tmp_path / "broken.eam.alloy"))
    print("non-finite force, precision", precision, flush=True)
    dev = _device(bad, precision)
    with pytest.raises(M.MisaError) as e:
        dev.compute_forces()
    assert e.value.code == M.E_DEVICE
    assert dev.device_flags() & M.FLAG_NON_FINITE
    # the same through a step: good forces first, then the broken table
    dev.close()
    dev = _device(setfl_fecuni, precision)
    dev.compute_forces()
    dev.step(2)
    assert dev.device_flags() == 0
    dev.set_potential(bad)
    with pytest.raises(M.MisaError) as e:
        dev.step(1)
    assert e.value.code == M.E_DEVICE
    assert dev.device_flags() & M.FLAG_NON_FINITE
    dev.close()


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED_KAHAN], ids=["fp64", "mixed"])
def test_density_beyond_the_table_is_flagged(tmp_path, precision):
    """A perfect-lattice Fe site sees rho well above 2.0; with F(rho)
    tabulated up to 2.0 every site lies beyond the table.  The flag is always
    recorded; only the range check turns it into an error."""
    pot = setfl_gen.write_setfl(str(tmp_path / "short_rho.eam.alloy"), ("Fe", "Cu", "Ni"), rho_max=2.0)
    dev = _device(pot, precision)
    dev.compute_forces()  # range check off: recorded, no error
    assert dev.device_flags() == M.FLAG_RHO_RANGE
    dev.set_range_check(True)
    with pytest.raises(M.MisaError) as e:
        dev.step(1)
    assert e.value.code == M.E_DEVICE
    dev.close()
    # a full table raises nothing
    ok = setfl_gen.write_setfl(str(tmp_path / "ok.eam.alloy"), ("Fe", "Cu", "Ni"))
    dev = _device(ok, precision)
    dev.set_range_check(True)
    dev.compute_forces()
    dev.step(3)
    assert dev.device_flags() == 0
    dev.close()


def test_range_check_argument_is_checked():
    dev = M.EamDevice((4, 4, 4))
    assert dev.lib.misa_set_range_check(dev.h, 2) == M.E_ARG
    dev.close()
