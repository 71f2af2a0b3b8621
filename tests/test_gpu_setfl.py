"""The device path on production-style setfl files (needs an MI355X): a
non-default grid (Nr 8000, Nrho 3000, E-format, four values a line) against
the oracle at the FP64 bars, and a file listing the elements in another
order giving bit-identical forces (species follow atomic numbers)."""
import numpy as np
import pytest

import misa_eam as M
import oracle_py as O
import setfl_gen

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855
ELS = ("Fe", "Cu", "Ni")


def _forces(path, sim):
    dev = M.EamDevice(sim.cells)
    dev.make_offsets()
    dev.set_potential(path)
    dev.upload(sim.atoms().copy())
    dev.compute_forces()
    aos, _ = dev.download()
    th = dev.thermo()
    out = M.sub_box_view(aos, dev).copy(), th
    dev.close()
    return out


def test_own_grid_against_oracle(tmp_path):
    path = setfl_gen.write_setfl(str(tmp_path / "g.eam.alloy"), ELS, nr=8000, nrho=3000, rho_max=60.0,
                                 per_line=4, fmt="%24.16E")
    cells = (7, 6, 8)
    sim = O.Sim(cells, A, F, path)
    sim.build_world(77, 900.0, (900, 70, 30), 77)
    sb = sim.sub_box()
    sb["x"] = sb["x"] + np.random.default_rng(77).uniform(-0.15, 0.15, sb["x"].shape)
    got, th = _forces(path, sim)
    sim.prepare()
    want = sim.sub_box()
    np.testing.assert_allclose(got["rho"], want["rho"], rtol=1e-12)
    f, w = got["f"].reshape(-1, 3), want["f"].reshape(-1, 3)
    rms = np.sqrt(np.mean(np.sum(w * w, axis=1)))
    err = np.linalg.norm(f - w, axis=1) / np.maximum(np.linalg.norm(w, axis=1), rms)
    print("own grid: force err max %.3e" % err.max(), flush=True)
    assert err.max() < 1e-12
    assert abs(th["e_potential"] - sim.energy()["e_potential"]) / (2 * np.prod(cells)) < 1e-13


def test_element_order_in_file(tmp_path):
    a = setfl_gen.write_setfl(str(tmp_path / "a.eam.alloy"), ELS, nr=3000, nrho=3000)
    b = setfl_gen.write_setfl(str(tmp_path / "b.eam.alloy"), ("Ni", "Fe", "Cu"), nr=3000, nrho=3000)
    sim = O.Sim((6, 6, 6), A, F, a)
    sim.build_world(5, 600.0, (900, 70, 30), 5)
    ga, _ = _forces(a, sim)
    gb, _ = _forces(b, sim)
    for k in ("rho", "df", "f"):
        np.testing.assert_array_equal(ga[k], gb[k])
