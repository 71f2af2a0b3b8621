"""Frontend pieces on the device state (needs an MI355X): the atom dump of a
resident state equals the dump of the same state downloaded to the host."""
import numpy as np
import pytest

import misa_eam as M
import oracle_py as O

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855


def test_dump_of_resident_state(setfl_fecuni, tmp_path):
    cells = (6, 6, 8)
    sim = O.Sim(cells, A, F, setfl_fecuni)
    sim.build_world(11, 600.0, (985, 10, 5), 11)
    dev = M.EamDevice(cells)
    dev.make_offsets()
    dev.set_potential(setfl_fecuni)
    dev.upload(sim.atoms().copy())
    dev.set_timestep(0.0005)
    dev.compute_forces()
    dev.collision_step((3, 3, 4, 0), (1.0, 3.0, 5.0), 500.0)
    dev.step(40)
    assert dev.thermo()["n_inter"] > 0
    a = M.AtomDumpFile(tmp_path / "dev.out")
    a.frame_state(dev, 40)
    na = a.close()
    atoms, inter = dev.download()
    b = M.AtomDumpFile(tmp_path / "host.out")
    b.frame(atoms, inter, cells, 40)
    nb = b.close()
    assert na == nb == len(inter) + int((M.sub_box_view(atoms, dev)["type"] >= 0).sum())
    assert (tmp_path / "dev.out").read_bytes() == (tmp_path / "host.out").read_bytes()
    M.dump_text_state(dev, tmp_path / "dev.atom")
    M.dump_text(tmp_path / "host.atom", atoms, inter, cells)
    assert (tmp_path / "dev.atom").read_text() == (tmp_path / "host.atom").read_text()
    dev.close()
