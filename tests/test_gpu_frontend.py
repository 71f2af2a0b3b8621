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


CONFIG = """
simulation:
  phasespace: [6, 6, 6]
  cutoff_radius_factor: {f}
  lattice_const: {a}
  def_timesteps_length: 0.0005
potential:
  type: "setfl"
  file_path: "{pot}"
creation:
  create_phase: true
  create_seed: 4242
  create_t_set: 600
  alloy:
    create_seed: 7
    ratio:
      Fe: 985
      Cu: 10
      Ni: 5
output:
  dump:
    atoms_dump_mode: "copy"
    atoms_dump_file_path: "{out}/a.{{}}.out"
    origin_dump_path: "{out}/a.origin.out"
    atoms_dump_interval: 5
    by_frame: true
  thermo:
    interval: 3
  logs:
    logs_mode: "file"
    logs_filename: "{out}/md.log"
stages:
  - name: rescale
    step_length: 0.001
    steps: 4
    rescale:
      t: 600
      every_steps: 2
  - name: collision
    step_length: 0.0002
    steps: 8
    set_v:
      collision_step: 3
      lat: [2, 2, 2, 0]
      energy: 300
      direction: [1.0, 3.0, 5.0]
  - name: run
    steps: 6
"""


def test_stage_machine_matches_step_by_step(setfl_fecuni, tmp_path):
    """misa_md_simulate (MDSimulation over misa_step chunks) against the same
    stages replayed one step at a time with the reference's beforeStep /
    postStep rules: same final state, same dumps."""
    cfg = M.parse_config(text=CONFIG.format(f=F, a=A, pot=setfl_fecuni, out=tmp_path))
    rep = M.md_simulate(cfg, 0, M.PREC_FP64)
    assert rep.steps == 18 and rep.stages_run == 3 and rep.collisions == 1 and rep.rescales == 2
    assert rep.dump_frames == 3 and rep.thermo_lines == 6
    assert abs(rep.phy_time - (4 * 0.001 + 8 * 0.0002 + 6 * 0.0005)) < 1e-15
    log = (tmp_path / "md.log").read_text()
    assert log.count("kinetic energy =") == 6 and "collision at step 6" in log

    dev = M.EamDevice((6, 6, 6), A, F)
    dev.make_offsets()
    dev.set_potential(setfl_fecuni)
    dev.set_precision(M.PREC_FP64)
    dev.build_world(4242, 600.0, (985, 10, 5), 7)
    dev.set_timestep(0.001)
    dev.compute_forces()
    stages = cfg.stage_list()
    nxt, cur, stage = 0, 0, None
    for s in range(18):
        if stage is None or cur >= stage.steps:
            stage = stages[nxt]
            nxt += 1
            cur = 0
            dev.set_timestep(stage.step_length or 0.0005)
        if stage.collision_set and cur + 1 == stage.collision_step:
            d = M.AtomDumpFile(tmp_path / "b.origin.out")
            d.frame_state(dev, s)
            d.close()
            dev.collision_step(tuple(stage.collision_lat), tuple(stage.direction), stage.pka_energy)
        if stage.rescale_set and cur % stage.rescale_every == 0:
            dev.rescale(stage.rescale_t)
        dev.step(1)
        cur += 1
        if (s + 1) % 5 == 0:
            d = M.AtomDumpFile(tmp_path / ("b.%d.out" % (s + 1)))
            d.frame_state(dev, s + 1)
            d.close()
    th = dev.thermo()
    for k in ("kinetic", "e_potential", "temperature", "n_inter", "n_vacancy"):
        assert rep.thermo.__getattribute__(k) == th[k], k
    for name in ("origin", "5", "10", "15"):
        assert (tmp_path / ("a.%s.out" % name)).read_bytes() == (tmp_path / ("b.%s.out" % name)).read_bytes(), name
    dev.close()
