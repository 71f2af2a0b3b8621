"""The frontend's YAML configuration (ConfigParser into ConfigValues): the
reference's own example configuration (example/config.yaml, kept as the
fixture tests/golden/example_config.yaml) parses to the values its comments
state, missing sections give the reference's messages, and the defaults are
those of def_config_values.h and config_values.cpp."""
import os

import pytest

import misa_eam as M

HERE = os.path.dirname(os.path.abspath(__file__))
EXAMPLE = os.path.join(HERE, "golden", "example_config.yaml")

MINIMAL = """
simulation:
  phasespace: [6, 7, 8]
  cutoff_radius_factor: 1.9614711
  lattice_const: 2.855
potential:
  type: "setfl"
  file_path: "pot.eam.alloy"
creation:
  create_t_set: 300
  alloy:
    ratio:
      Fe: 985
      Cu: 10
      Ni: 5
output:
  dump:
    atoms_dump_mode: "copy"
  thermo:
    interval: 0
  logs:
    logs_mode: "console"
stages:
  - steps: 5
"""


def test_reference_example_configuration():
    c = M.parse_config(EXAMPLE)
    assert list(c.phase_space) == [50, 50, 50]
    assert c.cutoff_radius_factor == 1.96125 and c.lattice_const == 2.85532 and c.time_step_length == 0.001
    assert c.create_phase == 1 and c.create_seed == 466953 and c.create_t_set == 600.0
    assert c.alloy_create_seed == 1024 and list(c.alloy_ratio) == [97, 2, 1]
    assert c.potential_file_type == b"setfl" and c.potential_filename == b"FeCuNi.eam.alloy"
    assert c.atoms_dump_mode == M.DUMP_COPY and c.out_by_frame == 1 and c.atoms_dump_interval == 10
    assert c.atoms_dump_file_path == b"misa_mdl.{}.out" and c.origin_dump_path == b"misa_mdl.origin.out"
    assert c.thermo_interval == 0 and c.logs_mode == 0
    st = c.stage_list()
    assert len(st) == 3 and c.time_steps == 4 + 8 + 6
    assert (st[0].steps, st[0].step_length, st[0].rescale_set, st[0].rescale_t, st[0].rescale_every) == \
        (4, 0.001, 1, 600.0, 2)
    assert st[0].collision_set == 0
    assert (st[1].steps, st[1].step_length, st[1].collision_set, st[1].collision_step) == (8, 0.0001, 1, 2)
    assert list(st[1].collision_lat) == [2, 2, 2, 0] and st[1].pka_energy == 6.8
    assert list(st[1].direction) == [1.0, 1.0, 1.0] and st[1].rescale_set == 0
    assert (st[2].steps, st[2].step_length, st[2].collision_set, st[2].rescale_set) == (6, 0.001, 0, 0)


def test_defaults():
    c = M.parse_config(text=MINIMAL)
    assert list(c.phase_space) == [6, 7, 8]
    assert c.time_step_length == 0.001                     # default_time_length
    assert c.create_phase == 1 and c.create_seed == 1024   # default_random_seek
    assert c.alloy_create_seed == 1024 and list(c.alloy_ratio) == [985, 10, 5]
    assert c.atoms_dump_interval == 2 ** 64 - 1            # ULONG_MAX: no dump
    assert c.out_by_frame == 0 and c.atoms_dump_file_path == b"misa_mdl.out"
    assert c.origin_dump_path == b""                       # empty: no recoil, no origin dump
    st = c.stage_list()
    assert len(st) == 1 and st[0].steps == 5 and st[0].step_length == 0.0  # 0: the default at run time
    assert st[0].collision_set == 0 and st[0].rescale_set == 0


@pytest.mark.parametrize("drop, msg", [
    ("simulation", '"simulation" in config is not specified.'),
    ("creation", '"creation" in config is not specified.'),
    ("potential", '"potential" in config is not specified.'),
    ("output", "output section in config file is not specified."),
])
def test_missing_sections(drop, msg):
    lines, skip = [], False
    for ln in MINIMAL.splitlines():
        if ln and not ln.startswith(" "):
            skip = ln.startswith(drop + ":")
        if not skip:
            lines.append(ln)
    with pytest.raises(M.MisaError) as e:
        M.parse_config(text="\n".join(lines))
    assert msg in str(e.value)


@pytest.mark.parametrize("old, new, msg", [
    ("phasespace: [6, 7, 8]", "phasespace: [6, 7]", '"simulation.phasespace" must be a array with length 3.'),
    ("  cutoff_radius_factor: 1.9614711\n", "", "cutoff_radius_factor must be set in config."),
    ("  create_t_set: 300\n", "", "creation t_set must be specified."),
    ('    atoms_dump_mode: "copy"', '    atoms_dump_mode: "copy"\n    by_frame: true',
     "error format of dump file path"),
    ("  - steps: 5", "  - steps: 5\n    set_v:\n      lat: [1, 2, 3]\n      direction: [1, 0, 0]",
     'array length of value "collision.lat" must be 4.'),
    ("  - steps: 5", "  - steps: 5\n    set_v:\n      lat: [1, 2, 3, 0]\n      direction: [1, 0]",
     'array length of value "simulation.collision.collision_v" must be 3.'),
    ("    ratio:\n      Fe: 985\n      Cu: 10\n      Ni: 5", "    ratio: [1, 2, 3]", "alloy ratio must be a map."),
])
def test_reference_messages(old, new, msg):
    assert old in MINIMAL
    with pytest.raises(M.MisaError) as e:
        M.parse_config(text=MINIMAL.replace(old, new))
    assert msg in str(e.value)


def test_yaml_subset():
    """Comments, quoting, booleans, a sequence at its key's indentation."""
    text = MINIMAL.replace('  file_path: "pot.eam.alloy"', "  file_path: 'dir #1/pot''s.eam.alloy'  # quoted")
    text = text.replace("  create_t_set: 300", "  create_phase: yes\n  create_t_set: 300.5e0")
    text = text.replace("stages:\n  - steps: 5", "stages:\n- steps: 5\n  step_length: 0.0002\n- steps: 7")
    c = M.parse_config(text=text)
    assert c.potential_filename == b"dir #1/pot's.eam.alloy"
    assert c.create_phase == 1 and c.create_t_set == 300.5
    st = c.stage_list()
    assert [s.steps for s in st] == [5, 7] and st[0].step_length == 0.0002 and c.time_steps == 12
    with pytest.raises(M.MisaError) as e:
        M.parse_config(text=MINIMAL.replace("  lattice_const: 2.855", "\tlattice_const: 2.855"))
    assert "tab" in str(e.value)
    with pytest.raises(M.MisaError):
        M.parse_config(text=MINIMAL.replace("  lattice_const: 2.855", "  lattice_const: &a 2.855"))
    with pytest.raises(M.MisaError) as e:
        M.parse_config(path=os.path.join(HERE, "no_such_config.yaml"))
    assert "can not access the configure file" in str(e.value)


def test_read_mode_needs_a_file():
    text = MINIMAL.replace("  create_t_set: 300", "  create_phase: false")
    with pytest.raises(M.MisaError) as e:
        M.parse_config(text=text)
    assert "read phase file must be specified." in str(e.value)
    c = M.parse_config(text=text.replace("  create_phase: false", "  create_phase: false\n  read_phase_filename: x"))
    assert c.create_phase == 0 and c.read_phase_filename == b"x"
