"""The one-rank mirror of libcomm (BccDomain, fwCommLocalRegion, Packer,
neiSendReceive) and of kiwi::logs, on the host.

Domain fields are checked against the relations the reference's own code
relies on (the sub-box start output_base_interface.h:23-35 and atom.cpp:154
derive, the ghost-extended array the world builder fills); the forwarding
regions against what they must achieve (their landing layers tile the ghost
shell exactly once); the lattice halo against the periodic closed form (every
ghost slot holds its image's type and the image's position plus the box
length it was shifted by)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import misa_eam as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEOMS = [((20, 20, 20), 2.855, 5.6 / 2.855), ((6, 7, 9), 2.85532, 1.96125), ((5, 5, 5), 3.0, 1.1421)]


@pytest.mark.parametrize("cells, a, f", GEOMS)
def test_domain_fields(cells, a, f):
    d = M.host_domain(cells, a, f)
    g = int(np.ceil(f)) + 1
    assert d.cut_lattice == int(np.ceil(f)) and list(d.lattice_size_ghost) == [g] * 3
    assert list(d.phase_space) == list(cells) and d.lattice_const == a and d.cutoff_radius_factor == f
    assert list(d.grid_size) == [1, 1, 1] and list(d.grid_coord) == [0, 0, 0]
    assert list(d.rank_id_neighbours) == [0] * 6
    L = [c * a for c in cells]
    assert list(d.meas_global_length) == L
    assert list(d.meas_global_region) == [0, 0, 0] + L == list(d.meas_sub_box_region)
    assert list(d.sub_box_lattice_size) == list(cells)
    assert list(d.dbx_sub_box_lattice_size) == [2 * cells[0], cells[1], cells[2]]
    ext = [2 * (cells[0] + 2 * g), cells[1] + 2 * g, cells[2] + 2 * g]
    assert list(d.dbx_ghost_extended_lattice_size) == ext
    assert list(d.ghost_extended_lattice_size) == [c + 2 * g for c in cells]
    # the AtomElement array the world builder fills has exactly these dims
    assert len(M.host_build_world(cells, 1, 0.0, (1, 0, 0), 1, a, f)) == ext[0] * ext[1] * ext[2]
    # output_base_interface.h:23-35: the sub-box inside the ghost-extended array
    lo = [d.dbx_sub_box_lattice_region[k] - d.dbx_ghost_ext_lattice_region[k] for k in range(3)]
    assert lo == list(d.dbx_lattice_size_ghost) == [2 * g, g, g]          # atom.cpp:154-156
    hi = [lo[k] + d.dbx_sub_box_lattice_size[k] for k in range(3)]
    assert hi == [d.dbx_sub_box_lattice_region[3 + k] - d.dbx_ghost_ext_lattice_region[k] for k in range(3)]
    for reg, size in ((d.ghost_ext_lattice_region, d.ghost_extended_lattice_size),
                      (d.dbx_ghost_ext_lattice_region, d.dbx_ghost_extended_lattice_size),
                      (d.sub_box_lattice_region, d.sub_box_lattice_size)):
        assert [reg[3 + k] - reg[k] for k in range(3)] == list(size)


@pytest.mark.parametrize("cells, a, f", GEOMS)
def test_forwarding_regions_tile_the_ghost_shell(cells, a, f):
    d = M.host_domain(cells, a, f)
    ext = list(d.dbx_ghost_extended_lattice_size)
    gh, sub = list(d.dbx_lattice_size_ghost), list(d.dbx_sub_box_lattice_size)
    hits = np.zeros(ext[::-1], dtype=np.int32)  # (z, y, X)
    for dim in range(3):
        for dr in range(2):
            r = M.host_fw_region(cells, dim, dr, a, f)
            lo, hi = r[:3], r[3:]
            # the slab is ghost-wide, next to the face, inside the array
            assert hi[dim] - lo[dim] == gh[dim]
            assert lo[dim] == (gh[dim] if dr == 0 else sub[dim])
            for k in range(3):
                assert 0 <= lo[k] < hi[k] <= ext[k]
            # where it lands: moved across the box by the sub-box length
            shift = sub[dim] if dr == 0 else -sub[dim]
            lo[dim] += shift
            hi[dim] += shift
            hits[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]] += 1
    core = np.zeros_like(hits, dtype=bool)
    core[gh[2]:gh[2] + sub[2], gh[1]:gh[1] + sub[1], gh[0]:gh[0] + sub[0]] = True
    assert (hits[core] == 0).all()
    assert (hits[~core] == 1).all()


@pytest.mark.parametrize("cells, a, f", GEOMS[1:])
def test_lattice_halo_closed_form(cells, a, f):
    atoms = M.host_build_world(cells, 9, 300.0, (90, 7, 3), 9, a, f)
    d = M.host_domain(cells, a, f)
    ext = list(d.dbx_ghost_extended_lattice_size)
    gh, sub = list(d.dbx_lattice_size_ghost), list(d.dbx_sub_box_lattice_size)
    box = atoms.reshape(ext[2], ext[1], ext[0])
    rng = np.random.default_rng(1)
    core = (slice(gh[2], gh[2] + sub[2]), slice(gh[1], gh[1] + sub[1]), slice(gh[0], gh[0] + sub[0]))
    box[core]["x"] += rng.uniform(-0.2, 0.2, box[core]["x"].shape)
    box[core]["type"][rng.random(box[core]["type"].shape) < 0.02] = -1
    mask = np.ones(box.shape, dtype=bool)
    mask[core] = False
    box["type"][mask] = 77
    box["x"][mask] = np.nan
    M.host_exchange_atoms(atoms, cells, a, f)
    L = list(d.meas_global_length)
    z, y, X = np.nonzero(mask)
    # the periodic image in the sub-box and the box lengths crossed
    idx, shift = [], []
    for c, g0, s in ((X, gh[0], sub[0]), (y, gh[1], sub[1]), (z, gh[2], sub[2])):
        q = c - g0
        idx.append(g0 + np.mod(q, s))
        shift.append(np.floor_divide(q, s))
    img = box[idx[2], idx[1], idx[0]]
    ghost = box[z, y, X]
    np.testing.assert_array_equal(ghost["type"], img["type"])
    for k in range(3):
        np.testing.assert_array_equal(ghost["x"][:, k], img["x"][:, k] + shift[k] * L[k])


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_cpp_api(tmp_path):
    exe = tmp_path / "comm_logs_check"
    subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "misa-md_amd", "host"),
                    os.path.join(ROOT, "tests", "cpp", "comm_logs_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    assert out[0] == "domain %d %d 2" % (2 * (10 + 6), 12 + 6)
    assert out[1] == "multi-rank refused"
    per = lambda d: ["len%d0" % d, "send%d0" % d, "len%d1" % d, "send%d1" % d, "recv%d0" % d, "recv%d1" % d]
    assert out[2].split()[1:] == per(0) + per(1) + per(2)
    assert out[3].split()[1:] == per(2) + per(1) + per(0)
    assert out[4] == "[I] sim: step 3 of 10: 2.5"
    assert out[5] == "[E] dump: file x.out failed"
    assert len(out) == 6
