"""Atom dumps of the reference frontend (frontend/io) on the host: the
copy-mode binary layout that tools/dump2text/out_convert.cpp reads (128-byte
file header, 128-byte rank header holding the record count, 72-byte
AtomInfoDump records in blocks of 1024) and the debug-mode text dump.

The expected records are built here from the AtomElement array directly:
inter atoms in list order, then the occupied sub-box slots in the reference's
z, y, doubled-x order (atom_dump.cpp:66-84), each with the frame's step and
inter_type 0 (buffered_io.cpp:17-29).
"""
import os

import numpy as np
import pytest

import misa_eam as M

CELLS = (8, 8, 10)  # 1280 sites: a frame spans two 1024-record blocks


def _state(seed=7):
    atoms = M.host_build_world(CELLS, seed, 600.0, (90, 7, 3), seed)
    g = int(np.ceil(M.RC_FACTOR)) + 1
    dims = (2 * (CELLS[0] + 2 * g), CELLS[1] + 2 * g, CELLS[2] + 2 * g)
    sub = atoms.reshape(dims[2], dims[1], dims[0])[g:g + CELLS[2], g:g + CELLS[1], 2 * g:2 * g + 2 * CELLS[0]]
    rng = np.random.default_rng(seed)
    # a few vacancies (type INVALID), as decide leaves them
    flat = sub.reshape(-1)
    for k in rng.choice(flat.size, 11, replace=False):
        z, rem = divmod(int(k), sub.shape[1] * sub.shape[2])
        y, x = divmod(rem, sub.shape[2])
        sub[z, y, x]["type"] = -1
    inter = np.zeros(4, dtype=M.ATOM_DTYPE)
    inter["id"] = [9000001, 9000003, 9000002, 17]
    inter["type"] = [0, 1, 0, 2]
    inter["x"] = rng.uniform(0.0, 20.0, (4, 3))
    inter["v"] = rng.normal(size=(4, 3))
    return atoms, sub, inter


def _expected(sub, inter, step):
    lat = sub.reshape(-1)
    lat = lat[lat["type"] >= 0]
    both = np.concatenate([inter, lat])
    rec = np.zeros(len(both), dtype=M.DUMP_DTYPE)
    for k in ("id", "type", "x", "v"):
        rec[k] = both[k]
    rec["step"] = step
    return rec


def test_copy_mode_layout(tmp_path):
    atoms, sub, inter = _state()
    path = tmp_path / "md.out"
    d = M.AtomDumpFile(path)
    d.frame(atoms, inter, CELLS, 10)
    moved = atoms.copy()
    moved["x"] += 0.25
    d.frame(moved, inter[:2], CELLS, 20)
    n = d.close()
    e1 = _expected(sub, inter, 10)
    g = int(np.ceil(M.RC_FACTOR)) + 1
    dims = (2 * (CELLS[0] + 2 * g), CELLS[1] + 2 * g, CELLS[2] + 2 * g)
    sub2 = moved.reshape(dims[2], dims[1], dims[0])[g:g + CELLS[2], g:g + CELLS[1], 2 * g:2 * g + 2 * CELLS[0]]
    e2 = _expected(sub2, inter[:2], 20)
    assert n == len(e1) + len(e2) == 2 * (2 * np.prod(CELLS) - 11) + 6
    raw = np.fromfile(path, dtype=np.uint8)
    assert raw.size == 256 + n * 72
    assert not raw[:128].any() and not raw[136:256].any()
    assert int(raw[128:136].view("<u8")[0]) == n
    got = raw[256:].view(M.DUMP_DTYPE)
    np.testing.assert_array_equal(got, np.concatenate([e1, e2]))
    assert (got["inter_type"] == 0).all() and (got["_pad"] == 0).all()
    # the first frame starts with the inter atoms in list order
    np.testing.assert_array_equal(got["id"][:4], [9000001, 9000003, 9000002, 17])
    np.testing.assert_array_equal(M.read_dump(path), got)


def test_reader_interleaves_rank_blocks(tmp_path):
    """read_dump follows out_convert's block arithmetic: rank r's k-th block
    sits at block index n_ranks * k + r after the headers."""
    rec = np.zeros(2500, dtype=M.DUMP_DTYPE)
    rec["id"] = np.arange(2500)
    other = np.zeros(1100, dtype=M.DUMP_DTYPE)
    other["id"] = 100000 + np.arange(1100)
    bb = 1024 * 72
    blocks = {}
    for r, data in enumerate((rec, other)):
        for k in range(0, len(data), 1024):
            blocks[(k // 1024) * 2 + r] = data[k:k + 1024].tobytes()
    nb = max(blocks) + 1
    buf = bytearray(128 + 2 * 128 + nb * bb)
    buf[128:136] = np.uint64(len(rec)).tobytes()
    buf[256:264] = np.uint64(len(other)).tobytes()
    for i, b in blocks.items():
        o = 128 + 256 + i * bb
        buf[o:o + len(b)] = b
    path = tmp_path / "two.out"
    path.write_bytes(bytes(buf))
    a, b = M.read_dump(path, n_ranks=2)
    np.testing.assert_array_equal(a["id"], rec["id"])
    np.testing.assert_array_equal(b["id"], other["id"])


def test_empty_and_bad_calls(tmp_path):
    atoms, _, inter = _state(3)
    path = tmp_path / "empty.out"
    d = M.AtomDumpFile(path)
    assert d.close() == 0
    raw = np.fromfile(path, dtype=np.uint8)
    assert raw.size == 256 and not raw.any()
    with pytest.raises(M.MisaError):
        M.AtomDumpFile(tmp_path)  # a directory
    d = M.AtomDumpFile(tmp_path / "bad.out")
    with pytest.raises(M.MisaError):
        d.frame(atoms, inter, (0, 8, 10), 1)
    d.close()


def test_debug_mode_text(tmp_path):
    atoms, sub, inter = _state(5)
    path = tmp_path / "dump_0_40.atom"
    M.dump_text(path, atoms, inter, CELLS)
    lines = path.read_text().splitlines()
    lat = sub.reshape(-1)
    lat = lat[lat["type"] >= 0]
    want = ["print atoms"]
    want += ["%d %g %g %g" % (a["id"], a["x"][0], a["x"][1], a["x"][2]) for a in lat]
    want += ["print inter"]
    want += ["%d %g %g %g" % (a["id"], a["x"][0], a["x"][1], a["x"][2]) for a in inter]
    assert lines == want
    # std::ostream's default six significant digits
    assert lines[1].split()[1] == "%g" % lat[0]["x"][0]
    assert os.path.getsize(path) > 0
