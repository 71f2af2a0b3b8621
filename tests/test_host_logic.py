"""Host-side logic of the product library against the oracle (no GPU needed).

Covers: the C-ABI exports every declared symbol; NeighbourIndex offsets,
spline tables and the world builder are bit-identical to the oracle's; the
parity-plane decoding of the reference offsets and the pruned lattice list.
"""
import math
import os
import re

import numpy as np
import pytest

import misa_eam as M
import oracle_py as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
A = 2.855
F = 5.6 / 2.855


def declared_symbols():
    with open(os.path.join(ROOT, "include", "misa_eam.h")) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(misa_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    import ctypes
    lib = ctypes.CDLL(M.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 29
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(M.exported_symbols())


@pytest.mark.parametrize("cells, a, f", [((20, 20, 20), A, F), ((10, 12, 14), A, F), ((6, 7, 8), 2.85532, 1.96125),
                                         ((5, 5, 5), 0.86, 1.1421)])
def test_offsets_bit_identical_to_oracle(cells, a, f):
    g = math.ceil(f) + 1
    ref = O.nei_make(2 * (cells[0] + 2 * g), cells[1] + 2 * g, cells[2] + 2 * g, math.ceil(f), f)
    for which in range(4):
        got = M.host_offsets(cells, which, a, f)
        np.testing.assert_array_equal(got, ref[which])


def _decode(off, cells, g, parity):
    sx, sy = 2 * (cells[0] + 2 * g), cells[1] + 2 * g
    X = ((off + sx // 2) % sx) - sx // 2
    rest = (off - X) // sx
    y = ((rest + sy // 2) % sy) - sy // 2
    z = (rest - y) // sy
    pt = (parity + X) & 1
    dcx = (parity + X - pt) // 2
    h = 0.5 * (pt - parity)
    return pt, dcx, y, z, math.sqrt((dcx + h) ** 2 + (y + h) ** 2 + (z + h) ** 2)


def test_pruned_plane_offsets():
    """The lattice-lattice list keeps every reference offset with ideal separation
    below rc + 0.4a (+0.01a) and drops the rest (all beyond 2.37a, the next shell
    at sqrt(6) a = 2.449a), mapped to parity-plane slot deltas."""
    cells = (20, 20, 20)
    g = 3
    NX, NY, NZ = (c + 2 * g for c in cells)
    plane = NX * NY * NZ
    for parity in (0, 1):
        full = M.host_offsets(cells, parity)
        pruned = M.host_offsets(cells, 4 + parity)
        want = []
        dropped = []
        for off in full:
            pt, dcx, y, z, ln = _decode(int(off), cells, g, parity)
            d = (pt - parity) * plane + (z * NY + y) * NX + dcx
            (want if ln < F + 0.41 else dropped).append((d, ln))
        np.testing.assert_array_equal(pruned, [d for d, _ in want])
        assert len(pruned) == 112
        assert min(ln for _, ln in dropped) > F + 0.4 + 0.05
        assert max(ln for _, ln in want) < F + 0.4


def test_spline_tables_bit_identical_to_oracle(setfl_fecuni):
    pot = O.Potential(setfl_fecuni)
    for kind, count in ((0, 3), (1, 3), (2, 6)):
        for idx in range(count):
            got = M.host_potential_table(setfl_fecuni, kind, idx)
            want = pot.table(kind, idx)
            assert got.shape == want.shape
            np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("cells, seed, tset, ratio, alloy_seed", [
    ((20, 20, 20), 466953, 600.0, (1, 0, 0), 1),
    ((8, 9, 10), 2020, 300.0, (985, 10, 5), 1),
    ((6, 6, 6), 7, 1200.0, (980, 20, 0), 1234),
])
def test_world_builder_bit_identical_to_oracle(cells, seed, tset, ratio, alloy_seed):
    sim = O.Sim(cells, A, F, None)
    sim.build_world(seed, tset, ratio, alloy_seed)
    want = sim.sub_box()
    got = M.host_build_world(cells, seed, tset, ratio, alloy_seed)
    g = 3
    dims = (2 * (cells[0] + 2 * g), cells[1] + 2 * g, cells[2] + 2 * g)
    gs = got.reshape(dims[2], dims[1], dims[0])[g:g + cells[2], g:g + cells[1], 2 * g:2 * g + 2 * cells[0]]
    for field in ("id", "type", "x", "v"):
        np.testing.assert_array_equal(gs[field], want[field])
    if ratio[1] + ratio[2] > 0:
        assert np.any(gs["type"] > 0)


def test_setfl_generator_grid(setfl_fe):
    with open(setfl_fe) as fh:
        lines = [next(fh) for _ in range(6)]
    assert lines[3].split() == ["1", "Fe"]
    nrho, drho, nr, dr, cut = lines[4].split()
    assert int(nrho) == 5000 and int(nr) == 5000
    assert float(dr) * (int(nr) - 1) == pytest.approx(5.6, rel=1e-15)
    assert lines[5].split()[:3] == ["26", "55.844999999999999", "2.855"]
