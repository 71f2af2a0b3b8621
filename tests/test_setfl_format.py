"""Setfl files as production potentials lay them out (simulation.cpp:87-133
reads them through pot's SetflParser): any number of values a line, E
exponents, elements in any order (species are matched by atomic number,
atom_types.h:61-72), their own Nr/dr/Nrho/drho.  The host tables must not
depend on the layout, must follow the elements, and must equal the oracle's
for a non-default grid; broken files are refused."""
import numpy as np
import pytest

import misa_eam as M
import oracle_py as O
import setfl_gen

ELS = ("Fe", "Cu", "Ni")


def _tables(path, nel):
    return {(k, i): M.host_potential_table(path, k, i)
            for k, n in ((0, nel), (1, nel), (2, nel * (nel + 1) // 2)) for i in range(n)}


@pytest.mark.parametrize("per_line, fmt", [(1, "%24.16E"), (4, "%.16e"), (7, "%.17g")])
def test_layout_does_not_matter(tmp_path, per_line, fmt):
    a = setfl_gen.write_setfl(str(tmp_path / "a.eam.alloy"), ELS, nr=600, nrho=400)
    b = setfl_gen.write_setfl(str(tmp_path / "b.eam.alloy"), ELS, nr=600, nrho=400, per_line=per_line, fmt=fmt)
    ta, tb = _tables(a, 3), _tables(b, 3)
    for k in ta:
        np.testing.assert_array_equal(ta[k], tb[k])


def test_element_order_follows_the_file(tmp_path):
    a = setfl_gen.write_setfl(str(tmp_path / "a.eam.alloy"), ELS, nr=500, nrho=500)
    order = ("Ni", "Fe", "Cu")
    b = setfl_gen.write_setfl(str(tmp_path / "b.eam.alloy"), order, nr=500, nrho=500)
    ta, tb = _tables(a, 3), _tables(b, 3)
    ib = {e: i for i, e in enumerate(order)}
    for ia, e in enumerate(ELS):
        for k in (0, 1):
            np.testing.assert_array_equal(ta[(k, ia)], tb[(k, ib[e])])
    for i in range(3):
        for j in range(i + 1):
            p, q = sorted((ib[ELS[i]], ib[ELS[j]]), reverse=True)
            np.testing.assert_array_equal(ta[(2, i * (i + 1) // 2 + j)], tb[(2, p * (p + 1) // 2 + q)])


def test_own_grid_matches_oracle(tmp_path):
    path = setfl_gen.write_setfl(str(tmp_path / "g.eam.alloy"), ELS, nr=8000, nrho=3000, rho_max=60.0,
                                 per_line=4, fmt="%24.16E")
    pot = O.Potential(path)
    for (k, i), got in _tables(path, 3).items():
        assert got.shape == (7 * (8000 if k else 3000),)
        np.testing.assert_array_equal(got, pot.table(k, i))


@pytest.mark.parametrize("cut", ["truncate", "grid", "count", "word"])
def test_broken_files_are_refused(tmp_path, cut):
    good = setfl_gen.write_setfl(str(tmp_path / "ok.eam.alloy"), ("Fe",), nr=50, nrho=50)
    lines = open(good).read().splitlines()
    if cut == "truncate":
        lines = lines[:-3]
    elif cut == "grid":
        lines[4] = "50 0.1 50"
    elif cut == "count":
        lines[3] = "0"
    else:
        lines[9] = lines[9].replace(" ", " x ", 1)
    bad = tmp_path / "bad.eam.alloy"
    bad.write_text("\n".join(lines) + "\n")
    with pytest.raises(M.MisaError):
        M.host_potential_table(str(bad), 0, 0)
    with pytest.raises(M.MisaError):
        M.host_potential_table(str(tmp_path / "missing.eam.alloy"), 0, 0)
