"""The committed golden fixtures (tests/golden/, written by make_golden.py
from the oracle): the oracle must reproduce them bit for bit, and the device
must reach them from the same initial state to the parity bars of the other
device tests.
"""
import os

import numpy as np
import pytest

import oracle_py as O
from inter_cases import oracle_forces_with_inter

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
A = 2.855
F = 5.6 / 2.855
FIELDS = ("id", "type", "x", "v", "f", "rho", "df")


def _load(name):
    with np.load(os.path.join(HERE, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_oracle_reproduces_lattice_fixture(setfl_fecuni):
    g = _load("fecuni_4c_10steps.npz")
    sim = O.Sim(tuple(g["cells"]), A, F, setfl_fecuni)
    sim.atoms()[:] = g["init_aos"]
    sim.prepare()
    for k in FIELDS:
        np.testing.assert_array_equal(sim.sub_box()[k], g["first_" + k])
    sim.run(10)
    for k in FIELDS:
        np.testing.assert_array_equal(sim.sub_box()[k], g["last_" + k])


def test_oracle_reproduces_inter_fixture(setfl_fecuni):
    g = _load("fecuni_inter_12steps.npz")
    sim = O.Sim(tuple(g["cells"]), A, F, setfl_fecuni)
    sim.atoms()[:] = g["init_aos"]
    oracle_forces_with_inter(sim, g["init_inter"])
    for k in FIELDS:
        np.testing.assert_array_equal(sim.sub_box()[k], g["first_" + k])
        np.testing.assert_array_equal(sim.inter()[k], g["first_inter_" + k])
    sim.run(12)
    for k in FIELDS:
        np.testing.assert_array_equal(sim.sub_box()[k], g["last_" + k])
        np.testing.assert_array_equal(sim.inter()[k], g["last_inter_" + k])


def _ferr(got, want):
    got = np.asarray(got).reshape(-1, 3)
    want = np.asarray(want).reshape(-1, 3)
    rms = np.sqrt(np.mean(np.sum(want * want, axis=1)))
    return np.max(np.linalg.norm(got - want, axis=1) / np.maximum(np.linalg.norm(want, axis=1), rms))


@pytest.mark.gpu
def test_device_reaches_lattice_fixture(setfl_fecuni):
    import misa_eam as M
    g = _load("fecuni_4c_10steps.npz")
    dev = M.EamDevice(tuple(g["cells"]))
    dev.make_offsets()
    dev.set_potential(setfl_fecuni)
    dev.upload(g["init_aos"])
    dev.set_timestep(0.001)
    dev.compute_forces()
    got = M.sub_box_view(dev.download()[0], dev)
    assert _ferr(got["f"], g["first_f"]) < 1e-12
    np.testing.assert_allclose(got["rho"], g["first_rho"], rtol=1e-12)
    dev.step(10)
    got = M.sub_box_view(dev.download()[0], dev)
    np.testing.assert_array_equal(got["type"], g["last_type"])
    np.testing.assert_array_equal(got["id"], g["last_id"])
    assert np.abs(got["x"] - g["last_x"]).max() < 1e-11
    assert np.abs(got["v"] - g["last_v"]).max() < 1e-9
    assert _ferr(got["f"], g["last_f"]) < 1e-10


@pytest.mark.gpu
def test_device_reaches_inter_fixture(setfl_fecuni):
    import misa_eam as M
    g = _load("fecuni_inter_12steps.npz")
    dev = M.EamDevice(tuple(g["cells"]))
    dev.make_offsets()
    dev.set_potential(setfl_fecuni)
    dev.upload(g["init_aos"], g["init_inter"])
    dev.set_timestep(0.001)
    dev.compute_forces()
    aos, inter = dev.download()
    np.testing.assert_array_equal(inter["id"], g["first_inter_id"])
    assert _ferr(inter["f"], g["first_inter_f"]) < 1e-12
    dev.step(12)
    aos, inter = dev.download()
    got = M.sub_box_view(aos, dev)
    np.testing.assert_array_equal(got["type"], g["last_type"])
    np.testing.assert_array_equal(inter["id"], g["last_inter_id"])
    np.testing.assert_array_equal(inter["type"], g["last_inter_type"])
    assert np.abs(got["x"] - g["last_x"]).max() < 1e-10
    assert np.abs(inter["x"] - g["last_inter_x"]).max() < 1e-10
    assert _ferr(inter["f"], g["last_inter_f"]) < 1e-9
