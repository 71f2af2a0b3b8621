"""HIP path against the oracle on identical seeded inputs (needs an MI355X).

FP64 bars: rho and F' relative 1e-12; forces 1e-12 relative to the RMS force
(per-atom errors normalised by max(|f_i|, RMS|f|)); potential energy 1e-13
eV per atom.  Mixed precision: RMS relative force error <= 1e-5.
Integer work (offsets, ids, types, ownership) is compared bit-exactly.
"""
import sys

import numpy as np
import pytest

import misa_eam as M
import oracle_py as O

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _force_err(got, want):
    got = np.asarray(got).reshape(-1, 3)
    want = np.asarray(want).reshape(-1, 3)
    rms = np.sqrt(np.mean(np.sum(want * want, axis=1)))
    norm = np.maximum(np.linalg.norm(want, axis=1), rms)
    err = np.linalg.norm(got - want, axis=1) / norm
    return err.max(), np.sqrt(np.mean(err * err))


def _oracle(setfl, cells, seed, tset, ratio, disp=0.0, alloy_seed=1):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, tset, ratio, alloy_seed)
    if disp:
        sb = sim.sub_box()
        rng = np.random.default_rng(seed)
        sb["x"] = sb["x"] + rng.uniform(-disp, disp, sb["x"].shape)
    return sim


def _device(setfl, cells):
    dev = M.EamDevice(cells)
    dev.make_offsets()
    dev.set_potential(setfl)
    return dev


def test_hook_mode_matches_oracle_lattice_loops(setfl_fecuni):
    """misa_lat_rho / df / force on the reference AoS array versus latRho +
    rhoReverse, latDf, latForce + forceReverse of the oracle."""
    cells = (6, 7, 8)
    say("hook mode", cells)
    sim = _oracle(setfl_fecuni, cells, 3, 300.0, (90, 6, 4), disp=0.1, alloy_seed=3)
    sim.phase("exchange_atom_first")
    sim.phase("clear_force")
    aos = sim.atoms().copy()
    sim.phase("lat_rho")
    sim.phase("rho_reverse")
    dev = _device(setfl_fecuni, cells)
    dev.lat_rho(aos, sim.rc)
    want = sim.sub_box()
    got = M.sub_box_view(aos, dev)
    np.testing.assert_allclose(got["rho"], want["rho"], rtol=1e-12)
    sim.phase("lat_df")
    dev.lat_df(aos, sim.rc)
    np.testing.assert_allclose(got["df"], want["df"], rtol=1e-12)
    sim.phase("df_forward")
    aos = sim.atoms().copy()  # ghost F' forwarded by the reference's halo
    aos["f"] = 0.0
    sim.phase("lat_force")
    sim.phase("force_reverse")
    dev.lat_force(aos, sim.rc)
    mx, rms = _force_err(M.sub_box_view(aos, dev)["f"], sim.sub_box()["f"])
    say("hook force err max %.3e rms %.3e" % (mx, rms))
    assert mx < 1e-12


def test_resident_first_force_matches_oracle(setfl_fecuni):
    cells = (8, 8, 8)
    say("resident first force", cells)
    sim = _oracle(setfl_fecuni, cells, 5, 600.0, (985, 10, 5), disp=0.12, alloy_seed=5)
    dev = _device(setfl_fecuni, cells)
    dev.upload(sim.atoms().copy())
    dev.compute_forces()
    sim.prepare()
    aos, inter = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    np.testing.assert_allclose(got["rho"], want["rho"], rtol=1e-12)
    np.testing.assert_allclose(got["df"], want["df"], rtol=1e-12)
    mx, rms = _force_err(got["f"], want["f"])
    say("resident force err max %.3e rms %.3e" % (mx, rms))
    assert mx < 1e-12
    np.testing.assert_array_equal(got["type"], want["type"])
    np.testing.assert_array_equal(got["id"], want["id"])
    t = dev.thermo()
    e = sim.energy()
    n = 2 * np.prod(cells)
    assert abs(t["e_potential"] - e["e_potential"]) / n < 1e-13
    assert t["virial"] == pytest.approx(e["virial"], rel=1e-11, abs=1e-9)
    assert t["mvv"] == pytest.approx(sim.mvv(), rel=1e-13)
    assert t["temperature"] == pytest.approx(sim.temperature(), rel=1e-13)
    assert np.all(np.abs(t["force_sum"]) < 1e-9)


def test_device_world_builder_matches_oracle(setfl_fecuni):
    cells = (10, 9, 8)
    sim = O.Sim(cells, A, F, setfl_fecuni)
    sim.build_world(2020, 300.0, (985, 10, 5), 1)
    dev = _device(setfl_fecuni, cells)
    dev.build_world(2020, 300.0, (985, 10, 5), 1)
    aos, _ = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    for field in ("id", "type", "x", "v"):
        np.testing.assert_array_equal(got[field], want[field])


@pytest.mark.parametrize("cells, ratio, steps", [((20, 20, 20), (1, 0, 0), 100), ((8, 10, 12), (985, 10, 5), 50)])
def test_resident_trajectory_matches_oracle(setfl_fe, setfl_fecuni, cells, ratio, steps):
    """Whole velocity-Verlet steps on the device versus the oracle's loop body
    (C1: 20^3 Fe at 600 K, 100 steps of 1 fs)."""
    setfl = setfl_fe if ratio == (1, 0, 0) else setfl_fecuni
    say("trajectory", cells, ratio, steps)
    sim = _oracle(setfl, cells, 466953, 600.0, ratio)
    dev = _device(setfl, cells)
    dev.upload(sim.atoms().copy())
    dev.set_timestep(0.001)
    dev.compute_forces()
    dev.step(steps)
    sim.prepare()
    sim.run(steps)
    aos, inter = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    assert len(sim.inter()) == 0 and len(inter) == 0
    np.testing.assert_array_equal(got["type"], want["type"])
    dx = np.abs(got["x"] - want["x"]).max()
    dv = np.abs(got["v"] - want["v"]).max()
    mx, rms = _force_err(got["f"], want["f"])
    say("after %d steps: max|dx| %.3e max|dv| %.3e force err max %.3e" % (steps, dx, dv, mx))
    assert dx < 1e-9
    assert dv < 1e-7
    assert mx < 1e-9
    t = dev.thermo()
    assert t["temperature"] == pytest.approx(sim.temperature(), rel=1e-9)


@pytest.mark.parametrize("mode", [M.PREC_MIXED, M.PREC_MIXED_KAHAN])
def test_mixed_precision_force_error(setfl_fecuni, mode):
    cells = (10, 10, 10)
    sim = _oracle(setfl_fecuni, cells, 9, 300.0, (985, 10, 5), disp=0.1)
    dev = _device(setfl_fecuni, cells)
    dev.upload(sim.atoms().copy())
    dev.set_precision(mode)
    dev.compute_forces()
    sim.prepare()
    aos, _ = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    mx, rms = _force_err(got["f"], want["f"])
    say("mixed mode %d: force err max %.3e rms %.3e" % (mode, mx, rms))
    assert rms <= 1e-5
    np.testing.assert_allclose(got["rho"], want["rho"], rtol=1e-5)
