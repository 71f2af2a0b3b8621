"""Inter atoms on the device against the oracle (needs an MI355X).

Covers the ownership pass (new inter atoms, vacancy recombination), the
migration of inter atoms across periodic faces, their ghost images, the
nearest-site buckets, the inter-atom density and force, and the lattice
sites that see inter atoms.  The list order is the reference's, so the inter
lists are compared entry by entry: ids, types and the vacancy pattern
bit-exactly, positions, velocities and forces to FP64 rounding (the device
sums each atom's terms in another order than the reference's half lists).
"""
import sys

import numpy as np
import pytest

import misa_eam as M
import oracle_py as O
from inter_cases import crafted_inter_state, oracle_forces_with_inter

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


def _device(setfl, cells):
    dev = M.EamDevice(cells)
    dev.make_offsets()
    dev.set_potential(setfl)
    return dev


def _crafted(setfl, cells, seed):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, 300.0, (985, 10, 5), seed)
    sb = sim.sub_box()
    rng = np.random.default_rng(seed)
    sb["x"] = sb["x"] + rng.uniform(-0.08, 0.08, sb["x"].shape)
    inter = crafted_inter_state(sim, seed)
    return sim, sim.atoms().copy(), inter


def _all_forces(sub, inter):
    live = sub["type"] >= 0
    return np.concatenate([sub["f"][live].reshape(-1, 3), inter["f"].reshape(-1, 3)])


def _compare_state(dev, sim, dx_tol, dv_tol, f_tol, what):
    aos, inter = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    wi = sim.inter()
    say("%s: inter %d / %d, vacancies %d" % (what, len(inter), len(wi), int((want["type"] < 0).sum())))
    np.testing.assert_array_equal(got["type"], want["type"])
    live = want["type"] >= 0
    np.testing.assert_array_equal(got["id"][live], want["id"][live])
    np.testing.assert_array_equal(inter["id"], wi["id"])
    np.testing.assert_array_equal(inter["type"], wi["type"])
    dx = max(np.abs(got["x"] - want["x"]).max(), np.abs(inter["x"] - wi["x"]).max() if len(wi) else 0.0)
    dv = max(np.abs(got["v"] - want["v"]).max(), np.abs(inter["v"] - wi["v"]).max() if len(wi) else 0.0)
    mx, rms = _force_err(_all_forces(got, inter), _all_forces(want, wi))
    say("%s: max|dx| %.3e max|dv| %.3e force err max %.3e rms %.3e" % (what, dx, dv, mx, rms))
    assert dx < dx_tol
    assert dv < dv_tol
    assert mx < f_tol
    return got, inter


def test_inter_upload_download_round_trip(setfl_fecuni):
    sim, atoms, inter = _crafted(setfl_fecuni, (6, 7, 8), 31)
    dev = _device(setfl_fecuni, (6, 7, 8))
    dev.upload(atoms, inter)
    aos, back = dev.download()
    for field in ("id", "type", "x", "v"):
        np.testing.assert_array_equal(back[field], inter[field])
    got = M.sub_box_view(aos, dev)
    np.testing.assert_array_equal(got["type"], sim.sub_box()["type"])


def test_inter_capacity_is_checked(setfl_fecuni):
    sim, atoms, inter = _crafted(setfl_fecuni, (6, 6, 6), 4)
    dev = _device(setfl_fecuni, (6, 6, 6))
    dev.set_inter_capacity(8)
    with pytest.raises(M.MisaError, match="capacity"):
        dev.upload(atoms, inter)


def test_inter_first_force_matches_oracle(setfl_fecuni):
    """Forces of a state holding 19 inter atoms (some outside the box, some
    on octahedral sites next to the faces) after migration and ghost images."""
    cells = (6, 7, 8)
    say("inter first force", cells)
    sim, atoms, inter = _crafted(setfl_fecuni, cells, 17)
    dev = _device(setfl_fecuni, cells)
    dev.upload(atoms, inter)
    dev.compute_forces()
    oracle_forces_with_inter(sim, inter)
    got, gi = _compare_state(dev, sim, 0.0 + 1e-300, 1e-300, 1e-12, "first force")
    want, wi = sim.sub_box(), sim.inter()
    live = want["type"] >= 0
    np.testing.assert_allclose(got["rho"][live], want["rho"][live], rtol=1e-12)
    np.testing.assert_allclose(got["df"][live], want["df"][live], rtol=1e-12)
    np.testing.assert_allclose(gi["rho"], wi["rho"], rtol=1e-12)
    np.testing.assert_allclose(gi["df"], wi["df"], rtol=1e-12)
    assert np.all(got["f"][~live] == 0.0)
    t = dev.thermo()
    e = sim.energy()
    assert t["n_inter"] == len(wi)
    assert t["n_vacancy"] == int((~live).sum())
    assert abs(t["e_potential"] - e["e_potential"]) / (live.sum() + len(wi)) < 1e-13
    assert t["virial"] == pytest.approx(e["virial"], rel=1e-10, abs=1e-8)
    assert t["mvv"] == pytest.approx(sim.mvv(), rel=1e-13)
    assert np.all(np.abs(t["force_sum"]) < 1e-8)


def test_inter_steps_from_crafted_state(setfl_fecuni):
    """30 steps from the crafted state: recombination into the vacated sites,
    new inter atoms, migration across faces."""
    cells = (6, 7, 8)
    sim, atoms, inter = _crafted(setfl_fecuni, cells, 23)
    dev = _device(setfl_fecuni, cells)
    dev.upload(atoms, inter)
    dev.set_timestep(0.001)
    dev.compute_forces()
    oracle_forces_with_inter(sim, inter)
    for k in range(3):
        dev.step(10)
        sim.run(10)
        _compare_state(dev, sim, 1e-9, 1e-7, 1e-9, "crafted +%d steps" % (10 * (k + 1)))


@pytest.mark.parametrize("lat, direction", [((0, 0, 0, 0), (-1.0, -0.77, -0.31)), ((5, 5, 5, 1), (0.3, 1.0, 0.55))])
def test_cascade_matches_oracle(setfl_fecuni, lat, direction):
    """A 1 keV recoil in 10^3 FeCuNi at 600 K, 150 steps of 1 fs: tens of inter
    atoms, vacancies, recombinations and (from the corner site) migrations
    across every face."""
    cells = (10, 10, 10)
    say("cascade", lat, direction)
    sim = O.Sim(cells, A, F, setfl_fecuni)
    sim.build_world(466953, 600.0, (985, 10, 5), 1)
    dev = _device(setfl_fecuni, cells)
    dev.upload(sim.atoms().copy())
    dev.set_timestep(0.001)
    dev.compute_forces()
    sim.prepare()
    dev.step(2)
    sim.run(2)
    dev.collision_step(lat, direction, 1000.0)
    sim.collision(lat, direction, 1000.0)
    peak = 0
    for k in range(5):
        dev.step(30)
        sim.run(30)
        peak = max(peak, len(sim.inter()))
        _compare_state(dev, sim, 1e-8, 1e-5, 1e-7, "cascade +%d steps" % (30 * (k + 1)))
    assert peak >= 10
    t = dev.thermo()
    assert t["temperature"] == pytest.approx(sim.temperature(), rel=1e-9)


@pytest.mark.parametrize("mode", [M.PREC_MIXED, M.PREC_MIXED_KAHAN])
def test_inter_mixed_precision_force_error(setfl_fecuni, mode):
    cells = (6, 7, 8)
    sim, atoms, inter = _crafted(setfl_fecuni, cells, 17)
    dev = _device(setfl_fecuni, cells)
    dev.set_precision(mode)
    dev.upload(atoms, inter)
    dev.compute_forces()
    oracle_forces_with_inter(sim, inter)
    aos, gi = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    wi = sim.inter()
    np.testing.assert_array_equal(gi["id"], wi["id"])
    mx, rms = _force_err(_all_forces(got, gi), _all_forces(want, wi))
    say("inter mixed mode %d: force err max %.3e rms %.3e" % (mode, mx, rms))
    assert rms <= 1e-5
    np.testing.assert_allclose(gi["rho"], wi["rho"], rtol=1e-5)


def test_inter_rescale(setfl_fecuni):
    cells = (6, 7, 8)
    sim, atoms, inter = _crafted(setfl_fecuni, cells, 9)
    dev = _device(setfl_fecuni, cells)
    dev.upload(atoms, inter)
    oracle_forces_with_inter(sim, inter)
    dev.compute_forces()
    t0 = dev.thermo()
    assert t0["temperature"] == pytest.approx(sim.temperature(), rel=1e-13)
    dev.rescale(450.0)
    sim.rescale(450.0)
    _, gi = dev.download()
    np.testing.assert_allclose(gi["v"], sim.inter()["v"], rtol=1e-14)
    assert dev.thermo()["temperature"] == pytest.approx(450.0, rel=1e-12)


@pytest.mark.parametrize("mode", [M.PREC_FP64, M.PREC_MIXED])
def test_far_lattice_atoms_use_the_full_list(setfl_fecuni, mode):
    """Lattice atoms up to 0.45a from their sites (as a refilled vacancy can
    be) interact with sites beyond the pruned offset list; the sites around
    them must visit the whole reference list."""
    cells = (7, 6, 8)
    sim = O.Sim(cells, A, F, setfl_fecuni)
    sim.build_world(77, 300.0, (985, 10, 5), 7)
    sb = sim.sub_box()
    rng = np.random.default_rng(77)
    faces = [(0, 3, 1), (cells[2] - 1, 0, 2 * cells[0] - 1), (4, cells[1] - 1, 0), (cells[2] - 1, cells[1] - 1, 5)]
    for q in range(12):
        if q < len(faces):  # next to faces: the partners are periodic images
            k, j, i = faces[q]
        else:
            k, j, i = rng.integers(cells[2]), rng.integers(cells[1]), rng.integers(2 * cells[0])
        u = np.zeros(3)
        u[rng.integers(3)] = rng.choice([-1.0, 1.0]) * 0.45 * A
        sb[k, j, i]["x"] = sb[k, j, i]["x"] + u
    atoms = sim.atoms().copy()
    dev = _device(setfl_fecuni, cells)
    dev.set_precision(mode)
    dev.upload(atoms)
    dev.compute_forces()
    sim.prepare()
    aos, _ = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    mx, rms = _force_err(got["f"], want["f"])
    say("far lattice atoms, mode %d: force err max %.3e rms %.3e" % (mode, mx, rms))
    if mode == M.PREC_FP64:
        assert mx < 1e-12
        np.testing.assert_allclose(got["rho"], want["rho"], rtol=1e-12)
        t, e = dev.thermo(), sim.energy()
        assert abs(t["e_potential"] - e["e_potential"]) / (2 * np.prod(cells)) < 1e-13
    else:
        assert rms < 1e-5


def test_adaptive_dt_cascade_matches_oracle(setfl_fecuni):
    """A 5 keV recoil along [1 3 5] with dt = min(1 fs, 0.01 A / max|v|) taken
    every step on each side from its own velocities."""
    cells = (10, 10, 10)
    sim = O.Sim(cells, A, F, setfl_fecuni)
    sim.build_world(466953, 300.0, (985, 10, 5), 1)
    dev = _device(setfl_fecuni, cells)
    dev.upload(sim.atoms().copy())
    dev.set_timestep(0.001)
    dev.compute_forces()
    sim.prepare()
    lat, direction = (5, 5, 5, 0), (1.0, 3.0, 5.0)
    dev.collision_step(lat, direction, 5000.0)
    sim.collision(lat, direction, 5000.0)
    dev.set_adaptive_dt(0.001, 0.01)
    t_oracle = 0.0
    for k in range(4):
        for _ in range(25):
            t_oracle += sim.adaptive_dt(0.001, 0.01)
            sim.run(1)
        dev.step(25)
        t_dev, dt_dev = dev.sim_time()
        say("adaptive +%d steps: t %.6e ps (oracle %.6e), last dt %.3e" % (25 * (k + 1), t_dev, t_oracle, dt_dev))
        assert t_dev == pytest.approx(t_oracle, rel=1e-12)
        _compare_state(dev, sim, 1e-9, 1e-5, 1e-8, "adaptive +%d steps" % (25 * (k + 1)))
    assert dt_dev < 2e-5  # the recoil sets the step


def test_four_simultaneous_recoils(setfl_fecuni):
    """Several PKAs in one stage: setv per recoil, then one force evaluation."""
    cells = (10, 10, 10)
    sim = O.Sim(cells, A, F, setfl_fecuni)
    sim.build_world(2024, 300.0, (985, 10, 5), 1)
    dev = _device(setfl_fecuni, cells)
    dev.upload(sim.atoms().copy())
    dev.set_timestep(0.0002)
    dev.compute_forces()
    sim.set_dt(0.0002)
    sim.prepare()
    rng = np.random.default_rng(4)
    pkas = []
    for q in range(4):
        lat = (int(rng.integers(10)), int(rng.integers(10)), int(rng.integers(10)), int(rng.integers(2)))
        u = rng.normal(size=3)
        pkas.append((lat, tuple(u / np.linalg.norm(u))))
    for lat, d in pkas[:-1]:
        dev.setv(lat, d, 500.0)
        sim.setv(lat, d, 500.0)
    dev.collision_step(pkas[-1][0], pkas[-1][1], 500.0)
    sim.collision(pkas[-1][0], pkas[-1][1], 500.0)
    for k in range(3):
        dev.step(20)
        sim.run(20)
        _compare_state(dev, sim, 1e-9, 1e-6, 1e-8, "4 recoils +%d steps" % (20 * (k + 1)))
