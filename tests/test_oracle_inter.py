"""The oracle's inter-atom path (decide, exchangeInter, borderInter, the
nearest-site buckets, interRho and interForce) against what the mathematics
fixes: an O(N^2) minimum-image pair sum over lattice and inter atoms
together, finite differences of the energy with respect to an inter atom's
position, Newton's third law, and energy conservation through a cascade.
Each check fails for a dropped ghost image, a wrong periodic shift, a missed
bucket or a pair counted twice.
"""
import numpy as np
import pytest

import oracle_py as O
from inter_cases import crafted_inter_state, oracle_forces_with_inter

A = 2.855
F = 5.6 / 2.855
Z = {0: 26, 1: 29, 2: 28}


def _crafted(setfl, cells, seed, ratio=(985, 10, 5)):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, 300.0, ratio, seed)
    sb = sim.sub_box()
    rng = np.random.default_rng(seed)
    sb["x"] = sb["x"] + rng.uniform(-0.08, 0.08, sb["x"].shape)
    inter = crafted_inter_state(sim, seed)
    oracle_forces_with_inter(sim, inter)
    return sim


def _all_atoms(sim):
    sb = sim.sub_box().reshape(-1)
    live = sb[sb["type"] >= 0]
    inter = sim.inter()
    return live, inter


def _brute(sim, x, t):
    pot = sim.potential()
    L = np.array(sim.cells) * A
    n = len(x)
    rc2 = sim.rc * sim.rc
    rho = np.zeros(n)
    pairs = []
    for i in range(n):
        d = x[i] - x[i + 1:]
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, axis=1)
        for jj in np.nonzero(r2 < rc2)[0]:
            j = i + 1 + jj
            rho[i] += pot.charge_density(Z[t[j]], r2[jj])
            rho[j] += pot.charge_density(Z[t[i]], r2[jj])
            pairs.append((i, j, d[jj], r2[jj]))
    df = np.array([pot.d_embed(Z[t[i]], rho[i]) for i in range(n)])
    f = np.zeros((n, 3))
    for i, j, d, r2 in pairs:
        fp = pot.to_force(Z[t[i]], Z[t[j]], r2, df[i], df[j])
        f[i] += d * fp
        f[j] -= d * fp
    return rho, df, f


def test_inter_state_brute_force(setfl_fecuni):
    """rho, F' and f of lattice and inter atoms against an O(N^2) minimum-image
    pair sum, with inter atoms migrated across faces and imaged as ghosts."""
    sim = _crafted(setfl_fecuni, (5, 6, 5), 21)
    live, inter = _all_atoms(sim)
    n0 = len(live)
    assert len(inter) == 15 + 4
    # exchangeInter wrapped the atoms pushed out of the box
    for at in inter:
        c = O.voronoy(at["x"], A)
        assert 0 <= c[0] < 2 * sim.cells[0] and 0 <= c[1] < sim.cells[1] and 0 <= c[2] < sim.cells[2]
    assert len(sim.inter_ghosts()) > 0
    x = np.concatenate([np.stack(live["x"]), np.stack(inter["x"])])
    t = np.concatenate([live["type"], inter["type"]]).astype(int)
    rho, df, f = _brute(sim, x, t)
    got_rho = np.concatenate([live["rho"], inter["rho"]])
    got_df = np.concatenate([live["df"], inter["df"]])
    got_f = np.concatenate([np.stack(live["f"]), np.stack(inter["f"])])
    np.testing.assert_allclose(got_rho, rho, rtol=1e-12)
    np.testing.assert_allclose(got_df, df, rtol=1e-12)
    scale = np.sqrt(np.mean(f * f))
    assert np.max(np.abs(got_f - f)) < 1e-11 * max(scale, 1.0)
    assert np.abs(got_f[n0:] - f[n0:]).max() < 1e-11 * max(scale, 1.0)


def test_inter_newton_third_law(setfl_fecuni):
    sim = _crafted(setfl_fecuni, (6, 6, 6), 5)
    assert np.all(np.abs(sim.system_force()) < 1e-9)


def test_inter_force_is_minus_energy_gradient(setfl_fecuni):
    """Energy finite differences along the coordinates of inter atoms (one of
    them a periodic neighbour of ghosts) against their computed force."""
    cells = (5, 5, 5)
    base = _crafted(setfl_fecuni, cells, 8)
    inter0 = base.inter().copy()
    atoms0 = base.atoms().copy()
    f0 = np.stack(inter0["f"]).copy()
    h = 1e-5

    def energy_with(inter):
        sim = O.Sim(cells, A, F, setfl_fecuni)
        sim.atoms()[:] = atoms0
        oracle_forces_with_inter(sim, inter)
        return sim.energy()["e_potential"]

    for q in (0, 2, len(inter0) - 1):
        for c in range(3):
            ip = inter0.copy()
            ip[q]["x"][c] += h
            im = inter0.copy()
            im[q]["x"][c] -= h
            fd = -(energy_with(ip) - energy_with(im)) / (2 * h)
            assert fd == pytest.approx(f0[q][c], abs=2e-6, rel=1e-5)


def test_cascade_energy_conservation(setfl_fe):
    """A 300 eV recoil in 6^3 Fe: inter atoms appear, migrate and recombine;
    total energy stays put at dt = 0.2 fs."""
    cells = (6, 6, 6)
    sim = O.Sim(cells, A, F, setfl_fe, dt=0.0002)
    sim.build_world(466953, 100.0, (1, 0, 0))
    sim.prepare()
    sim.collision([0, 0, 0, 0], [-1.0, -0.6, -0.3], 300.0)
    n = 2 * np.prod(cells)

    def etot():
        return (sim.mvv() / 2 * 1.0364269e-4 + sim.energy()["e_potential"]) / n

    e = [etot()]
    seen = 0
    for _ in range(20):
        sim.run(20)
        seen = max(seen, len(sim.inter()))
        e.append(etot())
    assert seen >= 2
    assert np.ptp(e) < 2e-4


def test_adaptive_dt_bounds_the_displacement(setfl_fe):
    """dt = min(dt_max, d / max|v|): the step the oracle takes equals the
    formula on its own velocities and keeps every displacement near d."""
    cells = (6, 6, 6)
    sim = O.Sim(cells, A, F, setfl_fe)
    sim.build_world(466953, 300.0, (1, 0, 0))
    sim.prepare()
    sim.collision([3, 3, 3, 0], [1.0, 3.0, 5.0], 5000.0)
    for _ in range(30):
        sb = sim.sub_box()
        live = sb["type"] >= 0
        inter = sim.inter()
        v = np.concatenate([sb["v"][live].reshape(-1, 3), inter["v"].reshape(-1, 3)])
        vmax = np.sqrt(np.max(np.sum(v * v, axis=1)))
        x0 = np.concatenate([sb["x"][live].reshape(-1, 3), inter["x"].reshape(-1, 3)])
        ids0 = np.concatenate([sb["id"][live], inter["id"]])
        dt = sim.adaptive_dt(0.001, 0.01)
        assert dt == pytest.approx(min(0.001, 0.01 / vmax), rel=1e-15)
        sim.run(1)
        sb = sim.sub_box()
        live = sb["type"] >= 0
        inter = sim.inter()
        x1 = np.concatenate([sb["x"][live].reshape(-1, 3), inter["x"].reshape(-1, 3)])
        ids1 = np.concatenate([sb["id"][live], inter["id"]])
        o0, o1 = np.argsort(ids0), np.argsort(ids1)
        d = x1[o1] - x0[o0]
        L = np.array(cells) * A
        d -= L * np.round(d / L)
        assert np.max(np.linalg.norm(d, axis=1)) < 0.0101
