"""The oracle against what the mathematics fixes.

The reference's EAM arithmetic (pot v0.1.0) is not in its tree, so the
oracle's interpolation, density, embedding and pair force are pinned here
against the closed forms the synthetic tables are sampled from, the perfect
BCC crystal (closed-form density, zero force), an O(N^2) brute-force pair
sum over minimum images, finite differences of the oracle's own energy,
Newton's third law, time reversibility of velocity Verlet and the virial
theorem under uniform dilation.
"""
import math

import numpy as np
import pytest

import oracle_py as O
import setfl_gen as G

A = 2.855
F = 5.6 / 2.855
Z = {0: 26, 1: 29, 2: 28}
EL = {0: "Fe", 1: "Cu", 2: "Ni"}


def test_spline_matches_closed_forms(setfl_fecuni):
    pot = O.Potential(setfl_fecuni)
    rng = np.random.default_rng(1)
    for r in rng.uniform(1.2, 5.55, 200):
        for t in range(3):
            assert pot.charge_density(Z[t], r * r) == pytest.approx(float(G.rho_fn(EL[t], r)), rel=1e-9, abs=1e-12)
        for a in range(3):
            for b in range(3):
                assert pot.pair_phi(Z[a], Z[b], r * r) == pytest.approx(float(G.phi_fn(EL[a], EL[b], r)),
                                                                        rel=1e-8, abs=1e-11)
    for rho in rng.uniform(2.0, 35.0, 200):
        for t in range(3):
            assert pot.d_embed(Z[t], rho) == pytest.approx(float(G.dembed_fn(EL[t], rho)), rel=1e-9)
            assert pot.embed(Z[t], rho) == pytest.approx(float(G.embed_fn(EL[t], rho)), rel=1e-11)


def test_to_force_closed_form(setfl_fecuni):
    """fpair = -(phi'_ij + F'_i rho'_j + F'_j rho'_i) / r with distinct species and F'."""
    pot = O.Potential(setfl_fecuni)
    rng = np.random.default_rng(2)
    for _ in range(300):
        a, b = rng.integers(0, 3, 2)
        r = rng.uniform(1.5, 5.5)
        dfi, dfj = rng.uniform(-0.5, -0.05, 2)
        want = -(float(G.dphi_fn(EL[a], EL[b], r)) + dfi * float(G.drho_fn(EL[b], r))
                 + dfj * float(G.drho_fn(EL[a], r))) / r
        got = pot.to_force(Z[a], Z[b], r * r, dfi, dfj)
        assert got == pytest.approx(want, rel=1e-7, abs=1e-10)


def test_table_edges_clamp(setfl_fe):
    pot = O.Potential(setfl_fe)
    last = float(G.rho_fn("Fe", 5.6))  # tapered to zero
    assert abs(pot.charge_density(26, 5.6 ** 2)) < 1e-12
    assert abs(pot.charge_density(26, 7.0 ** 2) - last) < 1e-12
    # F' below the first knot uses the first segment (clamped knot index)
    assert math.isfinite(pot.d_embed(26, 0.0))


def _shells(a):
    out = {}
    for i in range(-3, 4):
        for j in range(-3, 4):
            for k in range(-3, 4):
                for h in (0.0, 0.5):
                    d = math.sqrt((i + h) ** 2 + (j + h) ** 2 + (k + h) ** 2) * a
                    if 0 < d < 5.6:
                        key = round(d, 9)
                        out[key] = out.get(key, 0) + 1
    return out


def test_perfect_bcc_density_and_zero_force(setfl_fe):
    sim = O.Sim((6, 6, 6), A, F, setfl_fe)
    sim.build_world(5, 0.0, (1, 0, 0))
    sim.prepare()
    sb = sim.sub_box()
    rho_closed = sum(n * float(G.rho_fn("Fe", d)) for d, n in _shells(A).items())
    assert sum(_shells(A).values()) == 58
    np.testing.assert_allclose(sb["rho"], rho_closed, rtol=2e-9)  # spline vs closed form
    assert np.max(np.abs(np.stack(sb["f"].reshape(-1)))) < 1e-11


def _displaced_sim(setfl, cells, seed, amp=0.08, ratio=(90, 6, 4)):
    sim = O.Sim(cells, A, F, setfl)
    sim.build_world(seed, 300.0, ratio, alloy_seed=seed)
    sb = sim.sub_box()
    rng = np.random.default_rng(seed)
    x = sb["x"] + rng.uniform(-amp, amp, sb["x"].shape)
    sb["x"] = x
    sim.prepare()
    return sim


def test_brute_force_pair_sums(setfl_fecuni):
    """rho, F' and f of every atom against an O(N^2) minimum-image pair sum
    evaluated with the same pot functions."""
    cells = (5, 5, 5)
    sim = _displaced_sim(setfl_fecuni, cells, 11)
    pot = sim.potential()
    sb = sim.sub_box().reshape(-1)
    x = np.stack(sb["x"])
    t = sb["type"].astype(int)
    L = np.array(cells) * A
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
    assert abs(len(pairs) - 58 * n // 2) < 0.05 * 58 * n // 2  # displacements move pairs across rc
    np.testing.assert_allclose(sb["rho"], rho, rtol=1e-13)
    df = np.array([pot.d_embed(Z[t[i]], rho[i]) for i in range(n)])
    np.testing.assert_allclose(sb["df"], df, rtol=1e-13)
    f = np.zeros((n, 3))
    for i, j, d, r2 in pairs:
        fp = pot.to_force(Z[t[i]], Z[t[j]], r2, df[i], df[j])
        f[i] += d * fp
        f[j] -= d * fp
    got = np.stack(sb["f"])
    scale = np.sqrt(np.mean(f * f))
    assert np.max(np.abs(got - f)) < 1e-12 * max(scale, 1.0)


def _energy_of(sim):
    sim.phase("exchange_atom")
    sim.phase("clear_force")
    sim.phase("compute_eam")
    return sim.energy()["e_potential"]


def test_force_is_minus_energy_gradient(setfl_fecuni):
    sim = _displaced_sim(setfl_fecuni, (4, 4, 4), 3)
    sb = sim.sub_box()
    f0 = np.stack(sb.reshape(-1)["f"]).copy()
    h = 1e-5
    rng = np.random.default_rng(4)
    for flat in rng.choice(sb.size, 6, replace=False):
        k, j, i = np.unravel_index(flat, sb.shape)
        for c in range(3):
            x0 = sb[k, j, i]["x"][c]
            sb[k, j, i]["x"][c] = x0 + h
            ep = _energy_of(sim)
            sb[k, j, i]["x"][c] = x0 - h
            em = _energy_of(sim)
            sb[k, j, i]["x"][c] = x0
            fd = -(ep - em) / (2 * h)
            assert fd == pytest.approx(f0[flat][c], abs=2e-7, rel=1e-6)


def test_newton_third_law(setfl_fecuni):
    sim = _displaced_sim(setfl_fecuni, (6, 6, 6), 7)
    fs = sim.system_force()
    assert np.all(np.abs(fs) < 1e-10)  # the reference's DEV check is |sum f_x| <= 1e-5 (md_simulation.cpp:157)


def test_virial_is_minus_dilation_derivative(setfl_fecuni):
    """W = sum_pairs r.f = -dE/d(eps) for x -> (1+eps) x with the box scaled alike."""
    cells = (4, 4, 4)
    base = _displaced_sim(setfl_fecuni, cells, 9, amp=0.05)
    W = base.energy()["virial"]
    x0 = np.stack(base.sub_box().reshape(-1)["x"]).copy()
    types = base.sub_box().reshape(-1)["type"].copy()
    eps = 1e-6
    E = []
    for s in (1 + eps, 1 - eps):
        a = A * s
        sim = O.Sim(cells, a, 5.6 / a, setfl_fecuni)
        sim.build_world(1, 0.0, (1, 0, 0))
        sb = sim.sub_box()  # a strided view: write through it without reshaping
        sb["x"] = (x0 * s).reshape(sb["x"].shape)
        sb["type"] = types.reshape(sb.shape)
        sim.prepare()
        E.append(sim.energy()["e_potential"])
    dE = (E[0] - E[1]) / (2 * eps)
    assert W == pytest.approx(-dE, rel=1e-6)


def test_verlet_time_reversible(setfl_fe):
    sim = O.Sim((6, 6, 6), A, F, setfl_fe)
    sim.build_world(466953, 300.0, (1, 0, 0))
    sim.prepare()
    x0 = np.stack(sim.sub_box().reshape(-1)["x"]).copy()
    sim.run(40)
    sb = sim.sub_box()
    sb["v"] = -sb["v"]
    sim.run(40)
    x1 = np.stack(sim.sub_box().reshape(-1)["x"])
    assert np.max(np.abs(x1 - x0)) < 1e-9
    assert len(sim.inter()) == 0


def test_nve_energy_conservation(setfl_fe):
    sim = O.Sim((6, 6, 6), A, F, setfl_fe)
    sim.build_world(466953, 600.0, (1, 0, 0))
    sim.prepare()
    n = 2 * 6 ** 3
    e = []
    for _ in range(10):
        e.append(sim.mvv() / 2 * 1.0364269e-4 + sim.energy()["e_potential"])
        sim.run(10)
    e = np.array(e) / n
    assert np.ptp(e) < 2e-4  # eV/atom over 100 fs at dt = 1 fs
