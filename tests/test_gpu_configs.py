"""The HIP path at the full sizes of the five workloads (needs an MI355X).

* 100^3 FeCuNi (2M atoms, 985:10:5, 300 K, seed 2020): FP64 against the
  oracle (first forces, 20 velocity-Verlet steps); FP32-mixed forces against
  the oracle; the LDS force pass (which at this size walks several virtual
  blocks per persistent block) and the neighbour skin against their
  switched-off forms, bit for bit, through a recoil and adaptive steps.
* 200^3 (16M atoms): a perfect Fe lattice against the closed-form density and
  zero force; the 5 keV [1 3 5] cascade stepped until interstitials form, with
  and without the LDS pass, bit for bit; atom conservation (every id once,
  one vacancy per interstitial); energy conservation; FP64 momentum
  conservation and the mixed-vs-FP64 force error on the developed cascade.
* 64 boxes of 10^3..60^3 cells stepped as a batch against the same boxes
  stepped one by one, bit for bit.
* 300^3 (54M atoms), equilibrated, with four 10 keV recoils: LDS on/off
  bit for bit, energy conservation, FP64 momentum conservation,
  mixed-vs-FP64 forces.

Sizes that the oracle cannot step in seconds are checked through properties
that do not depend on size.
"""
import math
import os
import sys
import time

import numpy as np
import pytest

import misa_eam as M
import oracle_py as O

pytestmark = pytest.mark.gpu

A = 2.855
F = 5.6 / 2.855
FIELDS = ("id", "type", "x", "v", "f", "rho", "df")
MVV2E = 1.0364269e-4  # amu A^2/ps^2 -> eV (pre_define.h)
PKA_DIR = (1.0, 3.0, 5.0)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _force_err(got, want):
    got = np.asarray(got).reshape(-1, 3)
    want = np.asarray(want).reshape(-1, 3)
    rms = np.sqrt(np.mean(np.sum(want * want, axis=1)))
    norm = np.maximum(np.linalg.norm(want, axis=1), rms)
    err = np.linalg.norm(got - want, axis=1) / norm
    return float(err.max()), float(np.sqrt(np.mean(err * err)))


def _device(setfl, cells, precision=M.PREC_FP64, lds=M.LDS_AUTO, skin=0.2):
    dev = M.EamDevice(cells)
    dev.make_offsets()
    dev.set_potential(setfl)
    dev.set_precision(precision)
    dev.set_lds_tables(lds)
    dev.set_neighbour_skin(skin)
    return dev


def _same(a, b, what):
    sa, ia = a.download()
    sb, ib = b.download()
    va, vb = M.sub_box_view(sa, a), M.sub_box_view(sb, b)
    for k in FIELDS:
        np.testing.assert_array_equal(va[k], vb[k], err_msg="%s: %s" % (what, k))
        np.testing.assert_array_equal(ia[k], ib[k], err_msg="%s: inter %s" % (what, k))
    assert a.sim_time() == b.sim_time(), what
    return len(ia)


def _e_total(dev):
    th = dev.thermo()
    n = 2 * int(np.prod(dev.cells))
    return (th["kinetic"] * MVV2E + th["e_potential"]) / n, th


def _conserved(dev):
    """Every atom id once among the occupied sites and the inter atoms, and
    one vacancy per interstitial."""
    aos, inter = dev.download()
    sub = M.sub_box_view(aos, dev)
    live = sub["type"] >= 0
    ids = np.concatenate([sub["id"][live].ravel(), inter["id"]])
    n = 2 * int(np.prod(dev.cells))
    assert ids.size == n
    assert np.array_equal(np.sort(ids), np.arange(1, n + 1, dtype=ids.dtype))
    assert int((~live).sum()) == len(inter)
    return len(inter)


# --------------------------------------------------------------------------
# 100^3 cells, 2M atoms
# --------------------------------------------------------------------------
C2 = (100, 100, 100)


@pytest.fixture(scope="module")
def c2_oracle(setfl_fecuni):
    say("C2: oracle world 100^3")
    sim = O.Sim(C2, A, F, setfl_fecuni)
    sim.build_world(2020, 300.0, (985, 10, 5), 1)
    return sim


def test_c2_fp64_forces_and_steps_match_oracle(setfl_fecuni, c2_oracle):
    sim = c2_oracle
    atoms = sim.atoms().copy()
    dev = _device(setfl_fecuni, C2)
    dev.upload(atoms)
    dev.set_timestep(0.001)
    dev.compute_forces()
    t0 = time.time()
    sim.prepare()
    say("C2: oracle first force %.1f s" % (time.time() - t0))
    aos, _ = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    np.testing.assert_allclose(got["rho"], want["rho"], rtol=1e-12)
    np.testing.assert_allclose(got["df"], want["df"], rtol=1e-12)
    mx, rms = _force_err(got["f"], want["f"])
    say("C2 FP64 first force: err max %.3e rms %.3e" % (mx, rms))
    assert mx < 1e-12
    th, e = dev.thermo(), sim.energy()
    assert abs(th["e_potential"] - e["e_potential"]) / (2 * 100 ** 3) < 1e-13
    del aos, got
    dev.step(20)
    t0 = time.time()
    sim.run(20)
    say("C2: oracle 20 steps %.1f s" % (time.time() - t0))
    aos, inter = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    assert len(inter) == 0 and len(sim.inter()) == 0
    np.testing.assert_array_equal(got["type"], want["type"])
    np.testing.assert_array_equal(got["id"], want["id"])
    dx = float(np.abs(got["x"] - want["x"]).max())
    dv = float(np.abs(got["v"] - want["v"]).max())
    mx, rms = _force_err(got["f"], want["f"])
    say("C2 FP64 after 20 steps: max|dx| %.3e max|dv| %.3e force err max %.3e" % (dx, dv, mx))
    assert dx < 1e-10 and dv < 1e-8 and mx < 1e-10
    dev.close()


def test_c2_mixed_forces_match_oracle(setfl_fecuni):
    """The mixed first forces (LDS pass on, as the library picks it at this
    size) against the oracle's FP64 forces of the same state."""
    sim = O.Sim(C2, A, F, setfl_fecuni)
    sim.build_world(2020, 300.0, (985, 10, 5), 1)
    sim.prepare()
    dev = _device(setfl_fecuni, C2, M.PREC_MIXED_KAHAN)
    dev.upload(sim.atoms().copy())
    dev.compute_forces()
    aos, _ = dev.download()
    got, want = M.sub_box_view(aos, dev), sim.sub_box()
    mx, rms = _force_err(got["f"], want["f"])
    say("C2 mixed first force vs oracle: err max %.3e rms %.3e" % (mx, rms))
    assert rms <= 1e-5 and mx <= 1e-4
    np.testing.assert_allclose(got["rho"], want["rho"], rtol=1e-5)
    dev.close()


@pytest.mark.parametrize("what", ["lds", "skin"])
def test_c2_switches_give_one_trajectory(setfl_fecuni, what):
    """At 100^3 every persistent block of the LDS pass walks about eight
    virtual blocks, carrying its velocity and displacement maxima and the
    wave-mask ids between them: LDS on against off (and skin 0.2 against
    none), bit for bit, through fixed steps, a recoil and adaptive steps."""
    say("C2 switch", what)
    devs = []
    for on in (True, False):
        if what == "lds":
            d = _device(setfl_fecuni, C2, M.PREC_MIXED_KAHAN, lds=M.LDS_ON if on else M.LDS_OFF)
        else:
            d = _device(setfl_fecuni, C2, M.PREC_MIXED_KAHAN, skin=0.2 if on else 0.0)
        d.build_world(2020, 300.0, (985, 10, 5), 1)
        d.set_timestep(0.001)
        d.compute_forces()
        devs.append(d)
    a, b = devs
    _same(a, b, "first forces")
    for d in devs:
        d.step(10)
    _same(a, b, "10 fixed steps")
    for d in devs:
        d.collision_step((50, 50, 50, 0), PKA_DIR, 2000.0)
        d.set_adaptive_dt(0.001, 0.01)
        d.step(30)
    _same(a, b, "recoil and 30 adaptive steps")
    for d in devs:
        d.close()


# --------------------------------------------------------------------------
# 200^3 cells, 16M atoms
# --------------------------------------------------------------------------
C3 = (200, 200, 200)


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_MIXED_KAHAN], ids=["fp64", "mixed"])
def test_c3_perfect_lattice_closed_form(setfl_fe, precision):
    """Every site of a perfect Fe lattice has rho = the sum over the five
    shells within the cutoff (8, 6, 12, 24, 8 neighbours at sqrt(3)/2, 1,
    sqrt(2), sqrt(11)/2, sqrt(3) lattice constants) and zero force."""
    pot = O.Potential(setfl_fe)
    shells = [(8, math.sqrt(3) / 2), (6, 1.0), (12, math.sqrt(2)), (24, math.sqrt(11) / 2), (8, math.sqrt(3))]
    rho0 = sum(n * pot.charge_density(26, (r * A) ** 2) for n, r in shells)
    assert (math.sqrt(3) * A) < 5.6 < 2 * A  # the fifth shell is inside the cutoff, the sixth outside
    dev = _device(setfl_fe, C3, precision)
    dev.build_world(1, 0.0, (1, 0, 0), 1)
    dev.compute_forces()
    aos, _ = dev.download()
    sub = M.sub_box_view(aos, dev)
    rho, f = sub["rho"].ravel(), sub["f"].reshape(-1, 3)
    # FP64: positions up to 571 A carry rounding of ulp(571) = 1.1e-13 A, which
    # moves a 2.47 A distance by 5e-14 relative and rho by about 1e-13
    tol = 1e-12 if precision == M.PREC_FP64 else 2e-6
    say("C3 perfect lattice: rho0 %.15g, max rel dev %.3e, max |f| %.3e" %
        (rho0, float(np.abs(rho / rho0 - 1).max()), float(np.abs(f).max())))
    assert np.all(np.abs(rho / rho0 - 1) < tol)
    assert np.all(np.abs(f) < (1e-11 if precision == M.PREC_FP64 else 1e-5))
    dev.close()


def test_c3_developed_cascade_properties(setfl_fecuni):
    say("C3: cascade 200^3, LDS on and off")
    devs = [_device(setfl_fecuni, C3, M.PREC_MIXED_KAHAN, lds=m) for m in (M.LDS_AUTO, M.LDS_OFF)]
    for d in devs:
        d.build_world(2020, 300.0, (985, 10, 5), 1)
        d.set_timestep(0.001)
        d.compute_forces()
        d.step(50)
        d.rescale(300.0)
        d.collision_step((100, 100, 100, 0), PKA_DIR, 5000.0)
        d.set_adaptive_dt(0.001, 0.01)
    a, b = devs
    e0, _ = _e_total(a)
    steps = 0
    while True:  # until interstitials have formed (about 0.13 ps)
        for d in devs:
            d.step(500)
        steps += 500
        e1, th = _e_total(a)
        say("C3: %d steps, t %.4f ps, %d inter, dE %.3e eV/atom" % (steps, a.sim_time()[0], th["n_inter"], e1 - e0))
        if th["n_inter"] >= 10 or steps >= 5000:
            break
    assert th["n_inter"] >= 10
    ni = _same(a, b, "developed cascade")
    assert ni == th["n_inter"] == th["n_vacancy"]
    assert _conserved(a) == ni
    # the recoil adds 5000 eV / 16M = 3.1e-4 eV per atom; the step keeps 1e-3 of it
    assert abs(e1 - e0) < 1e-3 * 5000.0 / (2 * 200 ** 3)
    # FP64 forces of the same state: momentum, and the mixed forces' error
    aos, _ = a.download()
    fm = M.sub_box_view(aos, a)["f"].copy()
    live = M.sub_box_view(aos, a)["type"] >= 0
    del aos
    b.set_precision(M.PREC_FP64)
    b.compute_forces()
    th = b.thermo()
    say("C3 FP64 sum f", th["force_sum"])
    assert np.all(np.abs(th["force_sum"]) < 1e-6)
    aos, _ = b.download()
    f64 = M.sub_box_view(aos, b)["f"]
    mx, rms = _force_err(fm[live], f64[live])
    say("C3 developed cascade mixed vs FP64 forces (lattice atoms): max %.3e rms %.3e" % (mx, rms))
    assert rms <= 1e-5
    for d in devs:
        d.close()


# --------------------------------------------------------------------------
# 64 boxes of 10^3..60^3 cells
# --------------------------------------------------------------------------
def test_c4_batch_of_64_equals_single_boxes(setfl_fecuni):
    rs = np.random.RandomState(20240)
    boxes = []
    for r in range(64):
        n = int(rs.randint(10, 61))
        t = float(rs.uniform(300.0, 1200.0))
        c = int(rs.randint(0, 21))
        boxes.append((n, t, c, 20240 + r))
    say("C4: %d boxes, %d atoms" % (len(boxes), sum(2 * b[0] ** 3 for b in boxes)))
    pairs = []
    for n, t, c, seed in boxes:
        pair = []
        for _ in range(2):
            d = _device(setfl_fecuni, (n, n, n), M.PREC_MIXED_KAHAN)
            d.build_world(seed, t, (1000 - c, c, 0), seed)
            d.set_timestep(0.001)
            d.compute_forces()
            pair.append(d)
        pairs.append(pair)
    batch = M.EamBatch([p[0] for p in pairs])
    for k, nsteps in enumerate((6, 4)):
        batch.step(nsteps)
        for p in pairs:
            p[1].step(nsteps)
        for i, p in enumerate(pairs):
            _same(p[0], p[1], "box %d after replay %d" % (i, k))
    batch.close()
    for p in pairs:
        for d in p:
            d.close()


# --------------------------------------------------------------------------
# 300^3 cells, 54M atoms, four 10 keV recoils
# --------------------------------------------------------------------------
def test_c5_four_recoils_properties(setfl_fecuni):
    n = 300
    rng = np.random.default_rng(20240)
    pkas = []
    for _ in range(4):
        lat = tuple(int(v) for v in rng.integers(0, n, size=3)) + (int(rng.integers(2)),)
        u = rng.normal(size=3)
        pkas.append((lat, tuple(float(x) for x in u / np.linalg.norm(u))))
    say("C5: 300^3, recoils", pkas)
    a = _device(setfl_fecuni, (n, n, n), M.PREC_MIXED_KAHAN, lds=M.LDS_AUTO)
    a.build_world(2020, 300.0, (985, 10, 5), 1)
    state = a.download()
    b = _device(setfl_fecuni, (n, n, n), M.PREC_MIXED_KAHAN, lds=M.LDS_OFF)
    b.upload(*state)
    del state
    for d in (a, b):  # equilibration: 100 steps of 1 fs, rescaled to 300 K
        d.set_timestep(0.001)
        d.compute_forces()
        d.step(50)
        d.rescale(300.0)
        d.step(50)
        d.rescale(300.0)
        for lat, u in pkas[:-1]:
            d.setv(lat, u, 10000.0)
        d.collision_step(pkas[-1][0], pkas[-1][1], 10000.0)
        d.set_adaptive_dt(0.001, 0.01)
    e0, _ = _e_total(a)
    for d in (a, b):
        d.step(40)
    e1, th = _e_total(a)
    say("C5: 40 adaptive steps, t %.3e ps, dE %.3e eV/atom" % (a.sim_time()[0], e1 - e0))
    assert abs(e1 - e0) < 1e-3 * 4 * 10000.0 / (2 * n ** 3)
    _same(a, b, "4 recoils, 40 adaptive steps")
    aos, _ = a.download()
    fm = M.sub_box_view(aos, a)["f"].copy()
    del aos
    b.set_precision(M.PREC_FP64)
    b.compute_forces()
    th = b.thermo()
    say("C5 FP64 sum f", th["force_sum"])
    assert np.all(np.abs(th["force_sum"]) < 1e-6)
    aos, _ = b.download()
    f64 = M.sub_box_view(aos, b)["f"]
    mx, rms = _force_err(fm, f64)
    say("C5 mixed vs FP64 forces: max %.3e rms %.3e (rms |f| %.3e eV/A)" %
        (mx, rms, float(np.sqrt(np.mean(np.sum(f64 * f64, axis=-1))))))
    assert rms <= 1e-5
    del aos, fm
    a.close()
    b.close()
