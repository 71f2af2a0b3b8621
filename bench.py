"""EAM step throughput on one MI355X: atom-steps/s of the 16M-atom FeCuNi cascade.

Workload: BCC FeCuNi, 200^3 cells (16,000,000 atoms), Fe:Cu:Ni = 985:10:5
placed by the reference's world builder at 300 K, synthetic 3-element setfl
table (Nr = Nrho = 5000, cutoff 5.6 A).  The box is equilibrated (untimed:
steps of 1 fs with a rescale to 300 K), then a primary knock-on atom at the
box centre gets 5 keV along [1 3 5] and the cascade runs with an adaptive step
dt = min(1 fs, 0.01 A / max|v|).  A step is one complete velocity-Verlet step
of the resident state (rho/F' pass, force pass, step length, half kicks and
drift, ownership decisions with inter atoms, periodic halo), state resident
in HBM.  The timed steps run in the developed cascade (after 0.3 ps of
untimed FP32-mixed steps: a few hundred interstitials, thousands of atoms a
step taken off their sites and put back); the first steps after the recoil
are timed as well, as a side key.

Prints one JSON line: the FP32-mixed rate is `value`; the plain FP64 rate
from the same equilibrated state, the mixed-vs-FP64 force error, the
roofline of the dominant kernel (HIP events on the library's stream over the
timed steps), the NVE energy drift of both precisions over 10,000 steps of a
2M-atom FeCuNi box, and the single-core CPU oracle rate on a bounded sample
sit beside it.

Mixed precision here is FP32 splines and pair terms with FP32 Kahan-
compensated per-atom sums; positions, velocities, the Verlet update and all
reductions stay FP64.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "misa-md_amd"))

import misa_eam as M  # noqa: E402
import setfl_gen  # noqa: E402

METRIC = "atom-steps/s, 16M-atom FeCuNi EAM: FP32-mixed vs FP64 at ≤1e-5 force err"
HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per atom per launch of the force + Verlet pass:
# read x (24), type (1), F' (8 | 4), v (24); write x (24), v (24)
FORCE_BYTES = {"fp64": 105, "mixed": 101}
RHO_BYTES = {"fp64": 33, "mixed": 29}  # read x, type; write F'
STEP_BYTES = {"fp64": 138, "mixed": 130}
# HBM traffic of one force-pass launch, from the committed rocprofv3 counter
# summary of this workload (FETCH_SIZE + WRITE_SIZE, calibrated on a copy)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "kforce_mixed_fecuni_200c_traffic.txt")


# FP32 splines/pair terms, FP32 + Kahan per-atom sums
MIXED = M.PREC_MIXED_KAHAN
MVV2E = 1.0364269e-4  # amu A^2/ps^2 -> eV (pre_define.h)


def total_energy(th, natoms):
    """Kinetic (mvv/2 converted to eV) plus potential energy, per atom."""
    return (th["kinetic"] * MVV2E + th["e_potential"]) / natoms


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PKA_DIR = (1.0, 3.0, 5.0)  # the recoil direction of the cascade configuration


def equilibrate(dev, steps):
    """Untimed: fixed 1 fs steps with a rescale to 300 K every 50 steps."""
    dev.set_precision(MIXED)
    dev.set_adaptive_dt(0.001, 0.0)
    dev.compute_forces()
    for _ in range(max(1, steps // 50)):
        dev.step(50)
        dev.rescale(300.0)
    return dev.download()


def develop_cascade(dev, state, pka_site, pka_energy, t_target):
    """Untimed: from the equilibrated state, the recoil and FP32-mixed
    adaptive steps until `t_target` ps of cascade; returns the state (with its
    inter atoms), the simulated time and the steps taken."""
    dev.upload(*state)
    dev.set_precision(MIXED)
    dev.set_adaptive_dt(0.001, 0.0)
    dev.compute_forces()
    dev.collision_step(pka_site, PKA_DIR, pka_energy)
    dev.set_adaptive_dt(0.001, 0.01)
    steps = 0
    while True:
        dev.step(250)
        steps += 250
        t, _ = dev.sim_time()
        if steps % 1000 == 0:
            th = dev.thermo()
            log("cascade: %d steps, t %.4f ps, %d interstitials" % (steps, t, th["n_inter"]))
        if t >= t_target:
            break
    return dev.download(), t, steps


def run_developed(dev, state, mode, steps, warmup):
    """From a developed-cascade state (inter atoms included): warm-up steps,
    then `steps` timed adaptive steps, HIP events around every launch."""
    dev.upload(*state)
    dev.set_precision(mode)
    dev.set_adaptive_dt(0.001, 0.01)
    dev.compute_forces()
    if warmup:
        dev.step(warmup)
    natoms = int(np.prod(dev.cells)) * 2
    th0 = dev.thermo()
    e_start = total_energy(th0, natoms)
    t0_sim, _ = dev.sim_time()
    dc0 = dev.defect_counters()
    dev.set_profiling(True)
    t0 = time.perf_counter()
    dev.step(steps)  # returns after the device has finished
    secs = time.perf_counter() - t0
    tim = dev.last_timings()
    dev.set_profiling(False)
    dc1 = dev.defect_counters()
    t_sim, dt_last = dev.sim_time()
    th = dev.thermo()
    return secs, tim, {"t_window_ps": t_sim - t0_sim, "dt_last_ps": dt_last, "n_inter_start": th0["n_inter"],
                       "n_inter": th["n_inter"], "n_vacancy": th["n_vacancy"], "max_speed": th["max_speed"],
                       # per timed step: lattice atoms taken off their sites (moved beyond 0.2a), atoms
                       # put on a vacant site, and of those the ones farther than 0.2a from it (whose
                       # neighbourhoods visit the full neighbour list for a step)
                       "staged_per_step": (dc1["staged"] - dc0["staged"]) / steps,
                       "refilled_per_step": (dc1["refilled"] - dc0["refilled"]) / steps,
                       "refilled_far_per_step": (dc1["refilled_far"] - dc0["refilled_far"]) / steps,
                       "e_total_timed_start_eV_per_atom": e_start,
                       "e_total_timed_end_eV_per_atom": total_energy(th, natoms)}


def run_precision(dev, state, mode, steps, warmup, pka_energy, pka_site, pkas=None):
    """From the equilibrated state: the recoil(s), warm-up steps, then `steps`
    timed steps; HIP events around every launch on the library's stream split
    the timed region by kernel.  `pkas` lists (site, direction) pairs that
    recoil together (setv for each, then one collision step)."""
    dev.upload(*state)
    dev.set_precision(mode)
    dev.set_adaptive_dt(0.001, 0.0)
    dev.compute_forces()
    if pka_energy > 0:
        pkas = pkas or [(pka_site, PKA_DIR)]
        for lat, d in pkas[:-1]:
            dev.setv(lat, d, pka_energy)
        dev.collision_step(pkas[-1][0], pkas[-1][1], pka_energy)
        dev.set_adaptive_dt(0.001, 0.01)
    if warmup:
        dev.step(warmup)
    natoms = int(np.prod(dev.cells)) * 2
    e_start = total_energy(dev.thermo(), natoms)
    dev.set_profiling(True)
    t0 = time.perf_counter()
    dev.step(steps)  # returns after the device has finished
    secs = time.perf_counter() - t0
    tim = dev.last_timings()
    dev.set_profiling(False)
    t_sim, dt_last = dev.sim_time()
    th = dev.thermo()
    return secs, tim, {"t_sim_ps": t_sim, "dt_last_ps": dt_last, "n_inter": th["n_inter"],
                       "n_vacancy": th["n_vacancy"], "max_speed": th["max_speed"],
                       "e_total_timed_start_eV_per_atom": e_start,
                       "e_total_timed_end_eV_per_atom": total_energy(th, natoms)}


def traffic_bytes():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def force_error(dev):
    atoms, _ = dev.download()
    dev.set_precision(M.PREC_FP64)
    dev.compute_forces()
    a64, _ = dev.download()
    dev.set_precision(MIXED)
    dev.compute_forces()
    a32, _ = dev.download()
    sel = a64["type"] >= 0
    g = dev.ghost
    sb = M.sub_box_view(np.arange(len(a64)), dev).reshape(-1)
    f64 = a64["f"][sb]
    f32 = a32["f"][sb]
    rms = np.sqrt(np.mean(np.sum(f64 * f64, axis=1)))
    norm = np.maximum(np.linalg.norm(f64, axis=1), rms)
    err = np.linalg.norm(f32 - f64, axis=1) / norm
    del atoms, sel, g
    return float(err.max()), float(np.sqrt(np.mean(err * err))), float(rms)


def energy_drift(pot, cells, steps, every=500):
    """NVE from the world builder's 300 K state (FeCuNi 985:10:5, dt 1 fs):
    total energy per atom (FP64 kinetic + FP64 potential, whatever precision
    the step runs in) every `every` steps, for FP64 and mixed steps."""
    dev = M.EamDevice((cells, cells, cells))
    dev.make_offsets()
    dev.set_potential(pot)
    dev.build_world(2020, 300.0, (985, 10, 5), 1)
    dev.set_timestep(0.001)
    state = dev.download()
    natoms = 2 * cells ** 3
    out = {"workload": "BCC FeCuNi %d^3 cells (%d atoms), Fe:Cu:Ni 985:10:5, 300 K start, NVE, dt 1 fs, %d steps, "
                       "total energy sampled every %d steps" % (cells, natoms, steps, every)}
    for name, mode in (("fp64", M.PREC_FP64), ("mixed", MIXED)):
        dev.upload(*state)
        dev.set_precision(mode)
        dev.set_adaptive_dt(0.001, 0.0)
        dev.compute_forces()
        t, e = [0.0], []
        th = dev.thermo()
        e.append(total_energy(th, natoms))
        for k in range(steps // every):
            dev.step(every)
            th = dev.thermo()
            e.append(total_energy(th, natoms))
            t.append((k + 1) * every * 0.001)
        e = np.array(e)
        slope = float(np.polyfit(np.array(t), e - e[0], 1)[0]) if len(e) > 2 else 0.0
        out[name] = {"e0_eV_per_atom": float(e[0]), "drift_eV_per_atom": float(e[-1] - e[0]),
                     "max_dev_eV_per_atom": float(np.abs(e - e[0]).max()),
                     "slope_eV_per_atom_per_ps": slope, "temperature_end": th["temperature"]}
    dev.close()
    return out


def batch_workload(pot, args):
    """64 independent boxes stepped together through captured step graphs:
    n ~ U{10..60} cells per side, T ~ U[300, 1200] K, Cu per mille ~ U{0..20}
    (no Ni), seeds base + r, dt 1 fs, FP32-mixed, each box's launch shape
    picked by the autotuner.  Returns the aggregate atom-steps/s line."""
    rs = np.random.RandomState(args.batch_seed)  # MT19937
    boxes = []
    for r in range(args.batch_boxes):
        n = int(rs.randint(10, 61))
        t = float(rs.uniform(300.0, 1200.0))
        c = int(rs.randint(0, 21))
        boxes.append((n, t, c, args.batch_seed + r))
    devs, natoms, shapes = [], 0, []
    for n, t, c, seed in boxes:
        d = M.EamDevice((n, n, n))
        d.make_offsets()
        d.set_potential(pot)
        d.build_world(seed, t, (1000 - c, c, 0), seed)
        d.set_timestep(0.001)
        d.set_precision(MIXED)
        d.compute_forces()
        d.autotune(4)
        shapes.append(list(d.launch_shape()))
        devs.append(d)
        natoms += 2 * n ** 3
    log("batch of %d boxes, %d atoms" % (len(devs), natoms))
    b = M.EamBatch(devs)
    b.step(args.warmup)  # returns after the device has finished
    t0 = time.perf_counter()
    b.step(args.steps)
    secs = time.perf_counter() - t0
    temps = [d.thermo()["temperature"] for d in devs]
    b.close()
    for d in devs:
        d.close()
    return {
        "metric": "atom-steps/s, batch of %d FeCuNi EAM boxes (FP32-mixed)" % len(boxes),
        "value": natoms * args.steps / secs, "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": secs * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (mixed, FP32+Kahan sums)", "data": "synthetic",
        "config": {"workload": "batch: %d BCC boxes, n ~ U{10..60} cells, T ~ U[300,1200] K, Cu ~ U{0..20} per "
                               "mille, dt 1 fs, master seed %d" % (len(boxes), args.batch_seed),
                   "atoms": natoms, "boxes": [list(x) for x in boxes], "launch_shapes": shapes,
                   "parallelism": "single GPU"},
        "temperatures_end": temps,
    }


def large_workload(pot, args):
    """54M atoms (300^3 cells) FeCuNi at 300 K with 4 simultaneous 10 keV
    recoils at seeded sites along seeded isotropic directions, adaptive step;
    FP32-mixed and FP64 from the same state: rates, force error and energy
    change over the run."""
    n = args.cells
    natoms = 2 * n ** 3
    dev = M.EamDevice((n, n, n))
    dev.make_offsets()
    dev.set_potential(pot)
    dev.build_world(2020, 300.0, (985, 10, 5), 1)
    dev.set_timestep(0.001)
    log("equilibration")
    state = equilibrate(dev, args.equil)
    rng = np.random.default_rng(args.batch_seed)
    pkas = []
    for _ in range(4):
        lat = tuple(int(v) for v in rng.integers(0, n, size=3)) + (int(rng.integers(2)),)
        u = rng.normal(size=3)
        pkas.append((lat, tuple(float(x) for x in u / np.linalg.norm(u))))
    out = {}
    for name, mode in (("fp64", M.PREC_FP64), ("mixed", MIXED)):
        log("%s run" % name)
        secs, tim, cas = run_precision(dev, state, mode, args.steps, args.warmup, args.pka_energy, None, pkas)
        out[name] = {"value": natoms * args.steps / secs, "ms_per_step": secs * 1e3 / args.steps,
                     "force_ms": tim["force_ms"] / max(1, tim["launches"][1]),
                     "rho_ms": tim["rho_df_ms"] / max(1, tim["launches"][0]), "cascade": cas}
    del state
    log("force error")
    fmax, frms, f_scale = force_error(dev)
    dev.close()
    return {
        "metric": "atom-steps/s, 54M-atom FeCuNi EAM with 4 x 10 keV recoils (FP32-mixed)",
        "value": out["mixed"]["value"], "unit": "atom-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": out["mixed"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 (mixed, FP32+Kahan sums)", "data": "synthetic",
        "config": {"workload": "BCC FeCuNi %d^3 cells (%d atoms), 985:10:5, 300 K (%d equilibration steps), "
                               "4 PKAs of %.0f eV at %s, adaptive dt = min(1 fs, 0.01 A / max|v|)"
                               % (n, natoms, args.equil, args.pka_energy, str(pkas)),
                   "atoms": natoms, "parallelism": "single GPU"},
        "fp64": out["fp64"], "mixed": out["mixed"], "mixed_over_fp64": out["mixed"]["value"] / out["fp64"]["value"],
        "force_err_mixed_vs_fp64": {"max": fmax, "rms": frms, "rms_force_eV_per_A": f_scale},
    }


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count(), len(os.sched_getaffinity(0))


def cpu_baseline(setfl, setfl_fe, cells):
    """The oracle (the reference's single-rank step restated: half lists,
    104-B AoS, reference phase order) on one host thread: the timed workload
    (FeCuNi 985:10:5 at `cells`^3, 300 K) for one step after its first force
    evaluation, beside the 20^3 Fe box (600 K, 100 steps) and the 100^3
    FeCuNi box (3 steps).  `value` is the workload's rate."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    exe = oracle_py.BASELINE_BIN
    if not os.path.exists(exe):
        return None

    def run(pot, n, ratio, t, seed, steps):
        out = subprocess.run([exe, pot, str(n), str(n), str(n)] + [str(r) for r in ratio] +
                             [str(t), str(seed), str(steps)], check=True, capture_output=True, text=True,
                             timeout=900)
        return json.loads(out.stdout.strip().splitlines()[-1])

    samples = {}
    for name, pot, n, ratio, t, seed, steps in (
            ("fe_20c", setfl_fe, 20, (1, 0, 0), 600.0, 466953, 100),
            ("fecuni_100c", setfl, 100, (985, 10, 5), 300.0, 2020, 3),
            ("fecuni_%dc" % cells, setfl, cells, (985, 10, 5), 300.0, 2020, 1)):
        log("cpu baseline: %s, %d steps" % (name, steps))
        r = run(pot, n, ratio, t, seed, steps)
        samples[name] = {"atoms": int(r["atoms"]), "steps": steps, "seconds": r["seconds"],
                         "atom_steps_per_s": r["atom_steps_per_s"], "phase_s": r["phase_s"]}
    model, ncpu, avail = host_cpu()
    main = samples["fecuni_%dc" % cells]
    return {"value": main["atom_steps_per_s"], "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle on one host thread (g++ -O3, x86-64-v3): BCC FeCuNi 985:10:5 %d^3 cells (%d atoms), "
                      "300 K, 1 equilibrium step timed after the first force evaluation; beside it 20^3 Fe 600 K "
                      "100 steps and 100^3 FeCuNi 3 steps" % (cells, main["atoms"]),
            "host_cpu": model, "host_nproc": ncpu, "host_cpus_available": avail, "samples": samples}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--early-steps", type=int, default=20, help="timed steps right after the recoil (side key)")
    ap.add_argument("--develop-ps", type=float, default=0.3, help="untimed cascade time before the timed steps")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=200)
    ap.add_argument("--equil", type=int, default=100, help="untimed equilibration steps")
    ap.add_argument("--pka-energy", type=float, default=5000.0, help="eV; 0 times equilibrium steps of 1 fs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--drift-steps", type=int, default=10000, help="NVE steps of the energy-drift leg; 0 skips it")
    ap.add_argument("--drift-cells", type=int, default=100)
    ap.add_argument("--workload", choices=("cascade", "batch", "large"), default="cascade",
                    help="cascade: the 16M-atom cascade (default); batch: 64 boxes through captured graphs; "
                         "large: 54M atoms with 4 recoils (use --cells 300 --pka-energy 10000)")
    ap.add_argument("--batch-boxes", type=int, default=64)
    ap.add_argument("--batch-seed", type=int, default=20240)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the EAM step on one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.workload != "cascade":
        with tempfile.TemporaryDirectory() as tmp:
            pot = os.path.join(tmp, "fecuni.eam.alloy")
            setfl_gen.write_setfl(pot, ("Fe", "Cu", "Ni"))
            res = batch_workload(pot, args) if args.workload == "batch" else large_workload(pot, args)
        print(json.dumps(res))
        return
    n = args.cells
    natoms = 2 * n ** 3
    with tempfile.TemporaryDirectory() as tmp:
        pot = os.path.join(tmp, "fecuni.eam.alloy")
        setfl_gen.write_setfl(pot, ("Fe", "Cu", "Ni"))
        pot_fe = os.path.join(tmp, "fe.eam.alloy")
        setfl_gen.write_setfl(pot_fe, ("Fe",))
        log("building %d^3 cells (%d atoms)" % (n, natoms))
        dev = M.EamDevice((n, n, n))
        dev.make_offsets()
        dev.set_potential(pot)
        dev.build_world(2020, 300.0, (985, 10, 5), 1)
        dev.set_timestep(0.001)
        log("equilibration")
        state = equilibrate(dev, args.equil)
        site = tuple(n // 2 for _ in range(3)) + (0,)
        early = {}
        for name, mode in (("fp64", M.PREC_FP64), ("mixed", MIXED)):
            log("%s run, first steps after the recoil" % name)
            secs, tim, cas = run_precision(dev, state, mode, args.early_steps, args.warmup, args.pka_energy, site)
            early[name] = {"value": natoms * args.early_steps / secs, "ms_per_step": secs * 1e3 / args.early_steps,
                           "steps": args.early_steps, "cascade": cas}
        log("developing the cascade to %.2f ps (FP32-mixed, untimed)" % args.develop_ps)
        dstate, t_dev, dev_steps = develop_cascade(dev, state, site, args.pka_energy, args.develop_ps)
        del state
        runs, launched = {}, {}
        for name, mode in (("fp64", M.PREC_FP64), ("mixed", MIXED)):
            log("%s run, developed cascade" % name)
            runs[name] = run_developed(dev, dstate, mode, args.steps, args.warmup)
            launched[name] = dev.force_pass()
        del dstate
        log("force error")
        fmax, frms, f_scale = force_error(dev)
        th = dev.thermo()
        dev.close()
        drift = None
        if args.drift_steps > 0:
            log("energy drift")
            drift = energy_drift(pot, args.drift_cells, args.drift_steps)
        cpu = None if args.no_cpu else cpu_baseline(pot, pot_fe, n)

    s32, t32, c32 = runs["mixed"]
    s64, t64, c64 = runs["fp64"]
    rate32 = natoms * args.steps / s32
    rate64 = natoms * args.steps / s64
    avg_force_ms = t32["force_ms"] / max(1, t32["launches"][1])
    avg_rho_ms = t32["rho_df_ms"] / max(1, t32["launches"][0])
    achieved = natoms * FORCE_BYTES["mixed"] / (avg_force_ms * 1e-3) / 1e9
    avg_force64 = t64["force_ms"] / max(1, t64["launches"][1])
    avg_rho64 = t64["rho_df_ms"] / max(1, t64["launches"][0])

    def gbs(nbytes, ms):
        return natoms * nbytes / (ms * 1e-3) / 1e9

    def shares(t, secs):
        total = secs * 1e3
        return {"rho_pass": t["rho_df_ms"] / total, "force_pass": t["force_ms"] / total,
                "inter_and_ownership": t["other_ms"] / total}

    kernels = {
        "mixed": {"force_kernel": launched["mixed"], "force_ms": avg_force_ms, "force_GBs": achieved,
                  "rho_kernel": "kRhoDfM", "rho_ms": avg_rho_ms, "rho_GBs": gbs(RHO_BYTES["mixed"], avg_rho_ms),
                  "other_ms_per_step": t32["other_ms"] / args.steps, "share_of_step": shares(t32, s32)},
        "fp64": {"force_kernel": launched["fp64"], "force_ms": avg_force64,
                 "force_GBs": gbs(FORCE_BYTES["fp64"], avg_force64),
                 "rho_kernel": "kRhoDf", "rho_ms": avg_rho64, "rho_GBs": gbs(RHO_BYTES["fp64"], avg_rho64),
                 "other_ms_per_step": t64["other_ms"] / args.steps, "share_of_step": shares(t64, s64)},
    }
    res = {
        "metric": METRIC,
        "value": rate32,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s32 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (mixed: f32 splines/pair terms, f32+Kahan per-atom sums, f64 positions/Verlet/reductions)",
        "data": "synthetic",
        "config": {"workload": ("collision cascade: BCC FeCuNi %d^3 cells (%d atoms), Fe:Cu:Ni 985:10:5, "
                                "equilibrated at 300 K (%d steps of 1 fs), PKA %.0f eV at %s along [1 3 5], "
                                "adaptive dt = min(1 fs, 0.01 A / max|v|), developed for %.3f ps (%d steps, "
                                "untimed), then %d timed steps after %d warm-up; synthetic 3-element setfl "
                                "Nr=Nrho=5000 rc 5.6 A")
                               % (n, natoms, args.equil, args.pka_energy, str(site), t_dev, dev_steps, args.steps,
                                  args.warmup),
                   "atoms": natoms, "parallelism": "single GPU"},
        "cascade": {"developed_at_ps": t_dev, "mixed": c32, "fp64": c64},
        "fp64": {"value": rate64, "unit": "atom-steps/s", "ms_per_step": s64 * 1e3 / args.steps},
        "early_window": early,
        "kernels": kernels,
        "mixed_over_fp64": rate32 / rate64,
        "force_err_mixed_vs_fp64": {"max": fmax, "rms": frms, "rms_force_eV_per_A": f_scale,
                                    "within_1e-5": frms <= 1e-5},
        "energy_drift": drift,
        "roofline": {"bound": "hbm", "kernel": launched["mixed"] + " (force + Verlet pass, mixed)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_bytes(), "traffic_source": "committed rocprofv3 counter summary " +
                     os.path.relpath(TRAFFIC_FILE, ROOT) + " (FETCH_SIZE + WRITE_SIZE, calibrated on a copy)",
                     "algorithmic_bytes_per_launch": natoms * FORCE_BYTES["mixed"],
                     "algorithmic_bytes_per_atom": FORCE_BYTES["mixed"], "avg_launch_ms": avg_force_ms},
        "step_roofline_frac": rate32 * STEP_BYTES["mixed"] / (HBM_PEAK_GBS * 1e9),
        "thermo": {"temperature": th["temperature"], "e_potential_per_atom": th["e_potential"] / natoms},
        "cpu_baseline": cpu,
    }
    if frms > 1e-5:
        res["failed"] = "mixed-vs-FP64 RMS force error %.3e exceeds 1e-5" % frms
    print(json.dumps(res))


if __name__ == "__main__":
    main()
