"""EAM step throughput on one MI355X: atom-steps/s of the 16M-atom FeCuNi cascade.

Workload: BCC FeCuNi, 200^3 cells (16,000,000 atoms), Fe:Cu:Ni = 985:10:5
placed by the reference's world builder at 300 K, synthetic 3-element setfl
table (Nr = Nrho = 5000, cutoff 5.6 A).  The box is equilibrated (untimed:
steps of 1 fs with a rescale to 300 K), then a primary knock-on atom at the
box centre gets 5 keV along [1 3 5] and the cascade runs with an adaptive step
dt = min(1 fs, 0.01 A / max|v|).  A step is one complete velocity-Verlet step
of the resident state (rho/F' pass, force pass, step length, half kicks and
drift, ownership decisions with inter atoms, periodic halo), state resident
in HBM; the timed steps are the first steps of the cascade.

Prints one JSON line: the FP32-mixed rate is `value`; the plain FP64 rate
from the same equilibrated state, the mixed-vs-FP64 force error, the
roofline of the dominant kernel (HIP events on the library's stream over the
timed steps) and the single-core CPU oracle rate on a bounded sample sit
beside it.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


PKA_DIR = (1.0, 3.0, 5.0)  # the recoil direction of the cascade configuration


def equilibrate(dev, steps):
    """Untimed: fixed 1 fs steps with a rescale to 300 K every 50 steps."""
    dev.set_precision(M.PREC_MIXED)
    dev.set_adaptive_dt(0.001, 0.0)
    dev.compute_forces()
    for _ in range(max(1, steps // 50)):
        dev.step(50)
        dev.rescale(300.0)
    return dev.download()


def run_precision(dev, state, mode, steps, warmup, pka_energy, pka_site):
    """From the equilibrated state: the recoil, warm-up steps, then `steps`
    timed steps; HIP events around every launch on the library's stream split
    the timed region by kernel."""
    dev.upload(*state)
    dev.set_precision(mode)
    dev.set_adaptive_dt(0.001, 0.0)
    dev.compute_forces()
    if pka_energy > 0:
        dev.collision_step(pka_site, PKA_DIR, pka_energy)
        dev.set_adaptive_dt(0.001, 0.01)
    if warmup:
        dev.step(warmup)
    dev.set_profiling(True)
    t0 = time.perf_counter()
    dev.step(steps)  # returns after the device has finished
    secs = time.perf_counter() - t0
    tim = dev.last_timings()
    dev.set_profiling(False)
    t_sim, dt_last = dev.sim_time()
    th = dev.thermo()
    return secs, tim, {"t_sim_ps": t_sim, "dt_last_ps": dt_last, "n_inter": th["n_inter"],
                       "n_vacancy": th["n_vacancy"], "max_speed": th["max_speed"]}


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
    dev.set_precision(M.PREC_MIXED)
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


def cpu_baseline(setfl):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    exe = oracle_py.BASELINE_BIN
    if not os.path.exists(exe):
        return None
    cells, steps = 50, 20
    out = subprocess.run([exe, setfl, str(cells), str(cells), str(cells), "985", "10", "5", "300", "2020",
                          str(steps)], check=True, capture_output=True, text=True, timeout=600)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    return {"value": r["atom_steps_per_s"], "unit": "atom-steps/s", "cores": 1, "kind": "port",
            "sample": "oracle (reference single-rank step restated: half lists, 104-B AoS, reference phase order) "
                      "on FeCuNi 985:10:5 BCC %d^3 cells (%d atoms), 300 K, %d steps, one thread; %s" % (
                          cells, 2 * cells ** 3, steps, "phase seconds " + json.dumps(r["phase_s"]))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cells", type=int, default=200)
    ap.add_argument("--equil", type=int, default=100, help="untimed equilibration steps")
    ap.add_argument("--pka-energy", type=float, default=5000.0, help="eV; 0 times equilibrium steps of 1 fs")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs the EAM step on one GPU only; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    n = args.cells
    natoms = 2 * n ** 3
    with tempfile.TemporaryDirectory() as tmp:
        pot = os.path.join(tmp, "fecuni.eam.alloy")
        setfl_gen.write_setfl(pot, ("Fe", "Cu", "Ni"))
        log("building %d^3 cells (%d atoms)" % (n, natoms))
        dev = M.EamDevice((n, n, n))
        dev.make_offsets()
        dev.set_potential(pot)
        dev.build_world(2020, 300.0, (985, 10, 5), 1)
        dev.set_timestep(0.001)
        log("equilibration")
        state = equilibrate(dev, args.equil)
        site = tuple(n // 2 for _ in range(3)) + (0,)
        log("fp64 run")
        s64, t64, c64 = run_precision(dev, state, M.PREC_FP64, args.steps, args.warmup, args.pka_energy, site)
        log("mixed run")
        s32, t32, c32 = run_precision(dev, state, M.PREC_MIXED, args.steps, args.warmup, args.pka_energy, site)
        del state
        log("force error")
        fmax, frms, f_scale = force_error(dev)
        th = dev.thermo()
        dev.close()
        cpu = None if args.no_cpu else cpu_baseline(pot)

    rate32 = natoms * args.steps / s32
    rate64 = natoms * args.steps / s64
    avg_force_ms = t32["force_ms"] / max(1, t32["launches"][1])
    avg_rho_ms = t32["rho_df_ms"] / max(1, t32["launches"][0])
    achieved = natoms * FORCE_BYTES["mixed"] / (avg_force_ms * 1e-3) / 1e9
    avg_force64 = t64["force_ms"] / max(1, t64["launches"][1])
    avg_rho64 = t64["rho_df_ms"] / max(1, t64["launches"][0])

    def gbs(nbytes, ms):
        return natoms * nbytes / (ms * 1e-3) / 1e9

    kernels = {
        "mixed": {"kForceM_ms": avg_force_ms, "kForceM_GBs": achieved,
                  "kRhoDfM_ms": avg_rho_ms, "kRhoDfM_GBs": gbs(RHO_BYTES["mixed"], avg_rho_ms),
                  "other_ms_per_step": t32["other_ms"] / args.steps},
        "fp64": {"kForce_ms": avg_force64, "kForce_GBs": gbs(FORCE_BYTES["fp64"], avg_force64),
                 "kRhoDf_ms": avg_rho64, "kRhoDf_GBs": gbs(RHO_BYTES["fp64"], avg_rho64),
                 "other_ms_per_step": t64["other_ms"] / args.steps},
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
        "dtype": "f32 (mixed: f32 splines/pair terms, f64 positions/Verlet/sums)",
        "data": "synthetic",
        "config": {"workload": ("collision cascade: BCC FeCuNi %d^3 cells (%d atoms), Fe:Cu:Ni 985:10:5, "
                                "equilibrated at 300 K (%d steps of 1 fs), PKA %.0f eV at %s along [1 3 5], "
                                "adaptive dt = min(1 fs, 0.01 A / max|v|), timed steps %d after %d warm-up; "
                                "synthetic 3-element setfl Nr=Nrho=5000 rc 5.6 A")
                               % (n, natoms, args.equil, args.pka_energy, str(site), args.steps, args.warmup)
                               if args.pka_energy > 0 else
                               "BCC FeCuNi %d^3 cells (%d atoms), Fe:Cu:Ni 985:10:5, 300 K, dt 1 fs, synthetic "
                               "3-element setfl Nr=Nrho=5000 rc 5.6 A" % (n, natoms),
                   "atoms": natoms, "parallelism": "single GPU"},
        "cascade": {"mixed": c32, "fp64": c64},
        "fp64": {"value": rate64, "unit": "atom-steps/s", "ms_per_step": s64 * 1e3 / args.steps},
        "kernels": kernels,
        "mixed_over_fp64": rate32 / rate64,
        "force_err_mixed_vs_fp64": {"max": fmax, "rms": frms, "rms_force_eV_per_A": f_scale},
        "roofline": {"bound": "hbm", "kernel": "kForceM (force + Verlet pass, mixed)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic_bytes(), "algorithmic_bytes_per_launch": natoms * FORCE_BYTES["mixed"],
                     "algorithmic_bytes_per_atom": FORCE_BYTES["mixed"], "avg_launch_ms": avg_force_ms},
        "step_roofline_frac": rate32 * STEP_BYTES["mixed"] / (HBM_PEAK_GBS * 1e9),
        "thermo": {"temperature": th["temperature"], "e_potential_per_atom": th["e_potential"] / natoms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
