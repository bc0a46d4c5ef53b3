"""Precision study of the LJ time step on one GPU.

Workload (defaults): N = 131072, rho* = 0.3, T* = 1.0, hard-wall box, EVN,
dt = 0.002, 10000 steps, seeded lattice + Gaussian initial state.  The same
initial state is run in the three precision modes of the library:

  fp32          FP32 pair math, FP32 per-slice sums, FP64 cross-slice sums
  compensated   as fp32 with TwoSum-compensated per-particle sums
  fp64          double state, double pair math, the reference integrator's
                formulas in double (the reference of this study)

Reported: force / E_pot / virial error of the first state against fp64; the
deviation of E_pot and P from the fp64 run over the first 100 steps; the
total-energy drift per 1000 steps (least-squares slope of U/N) and the largest
excursion |U - U_0|/N of every mode over the whole run; time per step.

    python -m ljamd.precision [--N 131072] [--steps 10000] [--out result.json]
"""
import argparse
import ctypes as C
import json
import sys
import time

import numpy as np

from . import _lib as L
from .system import LJSystem

MODES = (("fp64", L.PREC_FP64), ("fp32", L.PREC_FP32), ("compensated", L.PREC_COMPENSATED))


def _forces(s):
    return s.get_array_f64(L.ARRAY_FORCE).reshape(-1, 4)[:, :3]


def run_mode(name, prec, a, log):
    with LJSystem(a.N, a.rho, a.T0, canonical=L.EVN, boundary=L.BC_WALL,
                  precision=prec) as s:
        s.sample_initial_state(a.seed)
        obs0 = s.observables()
        out = {"F0": _forces(s), "V0": obs0["V"], "W0": obs0["W"], "U0": obs0["U"]}
        hist = np.zeros((a.steps, 5))
        done, t_run = 0, 0.0
        n100 = min(100, a.steps)
        while done < a.steps:
            k = n100 if done == 0 else min(a.chunk, a.steps - done)
            t = time.perf_counter()
            hist[done:done + k] = s.Integrate(a.dt, k)
            t_run += time.perf_counter() - t
            done += k
            if done == n100:
                out["x100"] = s.get_array_f64(L.ARRAY_POS).reshape(-1, 4)[:, :3].copy()
            log(f"[{name}] step {done}/{a.steps}  U/N {hist[done - 1, 0] / a.N:+.9f}  "
                f"{1e3 * t_run / done:.3f} ms/step")
        us = C.c_double()
        L.check(L.lib().lj_time_force_kernel(s.h, 5, C.byref(us)))
    out.update(hist=hist, ms_per_step=1e3 * t_run / a.steps, force_us=us.value)
    return out


def drift(U, N, dt_steps=1000):
    k = np.arange(U.size, dtype=np.float64)
    slope = np.polyfit(k, U / N, 1)[0]
    return {"drift_per_1000_steps": float(slope * dt_steps),
            "max_excursion": float(np.max(np.abs(U - U[0])) / N)}


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("--N", type=int, default=131072)
    p.add_argument("--rho", type=float, default=0.3)
    p.add_argument("--T0", type=float, default=1.0)
    p.add_argument("--dt", type=float, default=0.002)
    p.add_argument("--steps", type=int, default=10000)
    p.add_argument("--chunk", type=int, default=1000)
    p.add_argument("--seed", type=int, default=5)
    p.add_argument("--out", default=None)
    a = p.parse_args(argv)
    a.chunk = max(1, min(a.chunk, a.steps))

    def log(msg):
        print(msg, flush=True)

    log(f"workload: N={a.N} rho={a.rho} T0={a.T0} hard wall EVN dt={a.dt} steps={a.steps} "
        f"seed={a.seed}")
    runs = {name: run_mode(name, prec, a, log) for name, prec in MODES}
    ref = runs["fp64"]
    n100 = min(100, a.steps)
    res = {"workload": {"N": a.N, "rho": a.rho, "T0": a.T0, "boundary": "hard wall",
                        "ensemble": "EVN", "dt": a.dt, "steps": a.steps, "seed": a.seed},
           "modes": {}}
    for name, r in runs.items():
        d = {"ms_per_step": r["ms_per_step"], "force_kernel_us": r["force_us"]}
        d.update(drift(r["hist"][:, 0], a.N))
        if name != "fp64":
            dF = r["F0"] - ref["F0"]
            n = np.linalg.norm(ref["F0"], axis=1)
            floor = np.sqrt(np.mean(n ** 2))
            d["state0"] = {
                "force_global_rel": float(np.linalg.norm(dF) / np.linalg.norm(ref["F0"])),
                "force_max_rel_floor": float(np.max(np.linalg.norm(dF, axis=1)
                                                    / np.maximum(n, floor))),
                "E_pot_rel": float(abs(r["V0"] - ref["V0"]) / abs(ref["V0"])),
                "W_rel": float(abs(r["W0"] - ref["W0"]) / abs(ref["W0"])),
            }
            h, h64 = r["hist"][:n100], ref["hist"][:n100]
            d["first_100_steps"] = {
                "E_pot_max_rel": float(np.max(np.abs(h[:, 3] - h64[:, 3]) / np.abs(h64[:, 3]))),
                "P_max_rel": float(np.max(np.abs(h[:, 4] - h64[:, 4]) / np.abs(h64[:, 4]))),
                "U_max_abs_per_N": float(np.max(np.abs(h[:, 0] - h64[:, 0])) / a.N),
                "x_max_abs_at_100": float(np.max(np.abs(r["x100"] - ref["x100"]))),
            }
        res["modes"][name] = d
    line = json.dumps(res, indent=1)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
