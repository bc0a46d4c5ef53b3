"""Benchmark of the MI355X Lennard-Jones time step.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload (the configuration the headline metric is quoted on): N = 16384,
rho* = 0.85, T* = 1.0, periodic box, TVN (kinetic-temperature rescaling every
step), dt* = 0.005, FP32, simple-cubic lattice + seeded Gaussian velocities.
A step is one full device-resident MDSystem::Integrate (drift, all-pairs
forces/energy/virial, TVN rescale, boundary conditions, observables),
replayed from a captured HIP graph.  value = N^2 * K / t pair interactions/s.

roofline: the force kernel alone, timed with HIP events on the stream it is
launched on; algorithmic work 20 FLOP per ordered pair, N^2 pairs per launch,
against the FP32 vector peak.  At this size the kernel evaluates each unordered
pair once (Newton's third law, N (N - 1) / 2 evaluations); it is still
credited N^2 ordered pairs, and the sliced kernel that evaluates both orders
is timed beside it.
cpu_baseline: the oracle's C restatement of the reference CPU path
(oracle/lj_oracle.c, bit-identical to the reference's MDSystem.cpp compiled
from source: tests/test_oracle_pins.py::test_oracle_bitexact_vs_reference and
the committed reference-run fixtures, tests/test_oracle_golden.py) running the
same workload for a bounded number of steps on one host thread, as the
reference does; the same code over 16 host threads beside it.
precision: the FP32, compensated and FP64 modes on the precision workload
(N = 131072, hard wall), force-pass time and E_pot / P deviation from FP64
over the first 100 steps.
shard: one GPU's share of the headline force pass sharded by i-particles over
8 GPUs (lj_time_forces_shard: 2048 rows against all N on the sliced kernel)
beside the whole pass on the same rule; the collectives are not run (one GPU).

Clocks: the GPU's shader clock ramps up over the first ~0.1 s of sustained
load, so every timed leg first runs force passes for at least SETTLE_MS
(untimed, bounded), reported as settle_ms.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lennard-jones-cuda_amd", "python"))

METRIC = "LJ pair interactions/s at N=16k (% FP32 peak); batched box-steps/s vs plain loop"
PEAK_FP32_TFLOPS = 157.3        # MI355X FP32 vector peak (packed FMA issue)
PEAK_FP32_SCALAR_TFLOPS = 78.65  # the same without packed FP32 issue
FLOP_PER_PAIR = 20
N, RHO, T0, DT, SEED = 16384, 0.85, 1.0, 0.005, 2024
SETTLE_MS = 250.0       # untimed force passes before each timed leg
SETTLE_MAX_MS = 3000.0  # bound of the settle phase
KERNEL_REPS = 200       # force launches averaged for roofline.kernel_us
# SQ_INSTS_VALU x 64 / N^2 of the pair-once step pass (12.35, rocprofv3) plus
# the half-rate v_rcp_f32, one per unordered pair
ISSUE_SLOTS_N3 = 12.85


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def profiled_traffic(kernel):
    """HBM bytes per force launch from the committed rocprofv3 summary."""
    import re

    path = os.path.join(ROOT, "profiles", f"{kernel}_n16384_periodic.md")
    try:
        m = re.search(r"traffic_bytes_per_launch:\s*(\d+)", open(path).read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def settle(s, min_ms=SETTLE_MS, max_ms=SETTLE_MAX_MS):
    """Untimed force passes on context s until min_ms of wall time have passed
    (at most max_ms): the shader clock settles under sustained load. Returns
    the milliseconds spent."""
    t0 = time.perf_counter()
    while True:
        s.time_force_kernel(100)
        ms = 1000.0 * (time.perf_counter() - t0)
        if ms >= min_ms or ms >= max_ms:
            return ms


def gpu_leg(steps, warmup, boundary=None):
    import ljamd

    bc = ljamd.BC_PERIODIC if boundary is None else boundary
    s = ljamd.LJSystem(N, RHO, T0, canonical=ljamd.TVN, boundary=bc, seed=SEED)
    settle_ms = settle(s)
    log(f"[bench] clocks settled ({settle_ms:.0f} ms of force passes); warmup {warmup} steps")
    s.Integrate(DT, max(warmup, 1))
    s.synchronize()
    log(f"[bench] timing {steps} steps")
    t0 = time.perf_counter()
    hist = s.Integrate(DT, steps)
    s.synchronize()
    t = time.perf_counter() - t0
    obs = s.observables()
    reps = KERNEL_REPS
    us, span = s.time_force_kernel_stamped(reps)
    pair_once, nslots = s.force_plan()
    s.close()
    return {"t": t, "hist": hist, "obs": obs, "force_us": us, "force_span_us": span, "reps": reps,
            "pair_once": pair_once, "nslots": nslots, "settle_ms": settle_ms}


def sliced_kernel_us(reps=KERNEL_REPS):
    """The sliced kernel (both orders of every pair) on the same workload."""
    import ljamd

    with ljamd.LJSystem(N, RHO, T0, canonical=ljamd.TVN, seed=SEED, jslices=32) as s:
        settle(s)
        return s.time_force_kernel_stamped(reps)[0]


def batch_leg(steps, plain_steps=40, nboxes=512):
    """512 independent periodic EVN boxes (N ~ U{200..4000}, rho* ~ U[0.05,0.9],
    T* ~ U[0.8,2.0], seeded) stepped by one captured batched graph, against a
    loop that launches each box's own single-box step graph."""
    import numpy as np

    import ljamd

    N, rho, T0 = ljamd.ensemble_boxes(nboxes, seed=7)
    pairs = float(np.sum(N.astype(np.float64) ** 2))
    b = ljamd.LJBatch(N, rho, T0, canonical=ljamd.EVN, seed=SEED)
    log("[bench] batch autotune")
    tuning, tune_us = b.autotune(reps=3)
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < SETTLE_MS / 1000.0:  # clocks settle
        b.time_force(20)
    b.integrate(DT, 3)
    log(f"[bench] batch timing {steps} steps (tuning {tuning})")
    t0 = time.perf_counter()
    b.integrate(DT, steps)
    tb = time.perf_counter() - t0
    force_us = b.time_force(5)
    b.close()
    log(f"[bench] plain loop over {nboxes} single-box contexts, {plain_steps} steps")
    ctxs = [ljamd.LJSystem(int(N[k]), float(rho[k]), float(T0[k]), canonical=ljamd.EVN,
                           seed=SEED + k) for k in range(nboxes)]
    for c in ctxs:  # builds each box's step graph
        c.integrate_async(DT, 1)
    for c in ctxs:
        c.synchronize()
    t0 = time.perf_counter()
    for _ in range(plain_steps):
        for c in ctxs:
            c.integrate_async(DT, 1)
    for c in ctxs:
        c.synchronize()
    tp = time.perf_counter() - t0
    for c in ctxs:
        c.close()
    bsps = nboxes * steps / tb
    psps = nboxes * plain_steps / tp
    return {"boxes": nboxes, "steps": steps, "box_steps_per_s": bsps,
            "plain_loop_steps": plain_steps, "plain_loop_box_steps_per_s": psps,
            "speedup_vs_plain_loop": bsps / psps, "pairs_per_step": pairs,
            "pair_interactions_per_s": pairs * steps / tb,
            "tuning": {"ipl": tuning[0], "jchunk": tuning[1], "blocks_per_cu": tuning[2],
                       "pair_once": tuning[3]},
            "force_pass_us": force_us, "ms_per_step": 1000.0 * tb / steps}


def adjoint_leg():
    """Differentiable workload: N=2048, rho*=0.85, T*=1, periodic, 200 EVN steps,
    dt=0.005, checkpoints every 20; gradient of <E_pot> + (<P> - 1)^2 with
    respect to eps, sigma and x0.  Cost relative to the forward trajectory."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import ljamd

    Na, rho, M, dt = 2048, 0.85, 200, 0.005
    with ljamd.LJSystem(Na, rho, 1.0, seed=4) as s:
        settle(s)
        s.trajectory_grad(dt, M, 20)  # warm: graph capture, scratch
        out, g2, _, t20 = s.trajectory_grad(dt, M, 20)
        _, _, _, t1 = s.trajectory_grad(dt, M, 1)
    return {"N": Na, "steps": M, "checkpoint_every": 20, "loss": out[0],
            "dL_deps": g2[0], "dL_dsigma": g2[1], "forward_ms": t20[0],
            "recompute_ms": t20[1], "reverse_ms": t20[2],
            "gradient_over_forward_k20": (t20[1] + t20[2]) / t20[0],
            "gradient_over_forward_k1": t1[2] / t1[0]}


def cpu_leg(budget_s=15.0):
    """The reference CPU path on the same workload, as the oracle restates it
    (oracle/lj_oracle.c: float state, double pair terms, float force sums over
    j ascending, MDSystem.cpp:232-311,438-510), one host thread as the
    reference runs. The restatement is bit-identical to the reference compiled
    from its sources (tests/test_oracle_pins.py::test_oracle_bitexact_vs_reference)
    and to the committed reference-run fixtures (tests/test_oracle_golden.py)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleSystem, lattice_state

    L, pos, vel = lattice_state(N, RHO, T0, SEED)
    o = OracleSystem(N, T0, RHO, 1, 0, nthreads=1)
    o.set_state(pos, vel)
    o.forces()
    o.params()
    t0 = time.perf_counter()
    o.integrate(DT, 1)
    t1 = time.perf_counter() - t0
    k = max(1, min(10, int(budget_s / max(t1, 1e-3)) - 1))
    t0 = time.perf_counter()
    o.integrate(DT, k)
    tk = time.perf_counter() - t0
    steps = k + 1
    tt = t1 + tk
    return {"value": N * N * steps / tt, "unit": "pair interactions/s", "cores": 1,
            "kind": "port",
            "sample": f"{steps} TVN steps of the oracle's restatement of MDSystem::Integrate "
                      f"(the reference CPU path, bit-identical to it) at N={N}, rho*={RHO}, "
                      f"1 host thread, {tt:.1f} s",
            "pin": "tests/test_oracle_pins.py::test_oracle_bitexact_vs_reference, "
                   "tests/test_oracle_golden.py",
            "steps_per_s": steps / tt}


def rdf_leg(reps=20):
    """The RDF pass (integer work, reported apart from the FLOP count) at the
    headline size: 500 bins uniform in r to L/2 (the workload's sampling) and
    the reference's 256 bins in r^2 with its float binning."""
    import ljamd

    with ljamd.LJSystem(N, RHO, T0, canonical=ljamd.TVN, seed=SEED) as s:
        settle(s)
        s.Integrate(DT, 20)
        s.synchronize()
        us_r = s.time_rdf(ljamd.RDF_RBINS, 500, reps=reps)
        us_g = s.time_rdf(ljamd.RDF_REF_GPU, 256, reps=reps)
        us_c = s.time_rdf(ljamd.RDF_REF_CPU, 256, reps=reps)
    # the same with hard walls (the workload's second run)
    with ljamd.LJSystem(N, RHO, T0, canonical=ljamd.TVN, boundary=ljamd.BC_WALL, seed=SEED) as s:
        s.Integrate(DT, 20)
        s.synchronize()
        wr = s.time_rdf(ljamd.RDF_RBINS, 500, reps=reps)
        wg = s.time_rdf(ljamd.RDF_REF_GPU, 256, reps=reps)
        wc = s.time_rdf(ljamd.RDF_REF_CPU, 256, reps=reps)
    pairs = float(N) * N
    return {"rbins_500_to_half_box_us": us_r, "rbins_pairs_per_s": pairs / (us_r * 1e-6),
            "ref_gpu_256_r2_bins_us": us_g, "ref_gpu_pairs_per_s": pairs / (us_g * 1e-6),
            "ref_cpu_256_r2_bins_us": us_c, "ref_cpu_pairs_per_s": pairs / (us_c * 1e-6),
            "hard_wall": {"rbins_500_to_half_box_us": wr, "ref_gpu_256_r2_bins_us": wg,
                          "ref_cpu_256_r2_bins_us": wc}}


def shard_leg(nshards=8, reps=200):
    """One share of the headline force pass sharded by i over nshards GPUs
    against the whole pass with the same slicing rule (the sliced kernel)."""
    import ljamd

    with ljamd.LJSystem(N, RHO, T0, canonical=ljamd.TVN, seed=SEED) as s:
        settle(s)
        share = s.time_forces_shard(0, nshards, reps=reps)
        whole = s.time_forces_shard(0, 1, reps=max(1, reps // 4))
        # the pair-once step pass with its items dealt out instead
        n3_share = s.time_force_items_share(0, nshards, reps=reps)
        n3_whole = s.time_force_items_share(0, 1, reps=max(1, reps // 4))
    return {"nshards": nshards, "rows": N // nshards, "share_force_us": share,
            "whole_force_us": whole, "speedup": whole / share,
            "kernel": "k_force (sliced) on a row range; rows of whole 64-particle tiles",
            "pair_once_items": {"share_force_us": n3_share, "whole_force_us": n3_whole,
                                "speedup": n3_whole / n3_share,
                                "kernel": "k_force_n3 step variant on 1/nshards of its items"}}


def precision_leg(steps=100):
    """The precision modes on the precision workload at full size (N = 131072,
    rho* = 0.3, T* = 1, hard wall, EVN, dt = 0.002, one seeded state): the
    force pass of each mode (HIP events) and, against the FP64 mode, the
    largest relative deviation of E_pot and P over the first `steps` steps."""
    import types

    import numpy as np
    from ljamd import precision as pr

    a = types.SimpleNamespace(N=131072, rho=0.3, T0=1.0, dt=0.002, steps=steps, chunk=steps,
                              seed=5)
    runs = {name: pr.run_mode(name, prec, a, lambda msg: None) for name, prec in pr.MODES}
    ref = runs["fp64"]["hist"]
    out = {"workload": "N=131072, rho*=0.3, T*=1.0, hard wall, EVN, dt*=0.002, seed 5",
           "steps": steps}
    for name, r in runs.items():
        d = {"force_kernel_us": r["force_us"], "ms_per_step": r["ms_per_step"]}
        if name != "fp64":
            h = r["hist"]
            d["E_pot_max_rel_vs_fp64"] = float(np.max(np.abs(h[:, 3] - ref[:, 3]) / np.abs(ref[:, 3])))
            d["P_max_rel_vs_fp64"] = float(np.max(np.abs(h[:, 4] - ref[:, 4]) / np.abs(ref[:, 4])))
        out[name] = d
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_parallel_leg(threads=16, budget_s=5.0):
    """The reference CPU path restated in the oracle (bit-identical to it: each
    particle's j order is unchanged) with the i loop split over `threads` host
    threads, on the same workload; a reported baseline, not a target."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import OracleSystem, lattice_state

    L, pos, vel = lattice_state(N, RHO, T0, SEED)
    o = OracleSystem(N, T0, RHO, 1, 0, nthreads=threads)
    o.set_state(pos, vel)
    o.forces()
    o.params()
    t0 = time.perf_counter()
    o.integrate(DT, 1)
    t1 = time.perf_counter() - t0
    k = max(1, min(40, int(budget_s / max(t1, 1e-3)) - 1))
    t0 = time.perf_counter()
    o.integrate(DT, k)
    tk = time.perf_counter() - t0
    steps, tt = k + 1, t1 + tk
    return {"value": N * N * steps / tt, "unit": "pair interactions/s", "cores": threads,
            "kind": "port", "host_cpus": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"{steps} TVN steps of the oracle's restatement of MDSystem::Integrate "
                      f"at N={N}, rho*={RHO}, {threads} host threads, {tt:.1f} s",
            "steps_per_s": steps / tt}


def profiled_sclk_ghz(kernel):
    """Shader clock under load recorded in the committed profile of the force
    kernel (amd-smi clock samples taken during a long bench run)."""
    import re

    path = os.path.join(ROOT, "profiles", f"{kernel}_n16384_periodic.md")
    try:
        m = re.search(r"sclk_ghz_measured:\s*([0-9.]+)", open(path).read())
        return float(m.group(1)) if m else None
    except OSError:
        return None


def profile_force(kernel, reps=20, warm=600, copy_mib=256, boundary="periodic", n=None):
    """Launch pattern for counter collection: a device-to-device copy of a
    known size (calibrates FETCH_SIZE / WRITE_SIZE per byte moved), then the
    force kernel alone `warm` times (the clocks settle over the first ~0.1 s
    of load) and `reps` times more, timed, at the headline workload (or at n
    particles, same density)."""
    n = N if n is None else n
    import ctypes as C

    import ljamd

    hip = C.CDLL("libamdhip64.so")
    nbytes = copy_mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, C.c_size_t(nbytes))
    for _ in range(3):
        hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    jsl = 32 if kernel == "sliced" else 0
    bc = ljamd.BC_WALL if boundary == "wall" else ljamd.BC_PERIODIC
    with ljamd.LJSystem(n, RHO, T0, canonical=ljamd.TVN, boundary=bc, seed=SEED, jslices=jsl) as s:
        s.time_force_kernel(warm)
        us, span = s.time_force_kernel_stamped(reps)
        pair_once, nslots = s.force_plan()
    print(json.dumps({"profile": "force", "N": n, "boundary": boundary, "warm": warm, "reps": reps,
                      "kernel_us": us, "kernel_us_back_to_back": span,
                      "pair_once": pair_once, "nslots": nslots, "copy_bytes": nbytes}))


SWEEP_N = (2048, 3072, 4096, 6144, 8192, 10000, 12288, 16384)
# kernel choices of the sweep: the default plan, the pair-once pass with and
# without split items, the sliced kernel (environment switches read at
# context creation)
SWEEP_PLANS = (("default", {}), ("pair_once_split", {"LJ_PAIR_ONCE": "1", "LJ_PAIR_ONCE_SPLIT": "1"}),
               ("pair_once_unsplit", {"LJ_PAIR_ONCE": "1", "LJ_PAIR_ONCE_SPLIT": "0"}),
               ("sliced", {"LJ_PAIR_ONCE": "0"}))


def sweep_n(steps=800, reps=200):
    """TVN step (rho* = 0.85, periodic, dt = 0.005) and force pass over N for
    each kernel choice: the best of three timed runs of `steps` steps after 400
    warmup steps, and the force pass alone averaged over `reps` launches (HIP
    events). One JSON line per (N, plan)."""
    import ljamd

    keys = ("LJ_PAIR_ONCE", "LJ_PAIR_ONCE_SPLIT")
    saved = {k: os.environ.get(k) for k in keys}
    with ljamd.LJSystem(N, RHO, T0, canonical=ljamd.TVN, seed=SEED) as s:
        s.time_force_kernel(1500)  # clocks settle
    try:
        for n in SWEEP_N:
            for name, env in SWEEP_PLANS:
                for k in keys:
                    os.environ.pop(k, None)
                os.environ.update(env)
                with ljamd.LJSystem(n, RHO, T0, canonical=ljamd.TVN, seed=SEED) as s:
                    pair_once, nslots = s.force_plan()
                    s.time_force_kernel(reps)
                    us = s.time_force_kernel(reps)
                    s.Integrate(DT, 400)
                    s.synchronize()
                    best = float("inf")
                    for _ in range(3):
                        t0 = time.perf_counter()
                        s.Integrate(DT, steps)
                        s.synchronize()
                        best = min(best, time.perf_counter() - t0)
                print(json.dumps({"sweep": "tvn_step", "N": n, "plan": name, "pair_once": bool(pair_once),
                                  "slots": nslots, "force_us": round(us, 2),
                                  "step_us": round(1e6 * best / steps, 2),
                                  "pairs_per_s": n * n * steps / best}), flush=True)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    # 1000 warmup steps (~0.11 s): the GPU's clocks settle over the first
    # ~100 ms of sustained load (0.117 ms/step after 20 warmup steps, 0.110
    # after 1000, on the same box)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--profile-force", nargs="?", const="pair-once", default=None,
                    choices=["pair-once", "sliced"],
                    help="only a calibration copy and the force kernel (for rocprofv3)")
    ap.add_argument("--profile-boundary", default="periodic", choices=["periodic", "wall"],
                    help="boundary of the --profile-force workload")
    ap.add_argument("--profile-n", type=int, default=None,
                    help="particles of the --profile-force workload (default: the headline N)")
    ap.add_argument("--profile-rdf", action="store_true",
                    help="only the RDF passes of the rdf leg (for rocprofv3)")
    ap.add_argument("--profile-batch", action="store_true",
                    help="only the batched-ensemble leg")
    ap.add_argument("--profile-precision", action="store_true",
                    help="only the precision modes' leg")
    ap.add_argument("--profile-shard", action="store_true",
                    help="only one GPU's share of the sharded force pass (for rocprofv3)")
    ap.add_argument("--profile-adjoint", action="store_true",
                    help="only the differentiable workload's leg (for rocprofv3)")
    ap.add_argument("--sweep-n", action="store_true",
                    help="only the TVN step and force pass over a range of N (kernel choice)")
    a = ap.parse_args()
    if a.sweep_n:
        sweep_n()
        return
    if a.profile_force:
        profile_force(a.profile_force, boundary=a.profile_boundary, n=a.profile_n)
        return
    if a.profile_shard:
        print(json.dumps({"profile": "shard", "N": N, **shard_leg()}))
        return
    if a.profile_rdf:
        print(json.dumps({"profile": "rdf", "N": N, **rdf_leg()}))
        return
    if a.profile_batch:
        print(json.dumps({"profile": "batch", **batch_leg(a.batch_steps)}))
        return
    if a.profile_precision:
        print(json.dumps({"profile": "precision", **precision_leg()}))
        return
    if a.profile_adjoint:
        print(json.dumps({"profile": "adjoint", **adjoint_leg()}))
        return
    if a.gpus != 1:
        print(f"bench.py: this library runs on one GPU only (asked for {a.gpus}); "
              "the multi-GPU sharding is a design note, not a runnable path", file=sys.stderr)
        sys.exit(2)

    g = gpu_leg(a.steps, a.warmup)
    pairs = float(N) * N
    value = pairs * a.steps / g["t"]
    ms_per_step = 1000.0 * g["t"] / a.steps
    achieved = FLOP_PER_PAIR * pairs / (g["force_us"] * 1e-6) / 1e12
    sliced_us = sliced_kernel_us()
    import ljamd

    wall_steps = min(a.steps, 300)
    gw = gpu_leg(wall_steps, a.warmup, boundary=ljamd.BC_WALL)
    wall = {"steps": wall_steps, "pair_interactions_per_s": pairs * wall_steps / gw["t"],
            "ms_per_step": 1000.0 * gw["t"] / wall_steps, "force_kernel_us": gw["force_us"],
            "force_tflops": FLOP_PER_PAIR * pairs / (gw["force_us"] * 1e-6) / 1e12}
    batched = None if a.no_batch else batch_leg(a.batch_steps)
    adjoint = adjoint_leg()
    rdf = rdf_leg()
    precision = precision_leg()
    shard = shard_leg()
    cpu = None if a.no_cpu else cpu_leg()
    if cpu is not None:
        cpu["all_cores"] = cpu_parallel_leg()
    kname = "k_force_n3" if g["pair_once"] else "k_force"
    sclk = profiled_sclk_ghz(kname)
    # 256 CU x 64 FP32 lanes x 2 FLOP (FMA) per clock, non-packed
    peak_sclk = 256 * 64 * 2 * sclk / 1e3 if sclk else None
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (sc lattice + seeded Gaussian velocities)",
        "config": {"workload": "LJ all-pairs MD step, N=16384, rho*=0.85, T*=1.0, TVN every "
                               "step, periodic, dt*=0.005, FP32",
                   "N": N, "rho": RHO, "T0": T0, "dt": DT, "ensemble": "TVN",
                   "boundary": "periodic"},
        "pct_fp32_peak": 100.0 * FLOP_PER_PAIR * value / (PEAK_FP32_TFLOPS * 1e12),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "frac_scalar_fp32": achieved / PEAK_FP32_SCALAR_TFLOPS,
                     "sclk_ghz_measured": sclk,
                     "sclk_ghz_source": f"profiles/{kname}_n16384_periodic.md (amd-smi samples "
                                        "under load); not measured in this run",
                     "peak_scalar_at_measured_sclk": peak_sclk,
                     "frac_scalar_at_measured_sclk": achieved / peak_sclk if peak_sclk else None,
                     "kernel": ("k_force_n3 (each unordered pair once, periodic fixed point)"
                                if g["pair_once"] else "k_force (sliced, periodic fixed point)"),
                     "kernel_us": g["force_us"],
                     "kernel_us_source": f"mean over {g['reps']} back-to-back launches on the "
                                         "kernel's stream of each launch's own start / stop HIP "
                                         "events (hipExtLaunchKernel), this run, after "
                                         f"{g['settle_ms']:.0f} ms of settling force passes",
                     "kernel_us_back_to_back": g["force_span_us"],
                     "traffic": profiled_traffic(kname),
                     "traffic_source": f"profiles/{kname}_n16384_periodic.md (rocprofv3 --pmc "
                                       "FETCH_SIZE and WRITE_SIZE, scaled by a calibration copy); "
                                       "not measured in this run",
                     "newton3": g["pair_once"],
                     "pair_evaluations_per_launch": (N * (N - 1) // 2 if g["pair_once"]
                                                     else N * N),
                     "issue_slots_per_ordered_pair": ISSUE_SLOTS_N3 if g["pair_once"] else 22.5,
                     "issue_slots_source": f"profiles/{kname}_n16384_periodic.md (rocprofv3 "
                                           "--pmc SQ_INSTS_VALU); a constant, not measured in "
                                           "this run",
                     "sliced_kernel_us": sliced_us,
                     "sliced_kernel_tflops": FLOP_PER_PAIR * pairs / (sliced_us * 1e-6) / 1e12},
        "settle_ms": g["settle_ms"],
        "cpu_baseline": cpu,
        "hard_wall": wall,
        "batched": batched,
        "adjoint": adjoint,
        "rdf": rdf,
        "precision": precision,
        "shard": shard,
        "final_T": g["obs"]["T"],
        "final_u": g["obs"]["U"] / N,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
