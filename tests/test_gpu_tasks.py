"""The task programs (lennard-jones-cuda_amd/tasks: run_isotherm,
run_fluctuations) on the device, against the reference's driver loops
(run-isotherm.cpp:84-166, run-fluctuations.cpp:79-199) restated here over the
oracle (the reference CPU path restated, bit-exact to it) from the same seeded
initial states.

The time averages of a run follow its trajectory, and the device's forces
round differently from the reference CPU path's, so two runs drift apart
slowly; the runs are short (300 steps of a dilute gas) and the averages are
compared within about 10x the measured differences (MI355X: 1e-5 relative on
the averages, which the 6 printed digits limit, 4e-5 on their error estimates).  The
statistics chain itself (device sums, time averages, printed columns) is
bit-exact elsewhere (tests/test_gpu_fluct.py, tests/test_tasks_cpu.py)."""
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import (ORC, OracleSystem, fl_ladder, fl_nsub_batch, fl_vz_batch, lattice_state,
                        tavg_eval)

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "lennard-jones-cuda_amd", "lib")


def _run(prog, params, cwd):
    path = os.path.join(cwd, "params.txt")
    with open(path, "w") as f:
        f.write("# test run\n")
        for k, v in params.items():
            f.write(f"{k} {v}\n")
    exe = os.path.join(LIB, prog)
    r = subprocess.run([exe, path], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       cwd=cwd, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _table(path):
    rows = [line.split() for line in open(path).read().splitlines()[1:] if line.strip()]
    return np.array([[float(x) for x in r] for r in rows])


def _steps(t, tend, dt):
    n = 0
    while t < tend:
        t += dt
        n += 1
    return n, t


def _oracle_isotherm(N, T0, rhos, dt, teq, tfin, seed):
    neq, t = _steps(0.0, teq, dt)
    nprod, t = _steps(t, tfin, dt)
    out = []
    for k, rho in enumerate(rhos):
        L, pos, vel = lattice_state(N, rho, T0, seed + k)
        o = OracleSystem(N, T0, rho, 1, 0, nthreads=8)
        o.set_state(pos, vel)
        o.forces()
        o.params()
        o.integrate(dt, neq)
        U0, P0 = o.s.U, o.s.P
        h = o.integrate(dt, nprod)
        u = np.concatenate([[U0], h[:-1, 0]]) / N
        p = np.concatenate([[P0], h[:-1, 4]])
        out.append((tavg_eval(u), tavg_eval(p), h[-1, 1]))
    return out, t


@pytest.mark.parametrize("batched", [1, 0])
def test_run_isotherm_vs_oracle(tmp_path, batched):
    N, T0, dt, teq, tfin, seed = 400, 1.4, 0.004, 0.2, 1.4, 3
    params = {"N": N, "T*": T0, "rho*_min": 0.05, "rho*_max": 0.25, "drho*": 0.1, "teq": teq,
              "tfin": tfin, "dt*": dt, "useCUDA": 1, "seed": seed, "timestamp": 0,
              "print_every": 100, "batched": batched, "output_prefix": "iso"}
    print(f"\n[case] run_isotherm batched={batched}", flush=True)
    out = _run("run_isotherm", params, str(tmp_path))
    rhos = []
    rho = 0.05
    while rho <= 0.25:
        rhos.append(rho)
        rho += 0.1
    tab = _table(tmp_path / "iso.N400.Tst1.4.dat")
    assert tab.shape == (len(rhos), 13)
    assert np.array_equal(tab[:, 0], np.round(np.array(rhos), 6))
    exp, t_end = _oracle_isotherm(N, T0, rhos, dt, teq, tfin, seed)
    assert np.allclose(tab[:, 1], t_end - dt, rtol=1e-6)
    worst = {}
    for k, (uo, po, Tend) in enumerate(exp):
        row = tab[k]
        for name, got, ref in (("<u*>", row[3], uo[0]), ("<p*>", row[6], po[0]),
                               ("T_end", row[2], Tend), ("<Z>", row[9], po[0] / rhos[k] / T0),
                               ("d<u*>", row[4], uo[3]), ("d<p*>", row[7], po[3])):
            d = abs(got - ref) / max(abs(ref), 1e-12)
            worst[name] = max(worst.get(name, 0.0), d)
    print("  relative differences to the oracle: " +
          ", ".join(f"{k} {v:.1e}" for k, v in worst.items()))
    for name in ("<u*>", "<p*>", "T_end", "<Z>"):
        assert worst[name] <= 1e-4, (name, worst[name])
    for name in ("d<u*>", "d<p*>"):  # error estimates: second moments of short series
        assert worst[name] <= 5e-4, (name, worst[name])
    # console: a row every 100 observations of each density, 300 observations
    rows = [line for line in out.splitlines() if line.strip() and line.split()[0][0].isdigit()]
    assert len(rows) == 3 * len(rhos)


def _oracle_fluctuations(N, T0, rho, dt, teq, tfin, seed, dalpha, every):
    neq, t = _steps(0.0, teq, dt)
    L, pos, vel = lattice_state(N, rho, T0, seed)
    o = OracleSystem(N, T0, rho, 1, 0, nthreads=8)
    o.set_state(pos, vel)
    o.forces()
    o.params()
    o.integrate(dt, neq)
    o.s.av_U_tot = o.s.av_T_tot = o.s.av_p_tot = 0.0
    o.s.av_iters = 0
    vmax = 3.0 * np.sqrt(T0)
    series, rows = [], []
    it = 0
    while t < tfin:
        o.integrate(dt, 1)
        t += dt
        it += 1
        p, v = o.pos.copy(), o.vel.copy()
        c = [fl_nsub_batch(p, L, dalpha, ty) for ty in (0, 1, 2, 3)]
        c.append(fl_vz_batch(v, vmax, dalpha, 2))
        c.append([ORC.orc_fl_nsub(N, p, L, 0.5, 2)])
        series.append(np.concatenate([np.asarray(x, np.float64) for x in c]))
        if it % every == 0:
            s = o.s
            Tav = s.av_T_tot / s.av_iters
            half = np.array(series)[:, -1]
            Nav, N2av = half.mean(), (half * half).mean()
            rows.append([t, s.U / N, s.T, s.P / (rho * s.T), s.av_U_tot / s.av_iters / N, Tav,
                         s.av_p_tot / s.av_iters / (rho * Tav), (N2av - Nav * Nav) / Nav / 0.5])
    return np.array(rows), np.array(series)


def test_run_fluctuations_vs_oracle(tmp_path):
    N, T0, rho, dt, teq, tfin, seed, dalpha, every = 400, 1.4, 0.05, 0.004, 0.2, 1.4, 2, 0.05, 100
    params = {"N": N, "T*": T0, "rho*": rho, "teq": teq, "tfin": tfin, "dt*": dt, "canonical": 1,
              "subvolume_spacing": dalpha, "useCUDA": 1, "seed": seed, "timestamp": 0,
              "print_every": every, "output_prefix": "fl"}
    print("\n[case] run_fluctuations", flush=True)
    out = _run("run_fluctuations", params, str(tmp_path))
    rows = np.array([[float(x) for x in line.split()] for line in out.splitlines()
                     if line.strip() and line.split()[0][0].isdigit() and len(line.split()) == 8])
    exp_rows, series = _oracle_fluctuations(N, T0, rho, dt, teq, tfin, seed, dalpha, every)
    assert rows.shape == exp_rows.shape == (3, 8)
    d = np.abs(rows[:, :7] - exp_rows[:, :7]) / np.maximum(np.abs(exp_rows[:, :7]), 1e-12)
    print(f"  console rows vs oracle: max relative difference {d.max():.1e}; "
          f"w column {rows[:, 7]} vs {exp_rows[:, 7]}")
    assert d.max() <= 1e-4
    assert np.allclose(rows[:, 7], exp_rows[:, 7], rtol=1e-4)
    prefix = tmp_path / f"fl.N{N}.Tst{T0}.rhost{rho}"
    td = _table(str(prefix) + ".TimeDep.txt")
    assert td.shape == (3, 13)
    assert np.allclose(td[:, :7], rows[:, :7], rtol=1e-5)
    # the counter files after 300 samples: <N> and scaled variances of each ladder
    alphas = fl_ladder(dalpha, 1.0 - 1e-9)
    m = alphas.size
    for k, name in enumerate(("flucsX", "flucsY", "flucsZ", "flucsCube")):
        f = _table(str(prefix) + f".{name}.dat")
        assert f.shape == (m, 6)
        assert np.allclose(f[:, 0], np.round(alphas, 6), rtol=1e-5)
        s = series[:, k * m:(k + 1) * m]
        assert np.allclose(f[:, 1], s.mean(axis=0), rtol=1e-5, atol=0.05), name
        wv = s.var(axis=0) / s.mean(axis=0)
        assert np.allclose(f[:, 2], wv, rtol=0.1, atol=0.02), name
    fv = _table(str(prefix) + ".flucsVz.dat")
    sv = series[:, 4 * m:4 * m + fv.shape[0]]
    assert np.allclose(fv[:, 1], sv.mean(axis=0), rtol=1e-5, atol=0.05)
    g = _table(str(prefix) + ".RDF.dat")
    assert g.shape == (256, 2) and np.all(g[:, 1] >= 0) and g[:, 1].max() > 0.5
