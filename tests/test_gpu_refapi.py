"""The reference-shaped API end to end: the tcsolver executable (main.c semantics over
domain_init, fluid_init, integrate, logging, save and statistics of include/*.h) run on an
initial-condition directory in the reference's NPY format, against the oracle stepped the
same way.

* the checkpoint written at the end (output/save/stepNNNNNNNNNN, save.c:74-110) holds the
  oracle's fields after the same steps (relative L2 <= 1e-10; p modulo its mean);
* the log files (logging/*.c formats) hold the oracle's diagnostics at the logged steps;
* output/stat/stepNNNNNNNNNN holds the oracle's running statistics (statistics.c:257-300);
* restarting from a periodic checkpoint reproduces the final checkpoint bit for bit.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import icgen
from harness import make_case, modulo_mean, oracle_for, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "simpletcsolver_amd", "tcsolver")
EPS = sys.float_info.epsilon


def write_ic_dir(dest, case, step=0, time=0.0):
    os.makedirs(dest, exist_ok=True)
    nx, ny, nz = case["nx"], case["ny"], case["nz"]
    np.save(f"{dest}/step.npy", np.array(step, dtype=np.uint64))
    np.save(f"{dest}/time.npy", np.array(time, dtype=np.float64))
    np.save(f"{dest}/is_curved.npy", np.array(bool(case["kw"]["is_curved"]), dtype=np.bool_))
    np.save(f"{dest}/xf.npy", case["xf"])
    np.save(f"{dest}/xc.npy", case["xc"])
    np.save(f"{dest}/glsizes.npy", np.array([nx, ny, nz], dtype=np.uint64))
    np.save(f"{dest}/lengths.npy", np.array(case["kw"]["lengths"], dtype=np.float64))
    for n in ("ux", "uy", "uz", "p", "t"):
        np.save(f"{dest}/{n}.npy", case["fields"][n])


def env_for(case, dt, timemax, log_rate, save_rate, stat_rate):
    kw = case["kw"]
    e = dict(os.environ)
    e.update({
        "Re": repr(kw["Re"]), "Pr": repr(kw["Pr"]), "timemax": repr(timemax), "wtimemax": "600",
        "log_rate": repr(log_rate), "save_rate": repr(save_rate), "save_after": "0",
        "stat_rate": repr(stat_rate), "stat_after": "0", "coef_dt_adv": "1", "coef_dt_dif": "1",
        "dt_fixed": repr(dt),
        "implicit_x": str(kw["implicit"][0]), "implicit_y": str(kw["implicit"][1]),
        "implicit_z": str(kw["implicit"][2]), "scalar": str(kw["scalar"]),
        "uy_xm": repr(kw["uy_walls"][0]), "uy_xp": repr(kw["uy_walls"][1]),
        "t_xm": repr(kw["t_walls"][0]), "t_xp": repr(kw["t_walls"][1]),
    })
    return e


def schedule(step, time, dt, timemax, log_rate, save_rate, stat_rate):
    """the event order of main.c:86-117 with the schedulers of logging/main.c:20-40,
    save.c:31-65 and statistics.c:56-101"""
    nxt = {"log": log_rate * math.ceil(max(EPS, time) / log_rate),
           "save": save_rate * math.ceil(max(EPS, max(time, 0.0)) / save_rate),
           "stat": stat_rate * math.ceil(max(EPS, max(time, 0.0)) / stat_rate)}
    rate = {"log": log_rate, "save": save_rate, "stat": stat_rate}
    events = []
    while True:
        step += 1
        time += dt
        events.append(("step", step))
        if timemax < time:
            break
        for kind in ("log", "save", "stat"):
            if nxt[kind] < time:
                events.append((kind, step))
                nxt[kind] += rate[kind]
    return step, events


def run_exe(cwd, icdir, env):
    for sub in ("log", "save", "stat"):
        os.makedirs(os.path.join(cwd, "output", sub), exist_ok=True)
    print(f"tcsolver {icdir} in {cwd}", flush=True)
    r = subprocess.run([EXE, icdir], cwd=cwd, env=env, capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:])
    print(r.stderr[-2000:])
    assert r.returncode == 0
    return r.stdout


def _rows(path):
    with open(path) as fh:
        return [[float(v) for v in line.split()] for line in fh if line.strip()]


@pytest.mark.parametrize("implicit", [(1, 0, 0), (1, 1, 1)])
def test_tcsolver_matches_oracle(need_gpu, tmp_path, implicit):
    assert os.path.exists(EXE), "build with make -C simpletcsolver_amd/csrc"
    case = make_case(nx=16, ny=12, nz=8, Re=300.0, implicit=implicit, amp=0.1, t_amp=0.05,
                     t_walls=(1.0, 0.0))
    dt, timemax, log_rate, save_rate, stat_rate = 1e-3, 0.0095, 0.002, 0.005, 0.002
    icdir = str(tmp_path / "ic")
    write_ic_dir(icdir, case)
    env = env_for(case, dt, timemax, log_rate, save_rate, stat_rate)
    run_exe(str(tmp_path / "run"), icdir, env)
    last, events = schedule(0, 0.0, dt, timemax, log_rate, save_rate, stat_rate)

    o = oracle_for(dict(case, kw=dict(case["kw"], dt_fixed=dt)))
    logs, saves = [], {}
    for kind, step in events:
        if kind == "step":
            o.integrate()
        elif kind == "log":
            logs.append((o.divergence(), o.injection(), o.dissipation(), o.total_energy()))
        elif kind == "save":
            saves[step] = o.npy_fields()
        elif kind == "stat":
            o.stat_collect()
    saves[last] = o.npy_fields()

    out = tmp_path / "run" / "output"
    for step, want in saves.items():
        d = out / "save" / f"step{step:010d}"
        assert int(np.load(d / "step.npy")) == step
        for n in ("ux", "uy", "uz", "p", "t"):
            got = np.load(d / f"{n}.npy")
            a, b = (modulo_mean(got, case), modulo_mean(want[n], case)) if n == "p" else (got, want[n])
            e = rel_l2(a, b)
            print(f"save {step} {n}: {e:.3e}")
            assert e <= 1e-10, (step, n)
        assert np.array_equal(np.load(d / "xf.npy"), case["xf"])

    div = _rows(out / "log" / "divergence.dat")
    inj = _rows(out / "log" / "injection.dat")
    dis = _rows(out / "log" / "dissipation.dat")
    ene = _rows(out / "log" / "total_energy.dat")
    assert len(div) == len(inj) == len(dis) == len(ene) == len(logs) > 0
    with open(out / "log" / "progress.dat") as fh:
        assert len(fh.read().splitlines()) == len(logs)
    for n, (d0, i0, s0, e0) in enumerate(logs):
        assert div[n][1] <= 1e-12 and d0 <= 1e-12
        np.testing.assert_allclose(inj[n][1:], i0, rtol=1e-9, atol=1e-14)
        np.testing.assert_allclose(dis[n][1:], s0, rtol=1e-9, atol=1e-14)
        np.testing.assert_allclose(ene[n][1:], e0, rtol=1e-9, atol=1e-14)

    nstat = sum(1 for k, _ in events if k == "stat")
    d = out / "stat" / f"step{last:010d}"
    assert int(np.load(d / "num.npy")) == nstat
    for name, want in o.stats().items():
        got = np.load(d / f"{name}.npy")
        assert got.shape == want.shape
        e = rel_l2(got, want)
        print(f"stat {name}: {e:.3e}")
        assert e <= 1e-10, name


def test_tcsolver_restart_is_bit_exact(need_gpu, tmp_path):
    case = make_case(nx=16, ny=12, nz=8, Re=300.0, implicit=(1, 1, 0), amp=0.1, t_amp=0.05)
    dt, timemax, log_rate, save_rate, stat_rate = 1e-3, 0.0095, 1.0, 0.005, 1.0
    icdir = str(tmp_path / "ic")
    write_ic_dir(icdir, case)
    env = env_for(case, dt, timemax, log_rate, save_rate, stat_rate)
    run_exe(str(tmp_path / "full"), icdir, env)
    last, events = schedule(0, 0.0, dt, timemax, log_rate, save_rate, stat_rate)
    mids = [s for k, s in events if k == "save"]
    assert mids and mids[0] < last
    full = tmp_path / "full" / "output" / "save"
    run_exe(str(tmp_path / "again"), str(full / f"step{mids[0]:010d}"), env)
    again = tmp_path / "again" / "output" / "save"
    for n in ("ux", "uy", "uz", "p", "t", "time", "step"):
        a = np.load(full / f"step{last:010d}" / f"{n}.npy")
        b = np.load(again / f"step{last:010d}" / f"{n}.npy")
        assert np.array_equal(a, b), n
