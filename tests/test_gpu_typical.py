"""The reference's documented typical case (docs/source/example/typical/data/exec.sh) rerun
through the drop-in: the tcsolver program on an initial-condition directory in the
reference's format with the reference generator's clipped-Chebyshev grid (64 x 8 x 32,
radius ratio 1/2, ly = pi/8, lz = 2, Re = 80, Pr = 1, CFL coefficients 1, random initial
fields).  The flow settles into steady Taylor vortices; the logged energy injection and
dissipation, normalised by the laminar torque as the reference's balance_main.py does, give
the Nusselt number that the documentation compares with Ostilla et al. (JFM 2013), 1.1375.
Two initial-field seeds reach the same state."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "simpletcsolver_amd", "tcsolver")
with open(os.path.join(ROOT, "tests", "golden", "typical_case.json")) as fh:
    CASE = json.load(fh)
LENGTHS = tuple(CASE["lengths"])


def _run(tmp, seed, timemax, precision=None):
    from simpletcsolver_amd import icdir
    icdir.write(os.path.join(tmp, "ic"), CASE["is_curved"], LENGTHS, tuple(CASE["glsizes"]), seed=seed)
    for sub in ("log", "save", "stat"):
        os.makedirs(os.path.join(tmp, "output", sub), exist_ok=True)
    env = dict(os.environ)
    env.update({"timemax": repr(timemax), "wtimemax": "600", "log_rate": "1.0", "save_rate": "1000",
                "save_after": "0", "stat_rate": "0.1", "stat_after": "1000",
                "coef_dt_adv": repr(CASE["coef_dt_adv"]), "coef_dt_dif": repr(CASE["coef_dt_dif"]),
                "Re": repr(CASE["Re"]), "Pr": repr(CASE["Pr"])})
    if precision:
        env["TC_PRECISION"] = precision
    print(f"tcsolver typical case, seed {seed}, timemax {timemax}", flush=True)
    r = subprocess.run([EXE, "ic"], cwd=tmp, env=env, capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    inj = np.loadtxt(os.path.join(tmp, "output", "log", "injection.dat"))
    dis = np.loadtxt(os.path.join(tmp, "output", "log", "dissipation.dat"))
    div = np.loadtxt(os.path.join(tmp, "output", "log", "divergence.dat"))
    return inj, dis, div


@pytest.mark.parametrize("seed", [0, 1])
def test_typical_case_nusselt(need_gpu, tmp_path, seed):
    from simpletcsolver_amd import icdir
    inj, dis, div = _run(str(tmp_path), seed, 200.0)
    f = icdir.laminar_torque_factor(CASE["Re"], LENGTHS)
    late = inj[:, 0] > 150.0
    nu_i = inj[late, 1].mean() / f
    nu_d = dis[late, 1].mean() / f
    print(f"seed {seed}: Nu from injection {nu_i:.5f}, from dissipation {nu_d:.5f} (Ostilla et al. 1.1375); "
          f"max div {div[:, 1].max():.1e}")
    # the discretisation's own value at this resolution is 1.1345 (0.26 % below the reference
    # value the documentation draws)
    assert abs(nu_i - CASE["nusselt"]) <= 5e-3 * CASE["nusselt"]
    # energy balance of the steady state: injection equals dissipation (velocity and scalar)
    assert abs(nu_i - nu_d) <= 1e-5 * nu_i
    assert abs(inj[late, 2].mean() - dis[late, 2].mean()) <= 1e-5 * abs(inj[late, 2].mean())
    assert div[:, 1].max() <= 1e-12


@pytest.mark.parametrize("precision", ["mixed", "mixed_refined"])
def test_typical_case_mixed_precision(need_gpu, tmp_path, precision):
    """the same run through the reference API with TC_PRECISION set: the steady Nusselt number
    within 1e-4 of the FP64 run's 1.13453; the refined form keeps FP64-level divergence"""
    from simpletcsolver_amd import icdir
    inj, dis, div = _run(str(tmp_path), 0, 200.0, precision)
    f = icdir.laminar_torque_factor(CASE["Re"], LENGTHS)
    late = inj[:, 0] > 150.0
    nu = inj[late, 1].mean() / f
    print(f"{precision}: Nu {nu:.5f}, max div {div[:, 1].max():.1e}")
    assert abs(nu - 1.13453) <= 1e-4 * 1.13453
    assert div[:, 1].max() <= (1e-12 if precision == "mixed_refined" else 1e-6)

