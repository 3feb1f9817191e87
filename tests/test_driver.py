"""The C driver (forestfv_public_amd/fv_run): the reference's time loop
(loop.c:12-30, main.c:107-134) on the C ABI alone, no Python in the loop."""
import os
import subprocess

import numpy as np
import pytest

import forestfv_public_amd as F
import oracle_lib as O
from cases import sod_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "forestfv_public_amd", "fv_run")


def run(*args, timeout=120):
    if not os.path.exists(EXE):
        pytest.fail("fv_run is not built (make -C forestfv_public_amd)")
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    return r


def drift(out, tag):
    line = [l for l in out.splitlines() if l.startswith(tag)][-1]
    w = line.split()
    return [float(w[i + 1]) for i, t in enumerate(w) if t in ("water", "air", "left", "right", "energy")]


def test_driver_prints_usage_without_arguments():
    r = run()
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F.PREC_EXACT, F.PREC_FP64])
def test_driver_sod_matches_oracle(tmp_path, prec):
    level, steps = 5, 12
    out = tmp_path / "sod.bin"
    r = run("sod", level, steps, prec, out)
    assert r.returncode == 0
    raw = np.fromfile(out, dtype=np.int32, count=2)
    n, nq = int(raw[0]), int(raw[1])
    S = np.fromfile(out, dtype=np.float64, offset=8).reshape(n, nq)
    p, m, S0 = sod_case(level=level)
    ref = O.OracleMesh(m).advance(p, S0, steps)
    if prec == F.PREC_EXACT:
        assert np.array_equal(S, ref)
    else:
        assert np.abs(S - ref).sum() / np.abs(ref).sum() < 1e-10
    assert max(drift(r.stdout, "fv_run sod:")) < 1e-13


@pytest.mark.gpu
def test_driver_device_and_host_regrids_give_the_same_run():
    """fv_dev_regrid in the loop leaves the same meshes and states as the host path."""
    outs = []
    for regrid in (1, 0):
        r = run("bubble", 3, 6, 3, 30, 5, F.PREC_FP64, 4, regrid)
        assert r.returncode == 0
        outs.append([l for l in r.stdout.splitlines() if l.startswith(("final mesh", "fv_run bubble:"))])
    assert outs[0] == outs[1] and len(outs[0]) == 2


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [F.PREC_FP64, F.PREC_MIXED])
def test_driver_bubble_amr_loop_conserves_in_a_closed_box(prec):
    # levels 3..5, two interface layers, regrid every 5 steps, walls on all sides
    r = run("bubble", 3, 5, 2, 20, 5, prec, 4)
    assert r.returncode == 0
    assert "final mesh" in r.stdout
    d = drift(r.stdout, "fv_run bubble:")
    assert len(d) == 3
    # merge averages and split distributes conservatively; the stage is a
    # conservative face-flux update with walls on every side
    assert max(d) < (1e-12 if prec == F.PREC_FP64 else 1e-6)


@pytest.mark.gpu
def test_driver_runs_the_bubble_to_its_first_collapse():
    """The whole loop on the device (stages, interface layers every step, device
    regrid every 5 steps, CFL step) on a coarse forest (levels 4-7): the bubble
    collapses near the Rayleigh time 0.915 R0 sqrt(rho / dp), later on this coarse
    mesh (the bubble radius is six cells), and the air mass stays in the bubble."""
    import json
    r = run("collapse", 4, 7, 8, 5, F.PREC_FP64, 3000, timeout=300)
    assert r.returncode == 0
    res = json.loads([l for l in r.stdout.splitlines() if l.startswith("{")][-1])
    assert res["collapsed"] is True
    assert 1.0 < res["t_collapse_over_rayleigh"] < 1.6
    assert 0.3 < res["R_min_over_R0"] < 0.8
    assert abs(res["air_mass_change"]) < 1e-4
    assert res["regrids"] > 50 and res["leaves_min"] < res["leaves_start"]
