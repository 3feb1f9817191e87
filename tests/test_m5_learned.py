"""Config 4's pieces on model 5: the MUSCL slope choice of the model-5 reconstruction (minmod,
monotonised central, or per leaf) and the learned indicator on a model-5 context.

Oracle pins (CPU): the MC slope of orc_m5_faceinterp on a line of cells against a brute-force
phi(r) b form of the monotonised central limiter, phi(r) = max(0, min(2r, (1 + r) / 2, 2)), a
different expression from the oracle's minmod(2a, 2b, (a + b) / 2); limiter 2 takes MC exactly on
the flagged leaves.

Device tests: face states and steps with limiters 1 and 2 equal the oracle; the model-5 features
and labels follow the plain indicator; the learned targets keep every plain refinement; a regrid on
the learned targets equals the oracle's criterionsmth + refineh of the same targets."""
import numpy as np
import pytest

import forestfv_public_amd as F
import oracle_lib as O
from test_m5_mixture import MIX, RHO, L_BOX, c2_init, c2_params, mix_state
from test_m5_regrid import c2_amr, c2_crit


def line_params(limiter):
    return F.make_m5_params(nmat=3, gamma=(3.0, 6.12, 1.4), pinf=(2e10, 3.43e8, 0.0), mush=(26e9, 0.0, 0.0),
                            amin=1e-6, avf_lim=0.5, prelax=0, order=2, rk_steps=3, fix=MIX, limiter=limiter)


def mc_phi(r):
    return np.maximum(0.0, np.minimum(np.minimum(2.0 * r, 0.5 * (1.0 + r)), 2.0))


def test_m5_mc_slope_brute_force_on_a_line():
    N, L = 24, 1e-2
    h = L / N
    m = F.Mesh.uniform((N, 1, 1), (0, 0, 0), h, 1, [2] * 6)
    x = m.centres()[:, 0]
    rng = np.random.default_rng(5)
    # a water density profile with smooth parts, a kink and a jump: every branch of phi
    rho_w = 1000.0 + 5.0 * np.sin(2 * np.pi * x / L) + np.where(x > 0.6 * L, 8.0, 0.0) + rng.normal(size=N) * 0.2
    S = np.array([mix_state(line_params(1), [1e-6, 1 - 2e-6, 1e-6], [2700.0, r, 1.2], 1e7, [0, 0, 0], np.zeros((3, 3)))
                  for r in rho_w])
    om = O.OracleMesh(m)
    SF1 = om.m5_faceinterp(line_params(1), S).reshape(N, 6, -1)
    SF0 = om.m5_faceinterp(line_params(0), S).reshape(N, 6, -1)
    q = S[:, 1]                               # alpha rho of the water
    hit = set()
    for i in range(1, N - 1):
        a, b = (q[i] - q[i - 1]) / h, (q[i + 1] - q[i]) / h
        if b != 0.0 and abs(q[i + 1] - q[i]) > 1e-10:
            r = a / b
            slope = mc_phi(r) * b
            hit.add("zero" if r <= 0 else "2r" if 2 * r <= min(0.5 * (1 + r), 2) else
                    "central" if 0.5 * (1 + r) <= 2 else "2")
        else:
            slope = 0.0
        # face 1 is +x, face 3 is -x (operatetp.c face order)
        assert np.isclose(SF1[i, 1, 1], q[i] + 0.5 * h * slope, rtol=1e-14, atol=1e-12), i
        assert np.isclose(SF1[i, 3, 1], q[i] - 0.5 * h * slope, rtol=1e-14, atol=1e-12), i
    assert {"zero", "2r", "central"} <= hit
    assert not np.allclose(SF0[:, :, 1], SF1[:, :, 1])
    # limiter 2: MC on the flagged leaves, minmod on the others
    flags = (rng.random(N) < 0.5).astype(np.uint8)
    O.m5_set_cell_limiter(flags)
    try:
        SF2 = om.m5_faceinterp(line_params(2), S).reshape(N, 6, -1)
    finally:
        O.m5_set_cell_limiter(None)
    assert np.array_equal(SF2[flags == 1], SF1[flags == 1])
    assert np.array_equal(SF2[flags == 0], SF0[flags == 0])


# ---------------------------------------------------------------- device

def say(*a):
    print(*a, flush=True)


@pytest.mark.gpu
@pytest.mark.parametrize("limiter", [1, 2])
def test_m5_device_limiters_match_oracle(limiter):
    p, m, S, crit = c2_amr(layers=2)
    p.limiter = limiter
    om = O.OracleMesh(m)
    p.dt = om.m5_cfl_dt(p, S, 0.4)
    rng = np.random.default_rng(limiter)
    flags = (rng.random(m.n) < 0.5).astype(np.uint8)
    say(f"[m5 limiter {limiter}] n={m.n}")
    dev = F.M5Device(p, m)
    dev.set_state(S)
    if limiter == 2:
        dev.set_limiter_flags(flags)
        assert np.array_equal(dev.get_limiter_flags(), flags)
        O.m5_set_cell_limiter(flags)
    try:
        dev.faceinterp()
        SF, _, _ = dev.stage_arrays()
        SFo = om.m5_faceinterp(p, S)
        dev.advance(3)
        Sd = dev.get_state()
        So = om.m5_advance(p, S, 3)
    finally:
        O.m5_set_cell_limiter(None)
        dev.close()
    scale = np.abs(SFo.reshape(m.n, -1)).max(axis=0) + 1e-300
    assert (np.abs(SF.reshape(m.n, -1) - SFo.reshape(m.n, -1)).max(axis=0) / scale).max() <= 1e-11
    scale = np.abs(So).max(axis=0) + 1e-300
    err = (np.abs(Sd - So).max(axis=0) / scale).max()
    say(f"[m5 limiter {limiter}] three steps, largest difference {err:.2e}")
    assert err <= 1e-10


@pytest.mark.gpu
def test_m5_learned_indicator_features_labels_and_forced_refinement():
    p, m, S, crit = c2_amr(layers=2)
    om = O.OracleMesh(m)
    p.dt = om.m5_cfl_dt(p, S, 0.4)
    dev = F.M5Device(p, m)
    dev.set_state(S)
    dev.set_forest(m)
    dev.advance(5)
    Sd = dev.get_state()
    SI, plain = dev.indicator(crit)
    mlp = F.MLP(seed=5)
    dev.sample_mlp(mlp, crit)
    X, cls, sm = mlp.dataset()
    assert X.shape == (m.n, 112) and np.all(X[:, 108:] == 0)
    c = 13
    assert np.all(X[:, 4 * c] == 0) and np.all(X[:, 4 * c + 1] == 0)
    # the centre's liquid fraction and |u| / c from the oracle's record of the state
    recs = [O.m5_record(p, s) for s in Sd[:200]]
    al = np.array([r["avf"][1] for r in recs], np.float32)
    assert np.allclose(X[:200, 4 * c + 2], np.clip(al, -8, 8), rtol=1e-6, atol=1e-7)
    lvl = m.arrays()["level"]
    assert np.array_equal(cls, np.where(plain > lvl, 0, np.where(plain < lvl, 2, 1)))
    # a barely trained network: its targets still keep every plain refinement, and its smooth flags
    # become the limiter flags
    mlp.train(1, batch=256, lr=2e-3)
    SI2, learned, smooth = dev.indicator_mlp(mlp, crit)
    assert np.array_equal(SI2, SI)
    assert np.all(learned[plain >= lvl] >= plain[plain >= lvl])
    assert np.array_equal(dev.get_limiter_flags(), smooth)
    dev.close()
    mlp.close()


@pytest.mark.gpu
def test_m5_regrid_on_learned_targets_matches_oracle():
    p, m, S, crit = c2_amr(layers=2)
    maxlevel = 5
    om = O.OracleMesh(m)
    p.dt = om.m5_cfl_dt(p, S, 0.4)
    mlp = F.MLP(seed=9)
    dev = F.M5Device(p, m)
    dev.set_state(S)
    dev.set_forest(m)
    dev.advance(2)
    dev.sample_mlp(mlp, crit)
    mlp.train(2, batch=256, lr=2e-3)
    S2 = dev.get_state()
    SI, learned, _ = dev.indicator_mlp(mlp, crit)
    # the oracle adapts the same state to the same targets
    m.set_SI(SI)
    split = O.OracleMesh(m).criterionsmth(maxlevel, learned)
    out = O.OracleMesh(m).m5_refineh(p, split, S2)
    nn = dev.regrid_mlp(mlp, crit, maxlevel)
    flags = dev.get_limiter_flags()
    root, level, ci, SIn = dev.get_forest()
    say(f"[m5 learned regrid] {m.n} -> {nn} leaves (oracle {len(out['level'])})")
    assert nn == len(out["level"])
    assert np.array_equal(level, out["level"]) and np.array_equal(ci, out["ci"]) and np.array_equal(SIn, out["SI"])
    scale = np.abs(out["S"]).max(axis=0) + 1e-300
    assert (np.abs(dev.get_state() - out["S"]).max(axis=0) / scale).max() <= 1e-12
    # the new leaves' limiter flags are the network's smooth flags on the new mesh
    _, _, smooth = dev.indicator_mlp(mlp, crit)
    dev.close()
    mlp.close()
    assert smooth.shape == (nn,) and np.array_equal(flags, smooth)
