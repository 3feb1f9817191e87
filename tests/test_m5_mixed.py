"""Mixed-precision model-5 stage (fv_m5_dev_set_precision(FV_PREC_MIXED)): FP32 limited
slopes, face records, HLLC, star states and GFM ghost flux; the records are formed in
FP64 and narrowed, and the residual, RK accumulation and relaxation stay FP64.

Measured against the FP64 device stage, which equals the oracle (test_m5_stage,
test_m5_consistent): relative L1 of every conserved field (momentum on the momentum
magnitude) and of the mixture pressure within 1e-4 after several steps, the bar the
mixed-precision path is held to; the GFM flags are the same, and the published and
consistent updates are both covered."""
import numpy as np
import pytest

import forestfv_public_amd as F
import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def field_l1(a, b, nm):
    """Per conserved field relative L1 of a against b (momentum rows on the momentum magnitude)."""
    num = np.abs(a - b).sum(0)
    den = np.abs(b).sum(0)
    mom = np.sqrt((b[:, nm:nm + 3] ** 2).sum(1)).sum()
    den[nm:nm + 3] = np.maximum(den[nm:nm + 3], mom)
    den[den == 0] = 1.0
    return num / den


def mixture_pressure(p, S):
    return np.array([(O.m5_record(p, s)["avf"] * O.m5_record(p, s)["p_hydro"]).sum() for s in S])


def run(p, m, S, steps, prec):
    dev = F.M5Device(p, m)
    dev.set_precision(prec)
    dev.set_state(S)
    dev.advance(steps)
    out = dev.get_state()
    dev.close()
    return out


def check(p, m, S, steps, tol=1e-4, what=""):
    a = run(p, m, S, steps, F.PREC_MIXED)
    b = run(p, m, S, steps, F.PREC_FP64)
    assert np.all(np.isfinite(a))
    e = field_l1(a, b, p.nmat)
    pa, pb = mixture_pressure(p, a), mixture_pressure(p, b)
    ep = np.abs(pa - pb).sum() / np.abs(pb).sum()
    say(f"[m5 mixed] {what} n={m.n} steps={steps}: max field L1 {e.max():.2e}, pressure L1 {ep:.2e}")
    assert e.max() <= tol, e
    assert ep <= tol
    assert not np.array_equal(a, b)     # the FP32 path ran
    return e, ep


# the published update at PRELAX 1 with the solid reaches NaN within 3 steps (oracle alike)
@pytest.mark.parametrize("table,prelax,fix", [("solid", 3, 0), ("fluids", 0, 0), ("solid", 3, 1), ("solid", 1, 1),
                                              ("fluids", 0, 1)])
def test_m5_mixed_steps_track_fp64(table, prelax, fix):
    from test_m5_consistent import fixed_case
    p, m, om, S = fixed_case(160 + prelax, True, prelax, table)
    p.fix = fix
    p.dt = 2e-9
    check(p, m, S, 3 if fix == 0 else 10, what=f"{table} prelax {prelax} fix {fix}")


def test_m5_mixed_solid_pulse_tracks_fp64():
    """Thirty steps of the consistent update's solid pulse at CFL 0.3 (test_m5_consistent)."""
    import test_m5_consistent as T
    p = T.params(prelax=0)
    m = T.line()
    cL = np.sqrt((3.0 * (1e7 + 2e10) + 4.0 / 3.0 * 26e9) / 2700.0)
    p.dt = 0.3 * (1e-2 / m.n) / cL
    check(p, m, T.pulse(p, m), 30, what="solid pulse")


def test_m5_mixed_face_flags_equal_fp64():
    from test_m5_consistent import fixed_case
    p, m, om, S = fixed_case(170, False, 3, "solid")
    out = []
    for prec in (F.PREC_MIXED, F.PREC_FP64):
        dev = F.M5Device(p, m)
        dev.set_precision(prec)
        dev.set_state(S)
        dev.faceinterp()
        dev.residual()
        out.append(dev.stage_arrays())
        dev.close()
    (SFa, Ra, ga), (SFb, Rb, gb) = out
    assert np.array_equal(ga, gb) and gb.any()
    scale = np.abs(Rb).max(0) + 1e-300
    say(f"[m5 mixed faces] SF max rel {(np.abs(SFa - SFb).max((0, 1)) / (np.abs(SFb).max((0, 1)) + 1e-300)).max():.2e}"
        f" RHS max rel {(np.abs(Ra - Rb).max(0) / scale).max():.2e}")
    assert (np.abs(SFa - SFb).max((0, 1)) / (np.abs(SFb).max((0, 1)) + 1e-300)).max() <= 1e-5
    assert (np.abs(Ra - Rb).max(0) / scale).max() <= 1e-3


def test_m5_set_precision_rejects_unknown_values():
    from test_m5_consistent import fixed_case
    p, m, om, S = fixed_case(171, False, 0, "fluids")
    dev = F.M5Device(p, m)
    with pytest.raises(F.FVError):
        dev.set_precision(7)
    dev.set_precision(F.PREC_EXACT)
    dev.close()
